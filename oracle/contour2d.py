"""Zero level set of a sampled signed distance as closed polylines: restatement of the reference's marching squares
(``scene/scene.py:762-806``, used by ``SdfRep.contour`` / ``lowered``, ``scene/scene.py:655-666``).  TEST
INFRASTRUCTURE ONLY.

Every cell (i outer, j inner) whose four nodes are not all of one sign gets, on each of its edges whose end nodes
differ (solid = d < 0), the vertex at the linear zero d[a] / (d[a] - d[b]); a vertex is keyed by the directed cell
edge, so the two cells that share a grid edge each number their own copy of its crossing (the loops close by
position, not by index); the
crossings are taken counter-clockwise around the cell and each solid -> fluid crossing is joined to the next
fluid -> solid crossing (the previous one, for a four-crossing cell whose node mean is solid), so the solid lies on
the left of every edge.  Border nodes all fluid: background 0; all solid: background 1; otherwise an error.
"""
import numpy as np


def contour(d, origin, spacing):
    """(vertices [V, 2], edges [E, 2] int32, background) in the reference's numbering."""
    d = np.asarray(d, dtype=float)
    nx, ny = d.shape
    neg = d < 0
    ring_border = [neg[0, j] for j in range(ny)] + [neg[nx - 1, j] for j in range(ny)] + \
        [neg[i, 0] for i in range(1, nx - 1)] + [neg[i, ny - 1] for i in range(1, nx - 1)]
    if any(ring_border) and not all(ring_border):
        raise ValueError("mixed border")
    background = 1 if all(ring_border) else 0
    ids, V, E = {}, [], []
    for i in range(nx - 1):
        for j in range(ny - 1):
            corners = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            f = [bool(neg[c]) for c in corners]
            if all(f) or not any(f):
                continue
            seq = []
            for k in range(4):
                a, b = corners[k], corners[(k + 1) % 4]
                if f[k] != f[(k + 1) % 4]:
                    key = (a, b)
                    if key not in ids:
                        t = d[a] / (d[a] - d[b])
                        ids[key] = len(V)
                        V.append([origin[0] + spacing * (a[0] + t * (b[0] - a[0])),
                                  origin[1] + spacing * (a[1] + t * (b[1] - a[1]))])
                    seq.append(("exit" if f[k] else "entry", ids[key]))
            m = len(seq)
            centre_solid = (d[i, j] + d[i + 1, j] + d[i + 1, j + 1] + d[i, j + 1]) / 4.0 < 0
            step = -1 if (m == 4 and centre_solid) else 1
            for k in range(m):
                if seq[k][0] == "exit":
                    q = (k + step) % m
                    while seq[q][0] != "entry":
                        q = (q + step) % m
                    E.append([seq[k][1], seq[q][1]])
    if not E:
        raise ValueError("no zero level set")
    return np.asarray(V, dtype=float), np.asarray(E, dtype=np.int32), background
