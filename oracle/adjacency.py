"""FP64 numpy restatement of the 2-D static cell list, fixed-capacity slot fill and winding indicator.
TEST INFRASTRUCTURE (see oracle/__init__).

Every floating-point operation is written in the reference's order with each product rounded separately
(numpy never contracts a*b+c), which is what the device adjacency must reproduce bit for bit:

* ``build_cell_list``    : ``scene/scene.py:100-119`` (``buildCellList``) on the support-inflated edge boxes of
                           ``scene/scene.py:273-279`` (``SurfaceRep._celllist``)
* ``build_ybins``        : ``scene/scene.py:314-331`` (``SurfaceRep._ybins``)
* ``indicator_points``   : ``scene/scene.py:333-371`` (``SurfaceRep.indicator``)
* ``cell_indicator``     : ``scene/scene.py:281-294``
* ``fixed_topology``     : ``scene/fixedadj.py:67-147`` (``_fixed_topology_kernel``): slots in ascending edge id of
                           the query's cell, kept iff the point-segment distance is below the support; indicator by
                           the cell cache, the background outside the grid, else y-bin ray casting with the boundary
                           angle sum.
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI_LITERAL = 6.283185307179586


def _cross2(ax, ay, bx, by):
    left = ax * by
    right = ay * bx
    return left - right


class CellList:
    def __init__(self, lo, cell, dims, start, items):
        self.lo, self.cell, self.dims, self.start, self.items = lo, cell, dims, start, items


def build_cell_list(box_lo, box_hi, cell):
    """every item into every cell its box overlaps; items of a cell in ascending item id (stable sort)."""
    gmin = box_lo.min(0)
    gmax = box_hi.max(0)
    dims = np.maximum(np.ceil((gmax - gmin) / cell).astype(np.int64), 1)
    nx, ny = int(dims[0]), int(dims[1])
    i0 = np.maximum(np.floor((box_lo - gmin) / cell).astype(np.int64), 0)
    i1 = np.minimum(np.floor((box_hi - gmin) / cell).astype(np.int64), dims - 1)
    cnt = np.maximum(i1 - i0 + 1, 0)
    per = cnt[:, 0] * cnt[:, 1]
    item = np.repeat(np.arange(len(box_lo)), per)
    start = np.cumsum(per) - per
    local = np.arange(int(per.sum())) - np.repeat(start, per)
    cx = i0[item, 0] + local % cnt[item, 0]
    cy = i0[item, 1] + local // cnt[item, 0]
    cid = cx * ny + cy
    order = np.argsort(cid, kind="stable")
    counts = np.bincount(cid[order], minlength=nx * ny)
    st = np.concatenate([[0], np.cumsum(counts)])
    return CellList(gmin, float(cell), (nx, ny), st, item[order])


def edge_cell_list(verts, edges, support_max):
    a = verts[edges[:, 0]]
    b = verts[edges[:, 1]]
    lo = np.minimum(a, b) - support_max
    hi = np.maximum(a, b) + support_max
    return build_cell_list(lo, hi, support_max)


def build_ybins(verts, edges):
    a = verts[edges[:, 0]]
    b = verts[edges[:, 1]]
    ylo = np.minimum(a[:, 1], b[:, 1])
    yhi = np.maximum(a[:, 1], b[:, 1])
    nb = max(1, int(math.sqrt(len(edges))) * 2)
    y0 = float(ylo.min())
    dy = max(float(yhi.max()) - y0, 1e-300) / nb
    i0 = np.clip(np.floor((ylo - y0) / dy).astype(np.int64), 0, nb - 1)
    i1 = np.clip(np.floor((yhi - y0) / dy).astype(np.int64), 0, nb - 1)
    cnt = i1 - i0 + 1
    item = np.repeat(np.arange(len(a)), cnt)
    st = np.cumsum(cnt) - cnt
    b_ = i0[item] + np.arange(int(cnt.sum())) - np.repeat(st, cnt)
    order = np.argsort(b_, kind="stable")
    counts = np.bincount(b_[order], minlength=nb)
    start = np.concatenate([[0], np.cumsum(counts)])
    return y0, dy, nb, start, item[order]


def _winding_row(px, py, h, verts, edges, bin_items, background, normalise=True):
    """one point: y-bin crossing count, boundary angle sum (scene/fixedadj.py:98-146)."""
    # the crossing count is an integer, so the per-item decisions are evaluated elementwise (numpy rounds every
    # product and difference separately, as the scalar form) and counted in any order
    its = np.asarray(bin_items, dtype=np.int64)
    a = verts[edges[its, 0]]
    c = verts[edges[its, 1]]
    ay, cyy = a[:, 1], c[:, 1]
    if normalise:
        arx, ary, crx, cry = (a[:, 0] - px) / h, (ay - py) / h, (c[:, 0] - px) / h, (cyy - py) / h
    else:
        arx, ary, crx, cry = a[:, 0] - px, ay - py, c[:, 0] - px, cyy - py
    cross = _cross2(arx, ary, crx, cry)
    up = (ay <= py) & (py < cyy) & (cross > 0.0)
    down = (cyy <= py) & (py < ay) & (cross < 0.0)
    w = float(int(up.sum()) - int(down.sum()))
    boundary = bool(((cross == 0.0) & (arx * crx + ary * cry <= 0.0)).any())
    if boundary:
        angle = 0.0
        for e in range(len(edges)):
            v0, v1 = edges[e]
            if normalise:
                arx, ary = (verts[v0, 0] - px) / h, (verts[v0, 1] - py) / h
                crx, cry = (verts[v1, 0] - px) / h, (verts[v1, 1] - py) / h
            else:
                arx, ary = verts[v0, 0] - px, verts[v0, 1] - py
                crx, cry = verts[v1, 0] - px, verts[v1, 1] - py
            cross = _cross2(arx, ary, crx, cry)
            if cross != 0.0:
                angle += math.atan2(cross, arx * crx + ary * cry)
        w = angle / TWO_PI_LITERAL
    return w + background


def indicator_points(verts, edges, background, pts, supports=None, ybins=None):
    """winding number + background (scene/scene.py:333-371); `supports` normalises the relative coordinates."""
    y0, dy, nb, start, items = ybins if ybins is not None else build_ybins(verts, edges)
    out = np.zeros(len(pts))
    for r, (px, py) in enumerate(pts):
        iy = math.floor((py - y0) / dy)
        if 0 <= iy <= nb:
            b = min(iy, nb - 1)
            its = items[start[b]:start[b + 1]]
        else:
            its = []
        h = 1.0 if supports is None else float(supports[r])
        out[r] = _winding_row(px, py, h, verts, edges, its, background, supports is not None)
    return out


def cell_indicator(verts, edges, background, cl: CellList, ybins):
    nx, ny = cl.dims
    counts = cl.start[1:] - cl.start[:-1]
    empty = counts == 0
    ind = np.zeros(nx * ny)
    idx = np.nonzero(empty)[0]
    if len(idx):
        cx, cy = idx // ny, idx % ny
        centres = cl.lo + (np.stack([cx, cy], 1).astype(np.float64) + 0.5) * cl.cell
        ind[idx] = indicator_points(verts, edges, background, centres, None, ybins)
    return ind, empty


class Static:
    """the static structures of one surface representation for one support_max."""

    def __init__(self, verts, edges, background, support_max):
        self.verts = np.asarray(verts, dtype=np.float64)
        self.edges = np.asarray(edges, dtype=np.int64)
        self.background = float(background)
        self.support_max = float(support_max)
        self.cl = edge_cell_list(self.verts, self.edges, self.support_max)
        self.ybins = build_ybins(self.verts, self.edges)
        self.cind, self.cempty = cell_indicator(self.verts, self.edges, self.background, self.cl, self.ybins)
        counts = self.cl.start[1:] - self.cl.start[:-1]
        self.K = max(1, int(counts.max()))


def fixed_topology(st: Static, lpos, lsup, valid=None):
    """(slots e[N*K] int32 with -1 for empty, counts[N], indicator[N]) -- bit-exact restatement."""
    lpos = np.asarray(lpos, dtype=np.float64)
    lsup = np.asarray(lsup, dtype=np.float64)
    N = len(lpos)
    K = st.K
    valid = np.ones(N, dtype=bool) if valid is None else np.asarray(valid, dtype=bool)
    e_out = np.full(N * K, -1, dtype=np.int32)
    cnt = np.zeros(N, dtype=np.int32)
    ind = np.zeros(N)
    cl = st.cl
    nx, ny = cl.dims
    V, E = st.verts, st.edges
    y0, dy, nb, bstart, bitems = st.ybins
    # vectorised slot fill: every (row, cell item) candidate, predicate elementwise in the reference's order
    px, py = lpos[:, 0], lpos[:, 1]
    cx = np.floor((px - cl.lo[0]) / cl.cell).astype(np.int64)
    cy = np.floor((py - cl.lo[1]) / cl.cell).astype(np.int64)
    inside = (cx >= 0) & (cx < nx) & (cy >= 0) & (cy < ny)
    cid = np.where(inside, cx * ny + cy, 0)
    ncand = np.where(inside & valid, cl.start[cid + 1] - cl.start[cid], 0)
    rows = np.repeat(np.arange(N), ncand)
    off = np.arange(int(ncand.sum())) - np.repeat(np.cumsum(ncand) - ncand, ncand)
    e = cl.items[cl.start[cid[rows]] + off]
    ax, ay = V[E[e, 0], 0], V[E[e, 0], 1]
    dx = V[E[e, 1], 0] - ax
    dyy = V[E[e, 1], 1] - ay
    qx, qy = px[rows], py[rows]
    t = np.clip(((qx - ax) * dx + (qy - ay) * dyy) / np.maximum(dx * dx + dyy * dyy, 1.0e-300), 0.0, 1.0)
    ex = qx - (ax + t * dx)
    ey = qy - (ay + t * dyy)
    keep = np.sqrt(ex * ex + ey * ey) < lsup[rows]
    kr = rows[keep]
    ke = e[keep]
    kc = np.bincount(kr, minlength=N)
    slot = np.arange(len(kr)) - np.repeat(np.cumsum(kc) - kc, kc)
    e_out[kr * K + slot] = ke
    cnt[:] = kc
    # indicator
    for r in range(N):
        if not valid[r]:
            continue
        h = lsup[r]
        if not inside[r]:
            ind[r] = st.background
            continue
        c = cid[r]
        if st.cempty[c]:
            ind[r] = st.cind[c]
            continue
        iy = math.floor((py[r] - y0) / dy)
        if 0 <= iy <= nb:
            b = min(iy, nb - 1)
            ind[r] = _winding_row(px[r], py[r], h, V, E, bitems[bstart[b]:bstart[b + 1]], st.background, True)
        else:
            ind[r] = st.background
    return e_out, cnt, ind


def brute_force_slots(verts, edges, lpos, lsup):
    """set-level reference without the cell list: every edge closer than the support, ascending id."""
    V = np.asarray(verts, dtype=np.float64)
    E = np.asarray(edges, dtype=np.int64)
    out = []
    for (px, py), h in zip(np.asarray(lpos), np.asarray(lsup)):
        ax, ay = V[E[:, 0], 0], V[E[:, 0], 1]
        dx, dyy = V[E[:, 1], 0] - ax, V[E[:, 1], 1] - ay
        t = np.clip(((px - ax) * dx + (py - ay) * dyy) / np.maximum(dx * dx + dyy * dyy, 1.0e-300), 0.0, 1.0)
        ex = px - (ax + t * dx)
        ey = py - (ay + t * dyy)
        out.append(np.nonzero(np.sqrt(ex * ex + ey * ey) < h)[0].tolist())
    return out
