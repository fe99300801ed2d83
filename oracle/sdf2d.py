"""Bilinear signed distance of a sampled field with node-blended jets: restatement of the reference's
``SdfRep._node_fields`` and ``SdfRep._signed`` (``scene/scene.py:665-706``) in numpy FP64.  TEST INFRASTRUCTURE ONLY
(see oracle/__init__).

Node jets: the value d, the central differences (d[i+1] - d[i-1]) / 2s in each axis and the five-point Laplacian, with
the grid padded by repeating its border nodes (``scene/scene.py:665-673``).  At a point: u = (x - origin) / s clamped to
the grid, the cell of the clamped point (the last cell on the far border), the four corner jets blended with the
bilinear weights in the order (0,0), (0,1), (1,0), (1,1); d grows by the distance from the grid, |(u - u_clamped) s|;
the normal is the blended gradient over its length (at least 1e-300); kappa = lap / (1 - d lap)
(``scene/scene.py:686-706``).
"""
import numpy as np


def node_fields(d, spacing):
    """[4, nx, ny]: value, d/dx, d/dy, Laplacian (border nodes repeated)."""
    d = np.asarray(d, dtype=np.float64)
    s = float(spacing)
    p = np.pad(d, 1, mode="edge")
    dx = (p[2:, 1:-1] - p[:-2, 1:-1]) / (2 * s)
    dy = (p[1:-1, 2:] - p[1:-1, :-2]) / (2 * s)
    lap = (p[2:, 1:-1] + p[:-2, 1:-1] + p[1:-1, 2:] + p[1:-1, :-2] - 4 * d) / (s * s)
    return np.stack([d, dx, dy, lap])


def signed(d, origin, spacing, pos):
    """(d, n [M, 2], kappa, gradNorm) at points pos [M, 2]."""
    nodes = node_fields(d, spacing)
    nx, ny = np.shape(d)
    pos = np.asarray(pos, dtype=np.float64)
    u = (pos - np.asarray(origin, dtype=np.float64)) / float(spacing)
    uc = np.stack([np.clip(u[:, 0], 0, nx - 1), np.clip(u[:, 1], 0, ny - 1)], 1)
    i0 = np.stack([np.clip(np.floor(uc[:, 0]), 0, nx - 2), np.clip(np.floor(uc[:, 1]), 0, ny - 2)], 1).astype(np.int64)
    t = uc - i0
    out = 0
    for dx in (0, 1):
        for dy in (0, 1):
            wgt = (t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1])
            out = out + wgt[None] * nodes[:, i0[:, 0] + dx, i0[:, 1] + dy]
    val, gx, gy, lap = out
    val = val + np.sqrt(np.sum(((u - uc) * float(spacing)) ** 2, axis=1))
    gn = np.maximum(np.sqrt(gx * gx + gy * gy), 1e-300)
    return val, np.stack([gx, gy], 1) / gn[:, None], lap / (1.0 - val * lap), gn
