"""Restatement of the conditioned small / sliver triangle route in numpy FP64.  TEST INFRASTRUCTURE ONLY.

Follows ``edge/warptriangle.py`` of the reference: ``prepare`` (:225-240), the cuts of a line at its closest point
and support-circle crossings (``_cuts`` :34-46), the inner segment of one outer node (``_line`` :63-99), one outer
panel of 8 or 16 nodes with 32- or 16-node inner rules (``_panel`` :102-116) and the per-row adaptive bisection with
its status, estimate and final assembly (``_triangle_kernel`` :119-181); the kernel in powers of u = R - r
(``edge/np_fem.py:100-114``) and the regular -W'(r)/r of the first piece near the origin (``_plan`` :184-197).
The kernel blocks come from this oracle's own table (``oracle/kernels.py``).

Pinned (tests/test_triangle2d_cpu.py): the reference's 40-digit golden fixtures of small and sliver elements
(``edge2d_golden.json``: elements 1e-6 h to 1e-2 h, aspect 1e-12 slivers), the edge-local closed form of
``oracle/edge2d.py`` run in 50-digit mpmath (no cancellation at that precision), reversal of the loop, rotation by
90 degrees and the support scaling.
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from math import comb

import numpy as np

from oracle import kernels as OK

_GL = {n: ((x + 1) / 2, w / 2) for n in (8, 16, 32) for x, w in [np.polynomial.legendre.leggauss(n)]}


@lru_cache(maxsize=None)
def _tables(kernel):
    """({R: dense b_k of pi W in powers of (R - r)}, ascending radii, first-piece coefficients of pi W, its upper end)."""
    ub = {}
    for R, coeffs in OK.blocks(kernel):
        d = ub.setdefault(R, {})
        for n, c in enumerate(coeffs):
            for k in range(n + 1):
                d[k] = d.get(k, 0) + c * comb(n, k) * R ** (n - k) * (-1) ** k
    radii = sorted(ub)
    dense = {float(R): np.array([float(ub[R].get(k, 0)) for k in range(max(ub[R]) + 1)]) for R in radii}
    c2pi, pieces = OK._SPEC[kernel]
    (lo, hi), e = pieces[0]
    origin = np.array([float(c) for c in OK._coeffs(c2pi * e)])
    return dense, [float(R) for R in radii], origin, float(min(Fraction(1, 8), Fraction(str(hi))))


def _kernel(kernel, r):
    """pi W(r) and -(pi W)'(r)/r for an array r (edge/warptriangle.py:84-92)."""
    dense, _, origin, origin_hi = _tables(kernel)
    W = np.zeros_like(r)
    dw = np.zeros_like(r)
    for R, b in dense.items():
        m = r < R
        z = np.where(m, R - r, 0.0)
        W = W + np.where(m, sum(b[k] * z ** k for k in range(len(b))), 0.0)
        dw = dw - np.where(m, sum(k * b[k] * z ** (k - 1) for k in range(1, len(b))), 0.0)
    radial = -dw / np.maximum(r, 1e-300)
    if origin[1] == 0.0:
        reg = np.zeros_like(r)
        for k in range(len(origin) - 1, 1, -1):
            reg = reg * r - k * origin[k]
        radial = np.where(r < origin_hi, reg, radial)
    return W, radial


def _cuts(cuts, a, b, radii):
    """closest point of a + t b and its support-circle crossings, kept in (0, 1) (edge/warptriangle.py:34-46)."""
    bb = float(b @ b)
    if bb > 0.0:
        t = -float(a @ b) / bb
        cuts.append(t)
        cl = a + t * b
        rr = float(cl @ cl)
        for R in radii:
            delta = (R * R - rr) / bb
            if delta >= 0.0:
                root = np.sqrt(delta)
                cuts += [t - root, t + root]
    return cuts


def _clip_sort(cuts):
    """[0, the cuts strictly inside (0, 1) in ascending order, 1] (the first two entries are 0 and 1)."""
    return [0.0] + sorted(t for t in cuts[2:] if 0.0 < t < 1.0) + [1.0]


def _line(kernel, u, a, b, c, scales, count):
    """the inner segment a + u b + t (1 - u) c, t in [0, 1], cut and integrated with `count` nodes per piece
    (edge/warptriangle.py:63-99); channels / scales, times (1 - u) / pi."""
    _, radii, _, _ = _tables(kernel)
    start, direction = a + u * b, (1.0 - u) * c
    cuts = _clip_sort(_cuts([0.0, 1.0], start, direction, radii))
    x, w = _GL[count]
    out = np.zeros(12)
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        if hi <= lo:
            continue
        t = lo + (hi - lo) * x
        y = start[None] + t[:, None] * direction[None]
        r = np.sqrt((y * y).sum(1))
        W, radial = _kernel(kernel, r)
        Gx, Gy = radial * y[:, 0], radial * y[:, 1]
        v = np.stack([W, y[:, 0] * W, y[:, 1] * W, Gx, Gy, y[:, 0] * Gx, y[:, 0] * Gy, y[:, 1] * Gx, y[:, 1] * Gy,
                      y[:, 0] ** 2 * W, y[:, 0] * y[:, 1] * W, y[:, 1] ** 2 * W], 1)
        out += ((hi - lo) * w * (1.0 - u) / np.pi) @ v / scales
    return out


def _panel(kernel, lo, hi, a, b, c, scales, kind):
    """outer Gauss panel: kind 1 = 16 x 32 nodes, 0 = 8 x 32, -1 = 16 x 16 (edge/warptriangle.py:102-116)."""
    x, w = _GL[16 if kind != 0 else 8]
    inner = 16 if kind < 0 else 32
    out = np.zeros(12)
    for xj, wj in zip(x, w):
        out += (hi - lo) * wj * _line(kernel, lo + (hi - lo) * xj, a, b, c, scales, inner)
    return out


def prepare(vertices, edges):
    """(order with the shortest edge first, |twice the signed area|, its sign) or None (edge/warptriangle.py:225-240)."""
    v = np.asarray(vertices, dtype=np.float64)
    if len(v) != 3 or len(edges) != 3:
        return None
    es = {tuple(int(x) for x in p) for p in edges}
    if es == {(0, 1), (1, 2), (2, 0)}:
        orient = [0, 1, 2]
    elif es == {(0, 2), (2, 1), (1, 0)}:
        orient = [0, 2, 1]
    else:
        return None
    vf = [[Fraction(float(x)) for x in p] for p in v]
    a, b, c = (vf[i] for i in orient)
    signed = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
    if not signed:
        return None
    short = int(np.argmin([np.dot(v[(i + 1) % 3] - v[i], v[(i + 1) % 3] - v[i]) for i in range(3)]))
    return [short, (short + 2) % 3, (short + 1) % 3], float(abs(signed)), float(1 if signed > 0 else -1)


def row(vertices, meta, x, h, kernel, background=0, tolerance=2e-12, max_panels=4096):
    """(values[12], estimate[12], status) of one particle in units of h (edge/warptriangle.py:119-181)."""
    order, area2, sign = meta
    V = np.asarray(vertices, dtype=np.float64)
    v0, v1, v2 = V[order[0]], V[order[1]], V[order[2]]
    e0, e1 = v1 - v0, v2 - v0
    longest2 = max(e0 @ e0, e1 @ e1, (e1 - e0) @ (e1 - e0))
    if longest2 > 0.02 * 0.02 * h * h and area2 > 0.01 * longest2:
        return np.zeros(12), np.zeros(12), 0
    status = 1
    a, b, c = (v0 - np.asarray(x, dtype=np.float64)) / h, e0 / h, e1 / h
    sx = max(abs(a[0]), abs(a[0] + b[0]), abs(a[0] + c[0]), 1e-150)
    sy = max(abs(a[1]), abs(a[1] + b[1]), abs(a[1] + c[1]), 1e-150)
    scales = np.array([1, sx, sy, sx, sy, sx * sx, sx * sy, sy * sx, sy * sy, sx * sx, sx * sy, sy * sy])
    _, radii, _, _ = _tables(kernel)
    cuts = _cuts([0.0, 1.0], a, b, radii)
    cuts = _cuts(cuts, a + c, b - c, radii)
    bc = b[0] * c[1] - b[1] * c[0]
    ac = a[0] * c[1] - a[1] * c[0]
    if bc != 0.0:
        cuts.append(-ac / bc)
        lc = np.sqrt(c @ c)
        for R in radii:
            cuts += [(-R * lc - ac) / bc, (R * lc - ac) / bc]
    cuts = _clip_sort(cuts)
    total, errors = np.zeros(12), np.zeros(12)
    visited = 0
    for lo0, hi0 in zip(cuts[:-1], cuts[1:]):
        stack = [(lo0, hi0)]
        while stack:
            lo, hi = stack.pop()
            if hi <= lo:
                continue
            fine = _panel(kernel, lo, hi, a, b, c, scales, 1)
            coarse = _panel(kernel, lo, hi, a, b, c, scales, 0)
            err, norm = np.abs(fine - coarse).max(), np.abs(fine).max()
            visited += 1
            accept = err <= tolerance * (hi - lo + norm)
            if not accept and len(stack) < 62 and visited < max_panels and hi - lo > 1e-14:
                mid = 0.5 * (lo + hi)
                stack += [(mid, hi), (lo, mid)]
            else:
                if not accept:
                    status = -1
                inner = _panel(kernel, lo, hi, a, b, c, scales, -1)
                total += fine
                errors += np.abs(fine - coarse) + np.abs(fine - inner)
    jac = area2 / (h * h)
    plane = np.array(OK.plane_vector(kernel))
    values = background * plane + sign * total * scales * jac
    if status < 0:
        values = np.full(12, np.nan)
    return values, errors * scales * jac, status


def channels(vertices, edges, positions, supports, kernel, background=0, tolerance=2e-12, max_panels=4096):
    """(values[N, 12], estimate[N, 12], status[N]) of every particle (status 0 rows zero)."""
    meta = prepare(vertices, edges)
    P = np.asarray(positions, dtype=np.float64).reshape(-1, 2)
    H = np.broadcast_to(np.asarray(supports, dtype=np.float64), (len(P),))
    vals, est, st = np.zeros((len(P), 12)), np.zeros((len(P), 12)), np.zeros(len(P), dtype=np.int32)
    if meta is None:
        return vals, est, st
    for i, (x, h) in enumerate(zip(P, H)):
        vals[i], est[i], st[i] = row(vertices, meta, x, h, kernel, background, tolerance, max_panels)
    return vals, est, st
