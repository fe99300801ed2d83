"""Restatement of the wedge-area and directional-extreme queries in scalar FP64.  TEST INFRASTRUCTURE ONLY.

Follows ``scene/fused.py`` of the reference: the angle wrap and the chord deficit (:157-176), the per-row cone area
(:179-240: (1/2) H^2 W indicator minus the signed chord deficits of the edges within H) and the directional extreme
(:243-309), over the rows' edges of this oracle's own fixed-capacity adjacency (``oracle/adjacency.py``), in the
body frame.  Pinned (tests/test_gpu_cone2d.py): polar midpoint grids with an even-odd point-in-polygon test (no edge
formulas) for the areas and a fine angular scan of the solid points within H for the extreme.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import adjacency as oadj

TWO_PI = 2.0 * math.pi


def wrap_pi(a):
    r = a + math.pi
    r = r - TWO_PI * math.floor(r / TWO_PI)
    return r - math.pi


def chord_deficit(l, h, z, H, be, pn):
    m = 0.5 * (l + h)
    k = math.floor((m - pn) / TWO_PI + 0.5)
    c_lo = max(l, pn + TWO_PI * k - be)
    c_hi = min(h, pn + TWO_PI * k + be)
    if not c_lo < c_hi:
        return 0.0
    return 0.5 * H * H * (c_hi - c_lo) - 0.5 * z * z * (math.tan(c_hi - pn) - math.tan(c_lo - pn))


def _rows(loop_ccw, background, pts, sup):
    n = len(loop_ccw)
    st = oadj.Static(loop_ccw, np.stack([np.arange(n), (np.arange(n) + 1) % n], 1), background, float(np.max(sup)))
    e, c, ind = oadj.fixed_topology(st, pts, sup)
    return st, [e[r * st.K:r * st.K + c[r]] for r in range(len(pts))], ind


def cone_area(loop_ccw, background, pts, sup, axes_body, half_angle):
    """(wedge area [N], full-disk area [N]) of the solid loop (counter-clockwise around the solid) in its frame."""
    st, rows, ind = _rows(loop_ccw, background, pts, sup)
    wedge, full = np.zeros(len(pts)), np.zeros(len(pts))
    al = float(half_angle)
    wedge_on = al < math.pi
    for q, (p, H) in enumerate(zip(pts, sup)):
        th = math.atan2(axes_body[q][1], axes_body[q][0])
        aw = af = 0.0
        for e in rows[q]:
            A = st.verts[st.edges[e, 0]] - p
            B = st.verts[st.edges[e, 1]] - p
            cr = A[0] * B[1] - A[1] * B[0]
            L = math.hypot(B[0] - A[0], B[1] - A[1])
            if not (L > 0 and abs(cr) >= 1e-300):
                continue
            sg = -1.0 if cr < 0 else 1.0
            z = abs(cr) / L
            if not z < H:
                continue
            pa = math.atan2(A[1], A[0])
            dphi = wrap_pi(math.atan2(B[1], B[0]) - pa)
            lo, hi = (pa + dphi, pa) if dphi < 0 else (pa, pa + dphi)
            dx, dy = (B[0] - A[0]) / L, (B[1] - A[1]) / L
            t = A[0] * dx + A[1] * dy
            pn = lo + wrap_pi(math.atan2(A[1] - t * dy, A[0] - t * dx) - lo)
            be = math.acos(z / H)
            af += sg * chord_deficit(lo, hi, z, H, be, pn)
            if wedge_on:
                c = lo + wrap_pi(th - lo)
                for kk in (-1, 0, 1):
                    l, h = max(lo, c + TWO_PI * kk - al), min(hi, c + TWO_PI * kk + al)
                    if l < h:
                        aw += sg * chord_deficit(l, h, z, H, be, pn)
            else:
                aw += sg * chord_deficit(lo, hi, z, H, be, pn)
        ww = 2.0 * al if wedge_on else TWO_PI
        wedge[q] = 0.5 * H * H * ww * ind[q] - aw
        full[q] = 0.5 * H * H * TWO_PI * ind[q] - af
    return wedge, full


def dir_extreme(loop_ccw, background, pts, sup, vecs_body):
    """[N]: max over the solid points within H of |w_hat . x_hat| (body frame)."""
    st, rows, ind = _rows(loop_ccw, background, pts, sup)
    out = np.zeros(len(pts))
    for q, (p, H) in enumerate(zip(pts, sup)):
        b = np.asarray(vecs_body[q], dtype=np.float64)
        nb = math.hypot(b[0], b[1])
        if not nb > 0:
            continue
        b = b / nb
        best = 1.0 if ind[q] > 0.5 else 0.0
        for e in rows[q]:
            A = st.verts[st.edges[e, 0]] - p
            D = st.verts[st.edges[e, 1]] - p - A
            a = D @ D
            if not a > 0:
                continue
            bq = A @ D
            disc = bq * bq - a * (A @ A - H * H)
            if not disc > 0:
                continue
            sq = math.sqrt(disc)
            t0, t1 = max(0.0, (-bq - sq) / a), min(1.0, (-bq + sq) / a)
            if not t0 < t1:
                continue
            f = 0.0
            crd = D[0] * b[1] - D[1] * b[0]
            if abs(crd) > 0:
                ts = -(A[0] * b[1] - A[1] * b[0]) / crd
                if t0 <= ts <= t1:
                    f = 1.0
            if f < 1.0:
                for t in (t0, t1):
                    r = A + t * D
                    rr = math.hypot(r[0], r[1])
                    if rr > 0:
                        f = max(f, abs(r @ b) / rr)
            best = max(best, f)
        out[q] = best
    return out
