"""Pins of the 2-D adjacency oracle (``oracle/adjacency.py``) by checks that do not share its code path.

* slot sets: a brute-force scan over every edge (no cell list) with the reference's FP64 point-segment predicate
  (``scene/scene.py:121-138``, ``scene/fixedadj.py:75-96``) gives the same edges per row, in ascending order, on
  uniform points, points on edges, at and next to vertices, at edge midpoints and on cell corners;
* cell list: every listed item's inflated box meets its cell, every edge whose inflated box meets a cell's interior
  is listed there (a geometric test over all (cell, edge) pairs), items ascending; K is the largest population;
* indicator: on dyadic polygons with a power-of-two support (every coordinate difference, normalisation and cross
  product is exact in FP64) the oracle's indicator equals an exact ``Fraction`` winding count off the boundary, 1/2
  on edge interiors and the exact interior-angle fraction at vertices; on the star workloads it equals an independent
  crossing-number test away from the wall.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import adjacency as oadj
from warpsphboundaries_amd import workloads as WL


def _loop_edges(n):
    return np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)


def _ccw(P, solid):
    """counter-clockwise around the solid (solid 'outside': clockwise around the fluid), background."""
    a = 0.5 * np.sum(P[:, 0] * np.roll(P[:, 1], -1) - np.roll(P[:, 0], -1) * P[:, 1])
    if (a > 0) != (solid == "inside"):
        P = P[::-1].copy()
    return np.ascontiguousarray(P), (0 if solid == "inside" else 1)


def _geometries():
    g = WL.rng(11)
    notch = np.array([[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.52, 0.9], [0.5, 0.3], [0.48, 0.9], [0.1, 0.9]])
    return {
        "box-outside": _ccw(np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]), "outside"),
        "64gon": _ccw(WL.regular_polygon((0.5, 0.5), 0.15, 64), "inside"),
        "star256": _ccw(WL.star_polygon(256, g=g), "inside"),
        "notch": _ccw(notch, "inside"),
    }


def _brute_slots(V, E, pts, sup, chunk=512):
    """every edge with point-segment distance < support, ascending id; the reference's operation order, no cells."""
    ax, ay = V[E[:, 0], 0][None], V[E[:, 0], 1][None]
    dx, dy = V[E[:, 1], 0][None] - ax, V[E[:, 1], 1][None] - ay
    dd = np.maximum(dx * dx + dy * dy, 1.0e-300)
    out = []
    for s in range(0, len(pts), chunk):
        px, py = pts[s:s + chunk, 0:1], pts[s:s + chunk, 1:2]
        t = np.clip(((px - ax) * dx + (py - ay) * dy) / dd, 0.0, 1.0)
        ex = px - (ax + t * dx)
        ey = py - (ay + t * dy)
        keep = np.sqrt(ex * ex + ey * ey) < sup[s:s + chunk, None]
        out.extend(np.nonzero(k)[0].tolist() for k in keep)
    return out


def _adversarial_points(V, h, n, seed, st):
    g = WL.rng(seed)
    lo, hi = V.min(0) - 2 * h, V.max(0) + 2 * h
    p, q = V, np.roll(V, -1, 0)
    k = g.integers(0, len(V), n // 4)
    t = g.uniform(0, 1, (n // 4, 1))
    corners = st.cl.lo + st.cl.cell * np.stack([g.integers(0, st.cl.dims[0] + 1, 200),
                                                 g.integers(0, st.cl.dims[1] + 1, 200)], 1).astype(np.float64)
    return np.concatenate([
        g.uniform(lo, hi, (n, 2)),
        p[k] + t * (q[k] - p[k]),                     # on the edges (to rounding)
        V[k[:256]],                                   # at the vertices
        p[k] + g.normal(0, 1e-9, (n // 4, 2)),        # next to the vertices
        0.5 * (p + q)[k[:256]],                       # edge midpoints
        corners,                                      # cell corners of the static grid
        p[k] + (h * (1 - 1e-12)) * np.stack([np.cos(t[:, 0] * 6.3), np.sin(t[:, 0] * 6.3)], 1),  # at the support
    ])


def _seg_dist(V, E, p, e):
    ax, ay = V[E[e, 0]]
    dx, dy = V[E[e, 1], 0] - ax, V[E[e, 1], 1] - ay
    t = min(max(((p[0] - ax) * dx + (p[1] - ay) * dy) / max(dx * dx + dy * dy, 1.0e-300), 0.0), 1.0)
    return math.hypot(p[0] - (ax + t * dx), p[1] - (ay + t * dy))


def _check_slots(V, E, bg, h, pts, sup):
    """slot rows equal the brute-force rows.  The one allowed difference is the reference's own: an edge whose
    rounded distance is below the support by a few ulps while the point sits on a cell boundary that the floor of the
    inflated box rounds to the other side (the cell list is complete up to the rounding of the cell indices)."""
    st = oadj.Static(V, E, bg, h)
    e, c, _ = oadj.fixed_topology(st, pts, sup)
    ref = _brute_slots(st.verts, st.edges, pts, sup)
    K = st.K
    ties = 0
    for r in range(len(pts)):
        row = e[r * K:r * K + c[r]].tolist()
        assert (e[r * K + c[r]:(r + 1) * K] == -1).all()
        assert row == sorted(set(row))
        if row != ref[r]:
            assert set(row) <= set(ref[r]), (r, pts[r], row, ref[r])
            for x in set(ref[r]) - set(row):
                d = _seg_dist(st.verts, st.edges, pts[r], x)
                assert sup[r] * (1 - 8e-16 * 8) <= d < sup[r], (r, pts[r], x, d, sup[r])
                f = (pts[r] - st.cl.lo) / st.cl.cell
                assert np.min(np.abs(f - np.round(f))) < 1e-9, (r, pts[r], f)
                ties += 1
    assert int(c.max(initial=0)) <= K
    assert ties <= max(2, len(pts) // 1000)
    return st, c


@pytest.mark.parametrize("name", ["box-outside", "64gon", "star256", "notch"])
def test_slots_equal_brute_force(name):
    V, bg = _geometries()[name]
    E = _loop_edges(len(V))
    h = 0.02 if name != "star256" else 0.01
    st = oadj.Static(V, E, bg, h)
    pts = _adversarial_points(V, h, 3000, 5, st)
    sup = np.where(WL.rng(6).uniform(0, 1, len(pts)) < 0.3, 0.6 * h, h)
    _, c = _check_slots(V, E, bg, h, pts, sup)
    assert c.sum() > len(pts) // 4


def test_slots_equal_brute_force_box_and_64gon_scene():
    """the two representations of the box + 64-gon scene (16,384 particles, w4 support 0.02), each its own list."""
    w = WL.config1()
    for loop, solid in w.loops:
        V, bg = _ccw(loop, solid)
        _check_slots(V, _loop_edges(len(V)), bg, w.h, w.positions, w.supports)


def test_slots_equal_brute_force_star4096_subsample():
    """the 4,096-edge star at h = 0.004 on a stratified 8,192-particle subsample plus adversarial points."""
    w = WL.subsample(WL.config2(n=65536), 8192, seed=8)
    V, bg = _ccw(w.loops[0][0], w.loops[0][1])
    E = _loop_edges(len(V))
    st = oadj.Static(V, E, bg, w.h)
    extra = _adversarial_points(V, w.h, 800, 9, st)
    pts = np.concatenate([w.positions, extra])
    sup = np.full(len(pts), w.h)
    _, c = _check_slots(V, E, bg, w.h, pts, sup)
    assert c[:len(w.positions)][w.near].min() >= 1      # near-wall rows all see the wall


@pytest.mark.parametrize("name", ["box-outside", "64gon", "star256", "notch"])
def test_cell_list_geometry(name):
    V, bg = _geometries()[name]
    E = _loop_edges(len(V))
    h = 0.02 if name != "star256" else 0.01
    st = oadj.Static(V, E, bg, h)
    cl = st.cl
    nx, ny = cl.dims
    a, b = V[E[:, 0]], V[E[:, 1]]
    blo, bhi = np.minimum(a, b) - h, np.maximum(a, b) + h
    populations = []
    for cx in range(nx):
        for cy in range(ny):
            c0 = cl.lo + cl.cell * np.array([cx, cy], dtype=np.float64)
            c1 = c0 + cl.cell
            items = cl.items[cl.start[cx * ny + cy]:cl.start[cx * ny + cy + 1]]
            populations.append(len(items))
            assert (np.diff(items) > 0).all()
            # listed => the inflated box meets the closed cell (with rounding slack of the floor divisions)
            tol = 1e-9 * cl.cell
            for e_ in items:
                assert (blo[e_] <= c1 + tol).all() and (bhi[e_] >= c0 - tol).all()
            # the inflated box meets the open cell => listed
            meets = ((blo < c1 - tol) & (bhi > c0 + tol)).all(1)
            assert set(np.nonzero(meets)[0]) <= set(items.tolist())
    assert st.K == max(max(populations), 1)


# ----------------------------------------------------------------------------------------------- exact winding
def _frac_winding(px, py, V):
    """exact winding number of the closed loop V around (px, py) (Fractions), or None on the loop."""
    w = 0
    n = len(V)
    for i in range(n):
        ax, ay = V[i]
        bx, by = V[(i + 1) % n]
        cr = (ax - px) * (by - py) - (ay - py) * (bx - px)
        if cr == 0 and (ax - px) * (bx - px) + (ay - py) * (by - py) <= 0:
            return None
        if ay <= py < by and cr > 0:
            w += 1
        elif by <= py < ay and cr < 0:
            w -= 1
    return w


def _boundary_fraction(px, py, V):
    """1/2 on an edge interior; the interior angle / 2 pi at a vertex (solid on the left of every edge)."""
    n = len(V)
    for i in range(n):
        if V[i] == (px, py):
            pv, nv = V[i - 1], V[(i + 1) % n]
            ux, uy = float(nv[0] - px), float(nv[1] - py)
            vx, vy = float(pv[0] - px), float(pv[1] - py)
            ang = math.atan2(ux * vy - uy * vx, ux * vx + uy * vy) % (2 * math.pi)
            return ang / (2 * math.pi)
    return 0.5


DYADIC = {
    "square": [(1, 1), (7, 1), (7, 7), (1, 7)],
    "L": [(1, 1), (9, 1), (9, 3), (3, 3), (3, 9), (1, 9)],
    "notch": [(1, 1), (9, 1), (9, 9), (6, 9), (5, 3), (4, 9), (1, 9)],
    "comb": [(0, 0), (12, 0), (12, 2), (10, 2), (10, 8), (8, 8), (8, 2), (6, 2), (6, 8), (4, 8), (4, 2), (2, 2),
             (2, 8), (0, 8)],
    "slanted": [(0, 0), (8, 1), (10, 6), (5, 4), (2, 9), (-1, 5)],
}


@pytest.mark.parametrize("shape", sorted(DYADIC))
@pytest.mark.parametrize("solid", ["inside", "outside"])
def test_indicator_exact_winding(shape, solid):
    scale = 0.0625                                              # coordinates k / 16: exact products
    P = np.array(DYADIC[shape], dtype=np.float64) * scale
    V, bg = _ccw(P, solid)
    E = _loop_edges(len(V))
    h = 0.125                                                    # power of two: the normalisation is exact
    st = oadj.Static(V, E, bg, h)
    g = WL.rng(17)
    lo, hi = V.min(0) - 3 * h, V.max(0) + 3 * h
    grid = np.stack(np.meshgrid(np.arange(lo[0], hi[0], scale / 4), np.arange(lo[1], hi[1], scale / 4)), -1).reshape(-1, 2)
    k = g.integers(0, len(V), 200)
    t = g.integers(1, 8, (200, 1)) / 8.0
    on_edges = V[k] + t * (np.roll(V, -1, 0)[k] - V[k])          # exact: dyadic t on dyadic edges
    pts = np.concatenate([grid, V, on_edges, g.uniform(lo, hi, (500, 2))])
    sup = np.full(len(pts), h)
    _, _, ind = oadj.fixed_topology(st, pts, sup)
    VF = [(Fraction(float(x)), Fraction(float(y))) for x, y in V]
    kinds = {"interior": 0, "boundary": 0}
    for r, (x, y) in enumerate(pts):
        w = _frac_winding(Fraction(float(x)), Fraction(float(y)), VF)
        if w is None:
            # the solid's share of an infinitesimal disk, whichever side is unbounded (winding + background)
            expect = _boundary_fraction(Fraction(float(x)), Fraction(float(y)), VF)
            kinds["boundary"] += 1
            assert abs(ind[r] - expect) <= 1e-14, (r, x, y, ind[r], expect)
        else:
            kinds["interior"] += 1
            assert ind[r] == w + bg, (r, x, y, ind[r], w + bg)
    assert kinds["boundary"] > 200 and kinds["interior"] > 1000


def test_indicator_matches_crossing_number_on_stars():
    """away from the wall the indicator is 0 / 1 and equals an independent crossing-number test."""
    w = WL.subsample(WL.config2(n=65536), 8192, seed=10)
    V, bg = _ccw(w.loops[0][0], w.loops[0][1])
    st = oadj.Static(V, _loop_edges(len(V)), bg, w.h)
    g = WL.rng(12)
    pts = np.concatenate([w.positions, g.uniform(0.15, 0.85, (4000, 2))])
    sup = np.full(len(pts), w.h)
    _, _, ind = oadj.fixed_topology(st, pts, sup)
    ref = WL.inside_polygon(pts, V).astype(np.float64)
    far = np.array([len(r) == 0 for r in _brute_slots(V, _loop_edges(len(V)), pts, np.full(len(pts), 1e-9))])
    assert far.sum() > 0.99 * len(pts)
    assert np.array_equal(ind[far], ref[far])
