"""SdfRep on the host (reference scene/scene.py:599-670, 762-806): the contour of the samples equals the oracle's
restatement of the reference's marching squares (same vertices, edges and numbering), is closed with the solid on the
left (area of an enclosed solid positive, of a tank negative), its vertices are the linear zeros on the grid edges,
it converges to the sampled shape (area within O(spacing^2)); saddle cells; argument errors.  No device calls."""
import math

import numpy as np
import pytest
import torch

from oracle import contour2d as OC
from warpsphboundaries_amd.scene import Body, SdfRep, SurfaceRep


def _area(V, E):
    return 0.5 * float(np.sum(V[E[:, 0], 0] * V[E[:, 1], 1] - V[E[:, 1], 0] * V[E[:, 0], 1]))


def _disk_sdf(c, R, sign=1.0):
    return lambda p: sign * ((p - torch.as_tensor(c, dtype=torch.float64)).norm(dim=1) - R)


@pytest.mark.parametrize("spacing", [0.02, 0.005])
def test_contour_of_an_enclosed_disk(spacing):
    c, R = (0.31, 0.47), 0.2
    rep = SdfRep.fromFunction(_disk_sdf(c, R), (0.0, 0.2), (0.65, 0.75), spacing)
    sr = rep.contour()
    V, E = sr.vertices.numpy(), sr.edges.numpy()
    Vo, Eo, bg = OC.contour(rep.d.numpy(), rep.origin.numpy(), spacing)
    np.testing.assert_array_equal(V, Vo)
    np.testing.assert_array_equal(E, Eo)
    assert sr.background == bg == 0 and rep.solid == "inside" and not Body(reps=[rep]).unbounded()
    # closed loops: every edge ends where another begins (a crossing on a grid edge is a vertex of each of the two
    # cells that share the edge, as in the reference's numbering, so the loops close by position)
    from collections import Counter
    starts = Counter(tuple(np.round(V[a], 12)) for a in E[:, 0])
    ends = Counter(tuple(np.round(V[b], 12)) for b in E[:, 1])
    assert starts == ends
    assert abs(_area(V, E) - math.pi * R * R) < 2.0 * spacing ** 2
    d = rep.signed(torch.as_tensor(V))[0]
    assert float(d.abs().max()) < 1e-14                                      # the linear zeros of the grid edges
    assert rep.lowered(0.01, "cpu") is sr


def test_contour_of_a_tank_and_a_fallback():
    rep = SdfRep.fromFunction(_disk_sdf((0.5, 0.5), 0.3, -1.0), (0.1, 0.1), (0.9, 0.9), 0.01)
    sr = rep.contour()
    Vo, Eo, bg = OC.contour(rep.d.numpy(), rep.origin.numpy(), 0.01)
    np.testing.assert_array_equal(sr.vertices.numpy(), Vo)
    np.testing.assert_array_equal(sr.edges.numpy(), Eo)
    # the reference's convention (scene/scene.py:1039-1051): a sampled field counts as unbounded through its fallback
    # surface only, so the tank contour alone leaves the body bounded
    assert sr.background == bg == 1 and rep.solid == "outside" and not Body(reps=[rep]).unbounded()
    assert _area(Vo, Eo) < 0 and abs(abs(_area(Vo, Eo)) - math.pi * 0.09) < 2e-4
    fb = SurfaceRep.box((0.2, 0.2), (0.8, 0.8), solid="outside")
    tank = SdfRep(rep.d, rep.origin, 0.01, fallback=fb)
    assert tank.lowered(0.01, "cpu") is fb and tank.solid == "outside" and Body(reps=[tank]).unbounded()


def test_fallback_decides_solid_and_unbounded():
    """with a fallback the sampled field's own border is never contoured: a grid whose border is partly solid is
    accepted, and the fallback's background alone decides solid / unbounded (a tank fallback over an all-fluid grid
    border is unbounded, a box obstacle over a mixed border is bounded)."""
    d = np.ones((6, 6))
    d[0, 2] = -1.0                                              # a mixed border: contour() would refuse it
    box = SurfaceRep.box((0.1, 0.1), (0.3, 0.3), solid="inside")
    rep = SdfRep(d, (0.0, 0.0), 0.1, fallback=box)
    assert rep.lowered(0.1, "cpu") is box and rep.solid == "inside" and not Body(reps=[rep]).unbounded()
    with pytest.raises(ValueError):
        rep.contour()
    tank = SdfRep(np.ones((6, 6)), (0.0, 0.0), 0.1, fallback=SurfaceRep.box((0.0, 0.0), (0.5, 0.5), solid="outside"))
    assert tank.solid == "outside" and Body(reps=[tank]).unbounded()


def test_keyword_thresholds_are_validated():
    """the reference's keyword arguments (scene/scene.py:614-630): unknown ones are rejected, bad values raise."""
    g = np.ones((3, 3))
    SdfRep(g, (0, 0), 1.0, maxKappaH=0.3, gradTol=0.2, minCurvatureDenominator=1.0, maxCurvatureVariationH2=2.0,
           backend="torch")
    with pytest.raises(TypeError):
        SdfRep(g, (0, 0), 1.0, fallbak=None)
    with pytest.raises(ValueError):
        SdfRep(g, (0, 0), 1.0, backend="x")
    with pytest.raises(ValueError):
        SdfRep(g, (0, 0), 1.0, minCurvatureDenominator=0.0)
    with pytest.raises(ValueError):
        SdfRep(g, (0, 0), 1.0, maxCurvatureVariationH2=float("inf"))


def test_saddle_cells_follow_the_centre_value():
    # two solid corners on a diagonal: joined across the cell when the node mean is solid, around the corners otherwise
    for centre_solid in (True, False):
        d = np.ones((4, 4))
        d[1, 1] = d[2, 2] = -1.0 if centre_solid else -0.3
        d[2, 1] = d[1, 2] = 0.5 if centre_solid else 1.0
        rep = SdfRep(d, (0.0, 0.0), 1.0)
        sr = rep.contour()
        Vo, Eo, _ = OC.contour(d, (0.0, 0.0), 1.0)
        np.testing.assert_array_equal(sr.vertices.numpy(), Vo)
        np.testing.assert_array_equal(sr.edges.numpy(), Eo)
        assert _area(Vo, Eo) > 0


def test_argument_errors():
    d = np.ones((5, 5))
    d[0, 2] = -1.0                                       # a solid node on the border, the others fluid
    with pytest.raises(ValueError):
        SdfRep(d, (0.0, 0.0), 0.1).contour()
    with pytest.raises(ValueError):
        SdfRep(np.ones((5, 5)), (0.0, 0.0), 0.1).contour()
    with pytest.raises(ValueError):
        SdfRep(np.ones(5), (0.0, 0.0), 0.1)
    with pytest.raises(ValueError):
        SdfRep(np.ones((3, 3)), (0.0, 0.0), 0.0)


# the reference's marching squares worked through by hand (scene/scene.py:762-806): cells i outer, j inner, the
# corners counter-clockwise from (i, j), on each side whose corners differ the vertex at t = d[a] / (d[a] - d[b])
# (origin 0, spacing 1: x = a + t (b - a)), numbered side by side; each solid -> fluid side joined to the next
# fluid -> solid side (the previous one in a saddle cell whose nodes average below zero)
HAND_CASES = {
    # one solid node in the middle of a 3 x 3 grid: four cells, two crossings each
    "single": (
        [[1.0, 1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 1.0, 1.0]],
        [(1.0, 0.5), (0.5, 1.0), (0.5, 1.0), (1.0, 1.5), (1.5, 1.0), (1.0, 0.5), (1.5, 1.0), (1.0, 1.5)],
        [(1, 0), (3, 2), (5, 4), (6, 7)]),
    # two solid nodes on a diagonal of the middle cell of a 4 x 4 grid, node mean -0.25: the saddle joins them
    "saddle": (
        [[1.0, 1.0, 1.0, 1.0], [1.0, -1.0, 0.5, 1.0], [1.0, 0.5, -1.0, 1.0], [1.0, 1.0, 1.0, 1.0]],
        [(1.0, 0.5), (0.5, 1.0), (0.5, 1.0), (1.0, 1.6666666666666665), (1.6666666666666667, 1.0), (1.0, 0.5),
         (1.6666666666666665, 1.0), (2.0, 1.3333333333333333), (1.3333333333333335, 2.0), (1.0, 1.6666666666666667),
         (1.3333333333333333, 2.0), (2.0, 2.5), (2.5, 2.0), (2.0, 1.3333333333333335), (2.5, 2.0), (2.0, 2.5)],
        [(1, 0), (3, 2), (5, 4), (6, 9), (8, 7), (11, 10), (13, 12), (14, 15)]),
}


@pytest.mark.parametrize("case", sorted(HAND_CASES))
def test_contour_numbering_matches_the_hand_worked_cases(case):
    d, V, E = HAND_CASES[case]
    Vo, Eo, bg = OC.contour(np.array(d), (0.0, 0.0), 1.0)
    assert bg == 0
    np.testing.assert_array_equal(Vo, np.array(V))
    np.testing.assert_array_equal(Eo, np.array(E))
    sr = SdfRep(np.array(d), (0.0, 0.0), 1.0).contour()
    np.testing.assert_array_equal(sr.vertices.numpy(), np.array(V))
    np.testing.assert_array_equal(sr.edges.numpy(), np.array(E))


def test_sdf_oracle_closed_forms():
    """the oracle's jets reduce to closed forms: the node values and central differences at the nodes, a linear field
    reproduced exactly inside the grid (gradient, unit normal, zero Laplacian), and the distance from the grid added
    beyond it."""
    from oracle import sdf2d as OS
    g = np.random.default_rng(3)
    d = g.normal(size=(7, 5))
    nodes = np.stack(np.meshgrid(np.arange(7), np.arange(5), indexing="ij"), -1).reshape(-1, 2) * 0.5 + 0.25
    val, n, kappa, gn = OS.signed(d, (0.25, 0.25), 0.5, nodes)
    np.testing.assert_array_equal(val, d.ravel())
    i, j = 3, 2
    gx = (d[i + 1, j] - d[i - 1, j]) / 1.0
    gy = (d[i, j + 1] - d[i, j - 1]) / 1.0
    k = i * 5 + j
    np.testing.assert_allclose(n[k] * gn[k], [gx, gy], rtol=1e-15)
    # a linear field a x + b y + c
    a, b, c = 0.6, -0.8, 0.1
    X, Y = np.meshgrid(0.25 + 0.5 * np.arange(7), 0.25 + 0.5 * np.arange(5), indexing="ij")
    lin = a * X + b * Y + c
    inner = g.uniform([0.75, 0.75], [2.75, 1.75], (200, 2))
    val, n, kappa, gn = OS.signed(lin, (0.25, 0.25), 0.5, inner)
    np.testing.assert_allclose(val, a * inner[:, 0] + b * inner[:, 1] + c, rtol=0, atol=1e-14)
    np.testing.assert_allclose(n, np.broadcast_to([a, b], n.shape), rtol=0, atol=1e-14)
    np.testing.assert_allclose(kappa, 0.0, atol=1e-12)
    # beyond the grid: the border value plus the distance from the grid
    far = np.array([[10.0, 1.0], [0.25, -3.0], [-2.0, -2.0]])
    clamp = np.clip(far, [0.25, 0.25], [3.25, 2.25])
    base = OS.signed(lin, (0.25, 0.25), 0.5, clamp)[0]
    np.testing.assert_allclose(OS.signed(lin, (0.25, 0.25), 0.5, far)[0], base + np.linalg.norm(far - clamp, axis=1),
                               rtol=1e-15)


def test_signed_matches_the_oracle_inside_and_outside_the_grid():
    """SdfRep.signed against the oracle restatement (scene/scene.py:665-706) of a sampled disk, at points inside the
    grid, on its border and far beyond it (where the nearest of two bodies is decided)."""
    from oracle import sdf2d as OS
    rep = SdfRep.fromFunction(_disk_sdf((0.5, 0.5), 0.3), (0.0, 0.0), (1.0, 1.0), 0.02)
    g = np.random.default_rng(4)
    pts = np.concatenate([g.uniform(-0.2, 1.2, (2000, 2)), [[10.0, 10.0], [-5.0, 0.3], [0.5, 7.0], [1.0, 1.0]]])
    got = rep.signed(torch.as_tensor(pts))
    ref = OS.signed(rep.d.numpy(), rep.origin.numpy(), rep.spacing, pts)
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a.numpy(), b, rtol=1e-14, atol=1e-15)
    # (10, 10) from a disk of radius 0.3 at (0.5, 0.5): the border value plus the distance beyond the grid
    assert abs(float(got[0][-4]) - (np.hypot(9.5, 9.5) - 0.3)) < 0.05
