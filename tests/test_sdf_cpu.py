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
    d, _ = rep.signed(torch.as_tensor(V))
    assert float(d.abs().max()) < 1e-14                                      # the linear zeros of the grid edges
    assert rep.lowered(0.01, "cpu") is sr


def test_contour_of_a_tank_and_a_fallback():
    rep = SdfRep.fromFunction(_disk_sdf((0.5, 0.5), 0.3, -1.0), (0.1, 0.1), (0.9, 0.9), 0.01)
    sr = rep.contour()
    Vo, Eo, bg = OC.contour(rep.d.numpy(), rep.origin.numpy(), 0.01)
    np.testing.assert_array_equal(sr.vertices.numpy(), Vo)
    np.testing.assert_array_equal(sr.edges.numpy(), Eo)
    assert sr.background == bg == 1 and Body(reps=[rep]).unbounded()
    assert _area(Vo, Eo) < 0 and abs(abs(_area(Vo, Eo)) - math.pi * 0.09) < 2e-4
    fb = SurfaceRep.box((0.2, 0.2), (0.8, 0.8), solid="outside")
    assert SdfRep(rep.d, rep.origin, 0.01, fallback=fb).lowered(0.01, "cpu") is fb


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
