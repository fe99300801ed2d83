"""FusedWall.cone_area / dir_extreme on the device (wsb_cone_area2d, wsb_dir_extreme2d; reference scene/fused.py:157-309,
741-835 and its tests/scene/test_cone_area_scene.py): against the oracle restatement (oracle/cone2d.py, the same
closed form: 1e-11 H^2 and 1e-13), and against brute force that uses no edge formula -- a polar midpoint grid about
each query with an even-odd point-in-polygon test and the wedge's fraction of each angular cell (areas, 1e-3 H^2 as
the reference's gate; the grid's own error is ~3e-4 H^2) and the largest |cos| over the same grid's solid points
(the extreme, which the grid approaches from below: within 2e-3).  A rotated, translated L-shape (solid inside), a
cavity (solid outside, background 1) and two bodies (the sum of the single-body results)."""
import math

import numpy as np
import pytest
import torch

from oracle import cone2d as OC
from warpsphboundaries_amd.fused import FusedGroup, FusedWall
from warpsphboundaries_amd.scene import Body, ParticleState, Scene, SurfaceRep, fixed_adjacency

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
LS = np.array([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], dtype=float)
CAV = np.array([(0, 0), (2, 0), (2, 1), (0, 1)], dtype=float)
CENTER, ANGLE, H = (0.3, -0.2), 0.7, 0.6


def _rot(a):
    return np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])


def _wall(bodies, pos, h):
    sc = Scene(bodies, DEV)
    ps = ParticleState(torch.tensor(pos, device=DEV), torch.full((len(pos),), h, dtype=torch.float64, device=DEV))
    return FusedWall(sc, fixed_adjacency(sc, ps, None, h), (FusedGroup("w2"),)), sc


def _inside(poly, x, y):
    """even-odd point in polygon, vectorised."""
    ins = np.zeros(x.shape, dtype=bool)
    n = len(poly)
    for i in range(n):
        (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % n]
        cond = (y0 > y) != (y1 > y)
        xi = x0 + (y - y0) * (x1 - x0) / np.where(y1 != y0, y1 - y0, 1.0)
        ins ^= cond & (x < xi)
    return ins


def _brute(poly_world, solid_inside, p, axis, al, h, nr=600, nt=1200):
    r = (np.arange(nr) + 0.5) / nr * h
    t = (np.arange(nt) + 0.5) / nt * 2 * math.pi
    R, T = np.meshgrid(r, t, indexing="ij")
    X, Y = p[0] + R * np.cos(T), p[1] + R * np.sin(T)
    solid = _inside(poly_world, X, Y)
    if not solid_inside:
        solid = ~solid
    dA = R * (h / nr) * (2 * math.pi / nt)
    th = math.atan2(axis[1], axis[0])
    dt = 2 * math.pi / nt                    # the fraction of each angular cell inside the wedge (cut cells partly)
    inwedge = np.clip((al - (np.abs((T - th + math.pi) % (2 * math.pi) - math.pi) - 0.5 * dt)) / dt, 0.0, 1.0)
    full = float(np.sum(dA * solid))
    wedge = float(np.sum(dA * solid * inwedge)) if al < math.pi else full
    cosmax = float(np.max(np.abs(np.cos(T - th)) * solid, initial=0.0))
    return wedge, full, cosmax


@pytest.mark.parametrize("al", [0.6, math.pi])
def test_rotated_l_shape_matches_oracle_and_brute(al):
    rng = np.random.default_rng(3)
    pos = rng.uniform(-2, 3, (200, 2))
    ax = rng.normal(size=(200, 2))
    wall, sc = _wall([Body(bodyId=0, reps=[SurfaceRep.polygon(LS)], center=CENTER, angle=ANGLE)], pos, H)
    got = wall.cone_area(torch.tensor(ax, device=DEV), al).cpu().numpy()
    ext = wall.dir_extreme(torch.tensor(ax[None], device=DEV)).cpu().numpy()[0]
    R = _rot(ANGLE)
    lp = sc.bodies[0].toLocal(torch.tensor(pos, device=DEV)).cpu().numpy()
    loop = sc.bodies[0].reps[0].vertices.cpu().numpy()
    w, f = OC.cone_area(loop, 0, lp, np.full(len(pos), H), ax @ R, al)
    assert np.abs(got[0] - w).max() <= 1e-11 * H * H and np.abs(got[1] - f).max() <= 1e-11 * H * H
    assert np.abs(ext - OC.dir_extreme(loop, 0, lp, np.full(len(pos), H), ax @ R)).max() <= 1e-13
    world = LS @ R.T + np.asarray(CENTER)
    near = np.nonzero(got[1] > 0.05 * H * H)[0][:12]
    assert len(near) >= 8
    for i in near:
        bw, bf, bc = _brute(world, True, pos[i], ax[i], al, H)
        assert abs(got[0, i] - bw) <= 1e-3 * H * H and abs(got[1, i] - bf) <= 1e-3 * H * H, (i, got[:, i], bw, bf)
        assert bc - 1e-9 <= ext[i] <= bc + 2e-3, (i, ext[i], bc)


def test_cavity_matches_brute():
    rng = np.random.default_rng(4)
    pos = np.concatenate([rng.uniform(0.05, 1.95, (30, 2)) * [1, 0.5], rng.uniform(-0.5, 2.5, (10, 2))])
    ax = rng.normal(size=(len(pos), 2))
    wall, sc = _wall([Body(bodyId=0, reps=[SurfaceRep.polygon(CAV, "outside")])], pos, H)
    for al in (0.4, math.pi):
        got = wall.cone_area(torch.tensor(ax, device=DEV), al).cpu().numpy()
        for i in range(0, len(pos), 3):
            bw, bf, _ = _brute(CAV, False, pos[i], ax[i], al, H)
            assert abs(got[0, i] - bw) <= 1e-3 * H * H and abs(got[1, i] - bf) <= 1e-3 * H * H, (al, i, got[:, i], bw, bf)


def test_two_bodies_is_the_sum_of_the_single_body_results():
    tri = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5]])
    b1 = dict(bodyId=0, reps=[SurfaceRep.polygon([(0, 0), (1, 0), (1, 1), (0, 1)])])
    b2 = dict(bodyId=1, reps=[SurfaceRep.polygon(tri)], center=(1.6, 0.2), angle=math.pi / 6)
    rng = np.random.default_rng(7)
    pos = rng.uniform(-0.5, 2.5, (200, 2))
    ax = torch.tensor(rng.normal(size=(200, 2)), device=DEV)
    both = _wall([Body(**b1), Body(**b2)], pos, 1.0)[0].cone_area(ax, 0.5).cpu().numpy()
    one = _wall([Body(**b1)], pos, 1.0)[0].cone_area(ax, 0.5).cpu().numpy()
    two = _wall([Body(**b2)], pos, 1.0)[0].cone_area(ax, 0.5).cpu().numpy()
    assert np.abs(both - (one + two)).max() <= 1e-11
    vec = torch.tensor(rng.normal(size=(2, 200, 2)), device=DEV)
    e2 = _wall([Body(**b1), Body(**b2)], pos, 1.0)[0].dir_extreme(vec).cpu().numpy()
    assert e2.shape == (2, 200) and ((e2 >= 0) & (e2 <= 1)).all() and (e2 == 1.0).any() and (e2 == 0.0).any()
