"""Periodic boxes on the device: a disk bundle that spans a periodic box, evaluated through its tiled images, against
the oracle's disk element summed over every image of every disk within reach of each particle (enumerated here by
brute force over box shifts); a compact polygon near the box edge equals the same polygon evaluated at the particles'
nearest images; signed distance (within the support's reach, for which the images are tiled) and inside across the box
boundary against brute force over the images."""
import itertools

import numpy as np
import pytest
import torch

from oracle import disk2d as OD
from warpsphboundaries_amd.fused import FusedGroup, FusedWall
from warpsphboundaries_amd.periodic import Periodic
from warpsphboundaries_amd.scene import Body, DiskArrayRep, ParticleState, Scene, SurfaceRep, fixed_adjacency

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
F64 = torch.float64
BOX = Periodic((0.0, 0.0), (1.0, 1.0))
L = np.array([1.0, 1.0])
H = 0.08


def _bundle(seed=4, n=60):
    g = np.random.default_rng(seed)
    c, r = [], []
    while len(c) < n:
        x = g.uniform(-0.5, 0.5, 2)
        rad = g.uniform(0.01, 0.04)
        ok = all(np.linalg.norm(min_img(x - y)) > rad + ry + 1e-3 for y, ry in zip(c, r))
        if ok:
            c.append(x)
            r.append(rad)
    return np.array(c), np.array(r)


def min_img(d):
    return d - L * np.round(d / L)


def _ps(pos, sup):
    n = len(pos)
    return ParticleState(torch.tensor(pos, dtype=F64, device=DEV), torch.tensor(sup, dtype=F64, device=DEV),
                         kinds=torch.zeros(n, dtype=torch.int32, device=DEV))


def test_spanning_bundle_equals_the_oracle_over_all_images():
    c, r = _bundle()
    center = np.array([0.5, 0.5])
    body = Body(reps=[DiskArrayRep(c, r)], center=tuple(center))
    sc = Scene([body], DEV)
    sc.setPeriodic(BOX, H)
    g = np.random.default_rng(7)
    pos = g.uniform(-1.0, 2.0, (3000, 2))             # anywhere: the box shifts bring them in
    sup = H * g.uniform(0.6, 1.0, len(pos))
    ps = _ps(pos, sup)
    adj = fixed_adjacency(sc, ps, None, H)
    got = FusedWall(sc, adj, (FusedGroup("w4", None, "fp64"),)).channels[0, 0].cpu().numpy()
    lpos = min_img(pos - center)
    ref = np.zeros((len(pos), 12))
    crossing = 0
    for i in range(len(pos)):
        for k in itertools.product((-1, 0, 1), repeat=2):
            ci = c + np.array(k) * L
            y = ci - lpos[i]
            D = np.linalg.norm(y, axis=1)
            m = np.nonzero(D < sup[i] + r)[0]
            if len(m) == 0:
                continue
            crossing += int(k != (0, 0)) * len(m)
            yy = y[m] / sup[i]
            ch = OD.circle_channels("w4", np.linalg.norm(yy, axis=1), r[m] / sup[i])
            ref[i] += OD.assemble(ch, yy).sum(0)
    assert crossing > 50
    # physical units of the fused channels
    s = sup[:, None]
    ref[:, [1, 2]] *= s
    ref[:, [3, 4]] /= s
    ref[:, [9, 10, 11]] *= s * s
    fam = ([0], [1, 2], [3, 4], [5, 6, 7, 8], [9, 10, 11])
    for f in fam:
        scale = max(float(np.abs(ref[:, f]).max()), 1e-6)
        assert np.abs(got[:, f] - ref[:, f]).max() <= 1e-12 * scale, (f, np.abs(got[:, f] - ref[:, f]).max() / scale)


def test_compact_polygon_sees_the_nearest_images():
    loop = np.array([[-0.15, -0.1], [0.15, -0.1], [0.05, 0.12]])
    body = Body(reps=[SurfaceRep.polygon(loop)], center=(0.92, 0.1), angle=0.3)
    sc = Scene([body], DEV)
    sc.setPeriodic(BOX, H)
    g = np.random.default_rng(8)
    pos = g.uniform(0.0, 1.0, (4000, 2))
    sup = np.full(len(pos), H)
    got = FusedWall(sc, fixed_adjacency(sc, _ps(pos, sup), None, H), (FusedGroup("w2", None, "fp64"),)).channels[0, 0]
    img = body.image(torch.tensor(pos, dtype=F64, device=DEV)).cpu().numpy()
    moved = Body(reps=[SurfaceRep.polygon(loop)], center=(0.92, 0.1), angle=0.3)
    sc2 = Scene([moved], DEV)
    ref = FusedWall(sc2, fixed_adjacency(sc2, _ps(img, sup), None, H), (FusedGroup("w2", None, "fp64"),)).channels[0, 0]
    assert torch.equal(got, ref)
    far = np.linalg.norm(pos - np.array([0.92, 0.1]), axis=1) > 0.5            # seen only through the box boundary
    assert (got.abs().sum(1).cpu().numpy()[far] > 0).sum() > 20


def test_signed_distance_and_inside_across_the_box():
    c, r = _bundle(seed=5, n=30)
    center = np.array([0.5, 0.5])
    sc = Scene([Body(reps=[DiskArrayRep(c, r)], center=tuple(center))], DEV)
    sc.setPeriodic(BOX, H)
    g = np.random.default_rng(9)
    pos = g.uniform(-0.5, 1.5, (2000, 2))
    d, n, hit = sc.signed_distance(torch.tensor(pos, dtype=F64, device=DEV))
    inside = sc.inside(torch.tensor(pos, dtype=F64, device=DEV)).cpu().numpy()
    lpos = min_img(pos - center)
    best = np.full(len(pos), np.inf)
    for k in itertools.product((-1, 0, 1), repeat=2):
        ci = c + np.array(k) * L
        dist = np.linalg.norm(lpos[:, None] - ci[None], axis=2) - r[None]
        best = np.minimum(best, dist.min(1))
    # the images are tiled for the reach of the particles' support (as the reference's tiledRep): within it the
    # nearest image is always listed; beyond it a farther listed disk may be the nearest one found
    got = d.cpu().numpy()
    near = best <= H
    assert near.sum() > 500
    np.testing.assert_allclose(got[near], best[near], rtol=0, atol=1e-14)
    assert (got[~near] >= best[~near] - 1e-14).all()
    assert np.array_equal(inside, best < 0)
    assert hit.all()
