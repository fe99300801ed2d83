"""Conditioned small / sliver triangle route on the device (wsb_triangle2d): against the host entry point (the same
arithmetic), inside FusedWall (rotated body, variable supports, two representations, quadrature and closed-form rows
side by side; reference tests/scene/test_complete_surface_moments.py:122-150) and in the server's per-scene path,
both against the oracle (oracle/triangle2d.py for the quadrature rows, the edge-local closed form for the others).

Tolerances: device vs host 1e-12 of the row's scaled signal (FMA contraction and summation order differ, the panel
decisions do not); FusedWall / server vs oracle 1e-10 relative + 1e-13 absolute in units of h (the quadrature rows
are the same algorithm; the closed-form rows carry the FP64 closed form's 1e-13).
"""
import numpy as np
import pytest
import torch

from oracle import triangle2d as OT
from tests.test_gpu_scene import _oracle_local
from warpsphboundaries_amd import triangle as T
from warpsphboundaries_amd.fused import FusedGroup, FusedWall, WallOutput
from warpsphboundaries_amd.scene import Body, ParticleState, Scene, SurfaceRep, fixed_adjacency

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
KERNELS = ["cubic", "w2", "w4", "w6"]
SMALL = np.array([[0.30, 0.20], [0.31, 0.203], [0.302, 0.209]])          # 1e-2 across
SLIVER = np.array([[0.1, 0.6], [1.3, 0.6 + 2e-7], [0.7, 0.6 + 5e-7]])    # aspect ~1e-6
BIG = np.array([[0.45, 0.35], [0.85, 0.40], [0.60, 0.75]])


def _scaled_close(a, b, V, meta, P, H, rel):
    for i in range(len(P)):
        s = np.abs(b[i]).max() + meta[1] / H[i] ** 2
        assert np.abs(a[i] - b[i]).max() <= rel * s, (i, np.abs(a[i] - b[i]).max(), s)


@pytest.mark.parametrize("k", KERNELS)
def test_device_matches_host_arithmetic(k):
    rng = np.random.default_rng(11)
    for V in (SMALL, SLIVER):
        for solid in ("inside", "outside"):
            rep = SurfaceRep.polygon(V, solid)
            P = V.mean(0) + rng.uniform(-0.6, 0.6, (96, 2))
            P[:8] = V.mean(0) + rng.uniform(-1e-3, 1e-3, (8, 2))
            H = rng.uniform(0.2, 1.2, len(P))
            valid = (rng.uniform(0, 1, len(P)) > 0.1).astype(np.uint8)
            Vr = rep.vertices.numpy()
            dv, de, ds = T.channels(Vr, torch.tensor(P, device=DEV), torch.tensor(H, device=DEV), rep._triangle, k,
                                    rep.background, valid=torch.tensor(valid, device=DEV))
            hv, he, hs = T.channels_host(Vr, P, H, rep._triangle, k, rep.background, valid=valid)
            ds, dv, de = ds.cpu().numpy(), dv.cpu().numpy(), de.cpu().numpy()
            assert np.array_equal(ds, hs)
            assert (hs == 1).sum() > 40                              # quadrature rows
            if V is SMALL:                                           # and rows where the triangle is not small
                assert ((hs == 0) & (valid == 1)).sum() > 10
            _scaled_close(dv, hv, Vr, rep._triangle, P, H, 1e-12)
            _scaled_close(de, he, Vr, rep._triangle, P, H, 1e-12)
    rep = SurfaceRep.polygon(SLIVER)
    v, e, s = T.channels(rep.vertices, torch.tensor([[0.5, 0.62]], device=DEV), torch.ones(1, dtype=torch.float64, device=DEV),
                         rep._triangle, k, tolerance=1e-30, max_panels=1)
    assert s.item() == -1 and torch.isnan(v).all() and torch.isfinite(e).all()


def _expected_local(reps, lp, h, k, smax):
    """[N, 12] units of h: per representation the oracle quadrature where it applies, else the closed form."""
    out = np.zeros((len(lp), 12))
    for rep in reps:
        V = rep.vertices.cpu().numpy()
        closed = _oracle_local(V, rep.background, lp, h, k, smax)
        if rep._triangle is None:
            out += closed
            continue
        qv, qe, qs = OT.channels(V, rep.edges.cpu().numpy(), lp, h, k, rep.background)
        out += np.where((qs != 0)[:, None], qv, closed)
    return out


@pytest.mark.parametrize("k", ["cubic", "w4"])
def test_fused_wall_rotated_body_two_triangles_variable_supports(k):
    angle, center = 0.43, np.array([0.7, -0.5])
    reps = [SurfaceRep.polygon(SMALL), SurfaceRep.polygon(BIG)]
    sc = Scene([Body(bodyId=0, reps=reps, center=tuple(center), angle=angle)], DEV)
    rng = np.random.default_rng(12)
    local = np.concatenate([SMALL.mean(0) + rng.uniform(-0.4, 0.4, (150, 2)), rng.uniform(0.2, 1.0, (150, 2))])
    h = rng.choice([0.3, 0.6, 1.0], len(local))
    R = np.array([[np.cos(angle), -np.sin(angle)], [np.sin(angle), np.cos(angle)]])
    ps = ParticleState(torch.tensor(local @ R.T + center, device=DEV), torch.tensor(h, device=DEV))
    wall = FusedWall(sc, fixed_adjacency(sc, ps, None, 1.0), (FusedGroup(k),))
    got = wall.channels[0, 0].cpu().numpy()
    lp = sc.bodies[0].toLocal(ps.positions).cpu().numpy()
    ref = _expected_local(reps, lp, h, k, 1.0)
    _, _, qs = OT.channels(reps[0].vertices.cpu().numpy(), reps[0].edges.cpu().numpy(), lp, h, k)
    assert (qs == 1).sum() > 50 and (qs == 0).sum() > 50            # both routes present
    # world frame, physical units
    unit = np.stack([np.ones_like(h), h, h, 1 / h, 1 / h, np.ones_like(h), np.ones_like(h), np.ones_like(h),
                     np.ones_like(h), h * h, h * h, h * h], 1)
    phys = ref * unit
    exp = phys.copy()
    exp[:, 1:3] = phys[:, 1:3] @ R.T
    exp[:, 3:5] = phys[:, 3:5] @ R.T
    exp[:, 5:9] = (R @ phys[:, 5:9].reshape(-1, 2, 2) @ R.T).reshape(-1, 4)
    m2 = R @ np.stack([phys[:, 9], phys[:, 10], phys[:, 10], phys[:, 11]], 1).reshape(-1, 2, 2) @ R.T
    exp[:, 9], exp[:, 10], exp[:, 11] = m2[:, 0, 0], m2[:, 0, 1], m2[:, 1, 1]
    np.testing.assert_allclose(got / unit, exp / unit, rtol=1e-10, atol=1e-12)
    out = wall.evaluate((WallOutput("lam", 0, "lam"),))
    np.testing.assert_allclose(out["lam"][0].cpu().numpy(), exp[:, 0], rtol=1e-10, atol=1e-12)


def test_exterior_tiny_triangle_keeps_the_whole_support_far_away():
    # reference tests/scene/test_complete_surface_moments.py:186-199: far from a tiny hole every row is the plane vector
    for k in KERNELS:
        rep = SurfaceRep.polygon(SMALL * 1e-3 + [0.2, 0.1], "outside")
        sc = Scene([Body(bodyId=0, reps=[rep])], DEV)
        pts = np.array([[2.0, 3.0], [-3.0, 2.0], [3.0, -2.0]])
        ps = ParticleState(torch.tensor(pts, device=DEV), torch.ones(3, dtype=torch.float64, device=DEV))
        got = FusedWall(sc, fixed_adjacency(sc, ps, None, 1.0), (FusedGroup(k),)).channels[0, 0].cpu().numpy()
        np.testing.assert_allclose(got, np.broadcast_to(T.plane_channels(k), got.shape), rtol=0, atol=3e-15)


def test_served_scene_with_a_conditioned_triangle():
    from warpsphboundaries_amd.serve import BoundaryServer
    box = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    rng = np.random.default_rng(13)
    pts = np.concatenate([SMALL.mean(0) + rng.uniform(-0.05, 0.05, (300, 2)), rng.uniform(0.02, 0.98, (300, 2))])
    h = 0.6
    for batched in (True, False):
        srv = BoundaryServer(streams=2, batched=batched, form="fp64")
        srv.add_scene("tri", [(box, "outside"), (SMALL, "inside")], h)
        srv.add_scene("plain", [(box, "outside")], h)
        assert srv.scenes["tri"].conditioned and not srv.scenes["plain"].conditioned
        reqs = [srv.submit("tri", pts, h, "w4"), srv.submit("plain", pts, h, "w4"), srv.submit("tri", pts[::2], h, "w4")]
        srv.run()
        reps = srv.scenes["tri"].reps
        ref = _expected_local(reps, pts, np.full(len(pts), h), "w4", h)
        unit = np.array([1, h, h, 1 / h, 1 / h, 1, 1, 1, 1, h * h, h * h, h * h])
        np.testing.assert_allclose(reqs[0].result / unit, ref, rtol=1e-10, atol=1e-12)
        assert np.array_equal(reqs[2].result, reqs[0].result[::2])
        ref_plain = _expected_local(srv.scenes["plain"].reps, pts, np.full(len(pts), h), "w4", h)
        np.testing.assert_allclose(reqs[1].result / unit, ref_plain, rtol=1e-10, atol=1e-12)
        if batched:
            with pytest.raises(NotImplementedError):
                srv.submit("tri", torch.tensor(pts, device=DEV), h, "w4")
