"""Scene-level contract on the device: provider / FusedWall outputs against the oracle, rotation covariance,
whole-support complement of an exterior solid, h-scaling, signed distance."""
import numpy as np
import pytest
import torch

from oracle import adjacency as oadj
from oracle import edge2d
from warpsphboundaries_amd import workloads as WL
from warpsphboundaries_amd.fused import FusedGroup, FusedWall, WallOutput
from warpsphboundaries_amd.provider import AnalyticBoundary, BoundaryProvider
from warpsphboundaries_amd.scene import Body, ParticleState, Scene, SurfaceRep, fixed_adjacency

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
OUTS = tuple(WallOutput(k, 0, k) for k in ["lam", "m1", "g0", "cov", "m2"])
NOTCH = np.array([[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.52, 0.9], [0.5, 0.3], [0.48, 0.9], [0.1, 0.9]])


def _oracle_local(loop_ccw, background, pts, sup, kname, smax):
    st = oadj.Static(loop_ccw, np.stack([np.arange(len(loop_ccw)), (np.arange(len(loop_ccw)) + 1) % len(loop_ccw)], 1),
                     background, smax)
    e, c, ind = oadj.fixed_topology(st, pts, sup)
    K = st.K
    rows = np.repeat(np.arange(len(pts)), c)
    sl = np.concatenate([e[r * K:r * K + k] for r, k in enumerate(c)])
    P = (st.verts[st.edges[sl, 0]] - pts[rows]) / sup[rows, None]
    Q = (st.verts[st.edges[sl, 1]] - pts[rows]) / sup[rows, None]
    out = np.zeros((len(pts), 12))
    np.add.at(out, rows, edge2d.edge_channels12(P, Q, kname))
    return out + ind[:, None] * edge2d.plane_vector(kname)[None]


def _scene(loop, solid, center=(0.0, 0.0), angle=0.0):
    rep = SurfaceRep.polygon(loop, solid)
    return Scene([Body(bodyId=0, reps=[rep], center=center, angle=angle)], DEV), rep


def test_fused_wall_rotated_body_matches_oracle():
    angle, center = 0.43, np.array([0.7, -0.5])
    sc, rep = _scene(NOTCH, "inside", tuple(center), angle)
    g = WL.rng(4)
    local = g.uniform(0.0, 1.0, (3000, 2))
    R = np.array([[np.cos(angle), -np.sin(angle)], [np.sin(angle), np.cos(angle)]])
    world = local @ R.T + center
    h = np.where(g.uniform(0, 1, len(local)) < 0.5, 0.05, 0.03)
    ps = ParticleState(torch.tensor(world, device=DEV), torch.tensor(h, device=DEV))
    wall = FusedWall(sc, fixed_adjacency(sc, ps, None, 0.05), (FusedGroup("w4"),))
    got = wall.evaluate(OUTS)
    lp = sc.bodies[0].toLocal(ps.positions).cpu().numpy()
    ref = _oracle_local(rep.vertices.cpu().numpy(), 0, lp, h, "w4", 0.05)
    # physical units, world frame
    m1 = (ref[:, 1:3] * h[:, None]) @ R.T
    g0 = (ref[:, 3:5] / h[:, None]) @ R.T
    cov = R @ ref[:, 5:9].reshape(-1, 2, 2) @ R.T
    m2 = R @ (np.stack([ref[:, 9], ref[:, 10], ref[:, 10], ref[:, 11]], 1).reshape(-1, 2, 2) * (h * h)[:, None, None]) @ R.T
    np.testing.assert_allclose(got["lam"][0].cpu().numpy(), ref[:, 0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(got["m1"][0].cpu().numpy(), m1, rtol=0, atol=1e-13)
    np.testing.assert_allclose(got["g0"][0].cpu().numpy(), g0, rtol=0, atol=1e-10)
    np.testing.assert_allclose(got["cov"][0].cpu().numpy(), cov, rtol=0, atol=1e-12)
    np.testing.assert_allclose(got["m2"][0].cpu().numpy(), m2, rtol=0, atol=1e-14)


def test_exterior_solid_complements_to_the_plane_vector():
    """inside + outside = the whole-support vector (reference tests/scene/test_complete_surface_moments.py:53-70)."""
    g = WL.rng(8)
    pts = g.uniform(0.0, 1.0, (2000, 2))
    h = 0.07
    for kname in ["cubic", "w2", "w4", "w6"]:
        outs = []
        for solid in ("inside", "outside"):
            sc, rep = _scene(NOTCH, solid)
            ps = ParticleState(torch.tensor(pts, device=DEV), torch.full((len(pts),), h, device=DEV, dtype=torch.float64))
            wall = FusedWall(sc, fixed_adjacency(sc, ps, None, h), (FusedGroup(kname),))
            outs.append(wall.channels[0, 0].cpu().numpy())
        unit = np.array([1, h, h, 1 / h, 1 / h, 1, 1, 1, 1, h * h, h * h, h * h])
        tot = (outs[0] + outs[1]) / unit                               # back to units of h
        tol = 5e-12 if kname == "w6" else 1e-12
        np.testing.assert_allclose(tot, np.broadcast_to(edge2d.plane_vector(kname), tot.shape), rtol=0, atol=tol)


def test_provider_protocol_and_lean_aggregate():
    sc, rep = _scene(NOTCH, "inside")
    ab = AnalyticBoundary(sc)
    assert isinstance(ab, BoundaryProvider)
    g = WL.rng(2)
    pts = torch.tensor(g.uniform(0, 1, (1000, 2)), device=DEV)
    ps = ParticleState(pts, torch.full((1000,), 0.04, device=DEV, dtype=torch.float64), kinds=torch.zeros(1000, dtype=torch.int32, device=DEV))
    fw = ab.aggregate(ps, 0.04, kernel="w2", lean=True)
    assert set(fw.out) == {"lam", "G", "Cov"}
    assert fw.out["lam"].shape == (1, 1000) and fw.out["Cov"].shape == (1, 1000, 2, 2)
    full = ab.aggregate(ps, 0.04, kernel="w2")                  # the full layout (cover and tensile groups)
    assert set(full.out) == {"lam", "G", "Cov", "cover", "tens"}
    torch.testing.assert_close(full.out["lam"], fw.out["lam"], rtol=0, atol=0)
    with pytest.raises(NotImplementedError):
        ab.aggregate(ps, 0.04, kernel="w6")                    # the wall operations: Wendland C2 and C4 only
    # lap identity 2 lam - tr cov
    lap = fw.evaluate((WallOutput("lap", 0, "lap"),))["lap"]
    torch.testing.assert_close(lap, 2 * fw.out["lam"] - fw.out["Cov"][..., 0, 0] - fw.out["Cov"][..., 1, 1])


def test_support_scaling():
    """m_alpha(T*s, x*s, h*s) = s^|alpha| m_alpha(T, x, h)."""
    g = WL.rng(5)
    pts = g.uniform(0.05, 0.95, (500, 2))
    res = []
    for s in (1.0, 3.0):
        sc, rep = _scene(NOTCH * s, "inside")
        ps = ParticleState(torch.tensor(pts * s, device=DEV), torch.full((500,), 0.08 * s, device=DEV, dtype=torch.float64))
        wall = FusedWall(sc, fixed_adjacency(sc, ps, None, 0.08 * s), (FusedGroup("w6"),))
        res.append(wall.channels[0, 0].cpu().numpy() / (0.08 * s) ** np.array([0, 1, 1, -1, -1, 0, 0, 0, 0, 2, 2, 2]))
    # in units of h the two scenes are the same scene
    np.testing.assert_allclose(res[1], res[0], rtol=0, atol=5e-12)


def test_signed_distance_matches_brute_force():
    sc, rep = _scene(NOTCH, "inside")
    g = WL.rng(12)
    pts = g.uniform(0.0, 1.0, (4000, 2))
    d, n, hit = sc.signed_distance(torch.tensor(pts, device=DEV), supportMax=0.05)
    V = rep.vertices.cpu().numpy()
    a, b = V, np.roll(V, -1, 0)
    e = b - a
    t = np.clip(((pts[:, None] - a[None]) * e[None]).sum(2) / (e * e).sum(1)[None], 0, 1)
    dist = np.linalg.norm(pts[:, None] - (a[None] + t[..., None] * e[None]), axis=2).min(1)
    ins = WL.inside_polygon(pts, V)
    np.testing.assert_allclose(d.cpu().numpy(), np.where(ins, -dist, dist), rtol=0, atol=1e-14)
    assert hit.all()
    np.testing.assert_allclose(np.linalg.norm(n.cpu().numpy(), axis=1), 1.0, atol=1e-14)
