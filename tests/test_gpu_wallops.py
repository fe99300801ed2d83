"""Wall-operation kernel groups of the full aggregate layout on the device, against the oracle and the reference's
own brute-force integrals (tests/scene/test_tensile_scene.py, test_viscosity_scene.py, test_cover_scene.py of the
reference):

* cone, lw2, lw4 (the FP64 closed form, like the scheme kernels): all twelve channels vs the oracle's FP64 edge
  reduction, 1e-12 in units of h (w-like kernels are within 1e-13; 10x margin);
* w2p5, w4p5 (gradient channels by chord quadrature): vs the oracle's edge reduction in 50-digit mpmath (the FP64
  monomial form itself is 3e-8 / 1e-3 off here), 1e-12 of max |G|;
* the reference's own golden values (its tests' smoke values, rtol 1e-6 / atol 1e-6 as there): the tensile vector
  above a floor, the wall Laplacian above a floor at six heights, the cover vector of a unit square; the tensile
  vector vs a 2000 x 2000 midpoint grid of W^4 grad W (5e-5, the reference's gate) and the wall Laplacian vs a polar
  midpoint grid of lap W (1e-3); signs (cover away from the wall, T into it, Delta-lambda > 0 on the fluid side) and
  zeros beyond H and deep inside the solid; the full aggregate layout's raw outputs times the factors equal the
  scene routes.
"""
import math

import numpy as np
import pytest
import torch

from oracle import adjacency as oadj
from oracle import arith as A
from oracle import edge2d
from warpsphboundaries_amd import wallops as WO
from warpsphboundaries_amd.fused import FusedGroup, FusedWall
from warpsphboundaries_amd.provider import AnalyticBoundary
from warpsphboundaries_amd.scene import Body, ParticleState, Scene, SurfaceRep, fixed_adjacency

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
LS = np.array([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], dtype=float)
FLOOR = [(-2.0, -2.0), (2.0, -2.0), (2.0, 0.0), (-2.0, 0.0)]


def _oracle(loop_ccw, background, pts, sup, k, smax, mode="f64"):
    st = oadj.Static(loop_ccw, np.stack([np.arange(len(loop_ccw)), (np.arange(len(loop_ccw)) + 1) % len(loop_ccw)], 1),
                     background, smax)
    e, c, ind = oadj.fixed_topology(st, pts, sup)
    K = st.K
    rows = np.repeat(np.arange(len(pts)), c)
    sl = np.concatenate([e[r * K:r * K + n] for r, n in enumerate(c)])
    P = (st.verts[st.edges[sl, 0]] - pts[rows]) / sup[rows, None]
    Q = (st.verts[st.edges[sl, 1]] - pts[rows]) / sup[rows, None]
    with A.precision(mode, 50):
        ch = A.to_float(edge2d.edge_channels12(A.asarray(P), A.asarray(Q), k))
    out = np.zeros((len(pts), 12))
    np.add.at(out, rows, ch)
    return out + ind[:, None] * np.asarray(edge2d.plane_vector(k))[None], len(rows)


def _channels(loop, k, pts, h, channels=None):
    rep = SurfaceRep.polygon(loop)
    sc = Scene([Body(bodyId=0, reps=[rep])], DEV)
    ps = ParticleState(torch.tensor(pts, device=DEV), torch.tensor(h, device=DEV))
    wall = FusedWall(sc, fixed_adjacency(sc, ps, None, float(h.max())), (FusedGroup(k, channels, "fp64"),))
    unit = np.stack([np.ones_like(h), h, h, 1 / h, 1 / h, np.ones_like(h), np.ones_like(h), np.ones_like(h),
                     np.ones_like(h), h * h, h * h, h * h], 1)
    return wall.channels[0, 0].cpu().numpy() / unit, rep.vertices.cpu().numpy()


@pytest.mark.parametrize("k", ["cone", "lw2", "lw4"])
def test_aux_kernels_match_oracle(k):
    rng = np.random.default_rng(31)
    pts = rng.uniform(-0.5, 2.5, (1500, 2))
    h = rng.choice([0.3, 0.6], len(pts))
    got, V = _channels(LS, k, pts, h)
    ref, npairs = _oracle(V, 0, pts, h, k, 0.6)
    assert npairs > 1000
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("k", ["w2p5", "w4p5"])
def test_tensile_gradient_matches_mpmath(k):
    rng = np.random.default_rng(32)
    pts = rng.uniform(-0.4, 2.4, (50, 2))
    h = np.full(len(pts), 0.6)
    got, V = _channels(LS, k, pts, h, (3, 4))
    ref, npairs = _oracle(V, 0, pts, h, k, 0.6, mode="mp")
    assert npairs > 50
    G, R = got[:, 3:5], ref[:, 3:5]
    scale = np.abs(R).max()
    assert scale > 0 and np.abs(G - R).max() <= 1e-12 * scale, (np.abs(G - R).max(), scale)
    assert not np.any(got[:, [0, 1, 2, 5, 6, 7, 8, 9, 10, 11]])       # the gradient channels only


def _w(fam, r, H):
    """the normalised Wendland kernel, W'(r) and lap W = W'' + W'/r (physical units)."""
    q = r / H
    u = 1 - q
    if fam == "w2":
        c = 7.0 / (math.pi * H * H)
        W = c * u ** 4 * (1 + 4 * q)
        dW = c / H * (-20 * q * u ** 3)
        d2W = c / H ** 2 * (-20 * u ** 3 + 60 * q * u ** 2)
    else:
        c = 9.0 / (math.pi * H * H)
        W = c * u ** 6 * (1 + 6 * q + 35 / 3 * q * q)
        dW = c / H * (-56 / 3 * q * u ** 5 * (1 + 5 * q))
        d2W = c / H ** 2 * (-56 / 3) * (u ** 5 * (1 + 5 * q) - 5 * q * u ** 4 * (1 + 5 * q) + 5 * q * u ** 5)
    return W, dW, d2W + dW / np.maximum(r, 1e-300)


def _dense_T(fam, pz, H, n=2000):
    """int over the floor's solid (y' < 0) within H of W^4 grad_x W, particle at (0, pz): midpoint grid."""
    xp = -H + (np.arange(n) + 0.5) / n * (2 * H)
    yp = -H + (np.arange(n) + 0.5) / n * H
    X, Y = np.meshgrid(xp, yp, indexing="ij")
    rx, ry = -X, pz - Y
    r = np.sqrt(rx * rx + ry * ry)
    m = r < H
    W, dW, _ = _w(fam, r[m], H)
    dA = (2 * H / n) * (H / n)
    return np.array([np.sum(W ** 4 * dW * rx[m] / r[m]), np.sum(W ** 4 * dW * ry[m] / r[m])]) * dA


def _polar_lap(fam, pz, H, nr=1000, nt=2000):
    """int over the floor's solid of lap W, particle at height pz: polar midpoint grid about the particle."""
    r = (np.arange(nr) + 0.5) / nr * H
    t = (np.arange(nt) + 0.5) / nt * 2 * math.pi
    R, T = np.meshgrid(r, t, indexing="ij")
    solid = pz + R * np.sin(T) < 0
    _, _, lap = _w(fam, R, H)
    return float(np.sum(lap * R * solid) * (H / nr) * (2 * math.pi / nt))


# golden values of the reference's own tests (tests/scene/test_tensile_scene.py:88, test_tensile_scene_c4.py:94,
# test_viscosity_scene.py:155-158, test_cover_scene.py:66): the exact routes' results at these points
TENSILE_SMOKE = {"w2": [(0.0, 0.3, 1.0, -0.141011574), (0.0, 0.15, 0.5, -72.1979258)],
                 "w4": [(0.0, 0.3, 1.0, -0.1978237590), (0.0, 0.15, 0.5, -101.2857646)]}
LAP_SMOKE = {"w2": {0.02: 0.4417774, 0.1: 1.959941, 0.3: 3.190942, 0.5: 1.943131, 0.7: 0.5164654, 0.9: 0.01570062},
             "w4": {0.02: 0.6086062, 0.1: 2.765873, 0.3: 4.083071, 0.5: 1.708461, 0.7: 0.2060951, 0.9: 8.409243e-4}}


@pytest.mark.parametrize("fam", ["w2", "w4"])
def test_tensile_floor_matches_reference_smoke_and_grid(fam):
    sc = Scene([Body(bodyId=0, reps=[SurfaceRep.polygon(FLOOR)])], DEV)
    for px, py, H, smoke in TENSILE_SMOKE[fam]:
        T = WO.tensile_vector_scene(sc, np.array([[px, py]]), H, fam).cpu().numpy()[0]
        assert abs(T[1] - smoke) <= 1e-6 * abs(smoke), (fam, T, smoke)
        assert T[1] < 0 and abs(T[0]) <= 1e-12 * abs(T[1])
        grid = _dense_T(fam, py, H)
        assert abs(T[1] - grid[1]) <= 5e-5 * abs(grid[1]), (T, grid)


@pytest.mark.parametrize("fam", ["w2", "w4"])
def test_wall_laplacian_floor_matches_reference_smoke(fam):
    floor = [(-5.0, -2.0), (5.0, -2.0), (5.0, 0.0), (-5.0, 0.0)]
    sc = Scene([Body(bodyId=0, reps=[SurfaceRep.polygon(floor)])], DEV)
    zs = sorted(LAP_SMOKE[fam])
    got = WO.lap_lambda_scene(sc, np.array([[0.0, z] for z in zs]), 1.0, fam)[0].cpu().numpy()
    for g, z in zip(got, zs):
        ref = LAP_SMOKE[fam][z]
        assert g > 0 and abs(g - ref) <= 1e-6 * ref, (fam, z, g, ref)
    H = 0.7
    got = WO.lap_lambda_scene(sc, np.array([[0.0, 0.3 * H], [0.0, 0.6 * H]]), H, fam)[0].cpu().numpy()
    for g, z in zip(got, (0.3 * H, 0.6 * H)):
        ref = _polar_lap(fam, z, H)
        assert abs(g - ref) <= 1e-3 * abs(ref), (g, ref)
    # beyond the support and deep inside the solid: zero (the indicator term cancels: 2 ind - tr(ind I) = 0)
    far = WO.lap_lambda_scene(sc, np.array([[0.0, 1.2], [0.0, -1.0]]), 1.0, fam)[0].cpu().numpy()
    assert np.abs(far).max() <= 1e-12


def test_cover_unit_square_matches_reference_smoke():
    sc = Scene([Body(bodyId=0, reps=[SurfaceRep.polygon([(0, 0), (1, 0), (1, 1), (0, 1)])])], DEV)
    got = WO.cover_vector_scene(sc, np.array([[0.3, 0.4], [1.3, 0.5]]), 1.0).cpu().numpy()
    np.testing.assert_allclose(got, [[-0.3457904, -0.1702250], [0.5929108, 0.0]], atol=1e-6)


@pytest.mark.parametrize("fam", ["w2", "w4"])
def test_full_aggregate_layout_outputs(fam):
    H = 0.6
    sc = Scene([Body(bodyId=0, reps=[SurfaceRep.polygon(FLOOR)])], DEV)
    pts = np.array([[0.0, 0.18], [0.1, 0.42], [0.0, 0.75]])           # two within the support, one beyond it
    ps = ParticleState(torch.tensor(pts, device=DEV), torch.full((len(pts),), H, dtype=torch.float64, device=DEV))
    ab = AnalyticBoundary(sc, form="fp64")
    fw = ab.aggregate(ps, H, kernel=fam, laplacian=True)
    assert set(fw.out) == {"lam", "G", "Cov", "cover", "lap", "tens"}
    assert fw.out["cover"].shape == (1, 3, 2) and fw.out["lap"].shape == (1, 3) and fw.out["tens"].shape == (1, 3, 2)
    np.testing.assert_allclose(WO.tensile_factor(H, fam) * fw.out["tens"][0].cpu().numpy(),
                               WO.tensile_vector_scene(sc, pts, H, fam).cpu().numpy(), rtol=1e-14, atol=0)
    np.testing.assert_allclose(WO.cover_factor(H) * fw.out["cover"][0].cpu().numpy(),
                               WO.cover_vector_scene(sc, pts, H).cpu().numpy(), rtol=1e-14, atol=0)
    np.testing.assert_allclose(WO.lap_factor(H, fam) * fw.out["lap"][0].cpu().numpy(),
                               WO.lap_lambda_scene(sc, pts, H, fam)[0].cpu().numpy(), rtol=1e-14, atol=0)
    cover, tens = fw.out["cover"][0].cpu().numpy(), fw.out["tens"][0].cpu().numpy()
    assert (WO.cover_factor(H) * cover[:2, 1] > 0).all() and (tens[:2, 1] < 0).all()
    assert not cover[2].any() and not tens[2].any() and fw.out["lap"][0, 2].item() == 0.0
    lean = ab.aggregate(ps, H, kernel=fam, lean=True)
    assert set(lean.out) == {"lam", "G", "Cov"}
    torch.testing.assert_close(lean.out["lam"], fw.out["lam"], rtol=0, atol=0)
