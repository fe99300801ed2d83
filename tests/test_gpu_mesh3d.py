"""3-D closed-surface path on the device: face adjacency (bit-exact), winding numbers, the 22 stored moments.

References: the reference's 70-digit fixtures (tests/golden/fluxbridge-*.json, copied data), the brute-force element
test of the oracle, the host evaluation of the same arithmetic, and the compact-form FP64 oracle with the snapped
winding number.  Tolerances: |device - fixture| <= 1e-12 (1 + |fixture|) (FP64 form); device vs host evaluation of
the same arithmetic (FMA contraction differs) 1e-11 (1 + |value|) per channel in units of h; device vs FP64 oracle 1e-12 (w6: 1e-11, the
oracle's own FP64 error at that size) in units of h; mixed vs FP64 1e-6 signal-relative per channel.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import mesh3d as OM
from warpsphboundaries_amd import mesh3d as MM
from warpsphboundaries_amd.plan3d import H_POWER3, REQUESTS22

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLD = os.path.join(os.path.dirname(__file__), "golden")
KERNELS = ["cubic", "w2", "w4", "w6"]


def _fixture_cases():
    out = []
    for fn, lim in (("fluxbridge-v1.json", 44), ("fluxbridge-nearplane-v1.json", 10 ** 6)):
        d = json.load(open(os.path.join(GOLD, fn)))
        for ci, c in enumerate(d["cases"]):
            if ci < lim and c.get("dimension", 3) == 3 and c.get("measure", "volume") == "volume":
                out.append(c)
    return out


def _cavity(level=2, h=0.08):
    Vb, Fb = MM.box_surface((0, 0, 0), (1, 1, 1))
    Vs, Fs = MM.icosphere(level, 0.25, (0.5, 0.5, 0.5))
    V = np.vstack([Vb, Vs])
    F = np.vstack([Fb, (Fs + len(Vb))[:, ::-1]])
    return V, F


def _points(n, seed=1):
    rng = np.random.default_rng(seed)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    p1 = 0.5 + dirs * (0.25 + rng.uniform(-0.06, 0.06, n))[:, None]
    p2 = rng.uniform(0, 1, (n, 3))
    p2[:, 0] = rng.uniform(-0.03, 0.05, n)
    p3 = rng.uniform(-0.05, 1.05, (n, 3))
    return np.vstack([p1, p2, p3])


def _units(h):
    return np.asarray(h, dtype=np.float64)[..., None] ** np.array(H_POWER3, dtype=np.float64)


def test_fixtures_fp64_and_mixed():
    worst = 0.0
    for c in _fixture_cases():
        V = np.array([[float(x) for x in v] for v in c["vertices"]])
        q = np.array([[float(x) for x in c["query"]]])
        h = float(c["h"])
        m = MM.Mesh3D(V, OM.tetrahedron_faces(V), h, device=DEV)
        Q = torch.tensor(q, device=DEV)
        H = torch.full((1,), h, dtype=torch.float64, device=DEV)
        adj = m.adjacency(Q, H)
        ref = np.array([float(v) for v in c["values"]])[:22]
        got = m.moments(adj, Q, H, c["kernel"], "fp64").cpu().numpy()[0]
        err = np.abs(got - ref) / (1 + np.abs(ref))
        worst = max(worst, err.max())
        assert err.max() <= 1e-12, (c.get("label", c["query"]), c["kernel"], err.max(), int(err.argmax()))
        mix = m.moments(adj, Q, H, c["kernel"], "mixed").cpu().numpy()[0]
        scale = np.abs(ref).max()
        assert (np.abs(mix - ref) <= 1e-6 * scale + 1e-30).all(), (c.get("label"), c["kernel"])
    print("fixture worst", worst)


def test_adjacency_winding_bitexact_and_host_equal():
    V, F = _cavity(2)
    h = 0.08
    P = _points(400)
    H = np.full(len(P), h)
    H[::3] *= 0.6
    m = MM.Mesh3D(V, F, h, device=DEV)
    adj = m.adjacency(torch.tensor(P, device=DEV), torch.tensor(H, device=DEV))
    q, el, off = adj.numpy()
    cols, offs = OM.face_adjacency(V[F], P, H)
    assert np.array_equal(off, offs) and np.array_equal(el, cols)
    hoff, hfaces, hchi = MM.Mesh3D(V, F, h, host_only=True).adjacency_host(P, H)
    assert np.array_equal(hoff, off) and np.array_equal(hfaces, el)
    chi = adj.chi.cpu().numpy()
    np.testing.assert_allclose(chi, hchi, rtol=0, atol=1e-14)
    assert np.abs(chi - OM.winding(V, F, P)).max() < 1e-12


def test_capacity_retry():
    V, F = _cavity(2)
    m = MM.Mesh3D(V, F, 0.08, device=DEV)
    P = torch.tensor(_points(200), device=DEV)
    H = torch.full((len(P),), 0.08, dtype=torch.float64, device=DEV)
    a = m.adjacency(P, H, capacity=3)
    b = m.adjacency(P, H)
    assert a.count == b.count
    assert torch.equal(a.faces[: a.count], b.faces[: b.count])


@pytest.mark.parametrize("search,layout", [("grid", "rows"), ("grid", "sorted"), ("bvh", "rows")])
def test_support_above_support_max_is_refused(search, layout):
    # the device flags the supports in the count pass; offsets[n] = -1 makes the one host read raise, and the flag
    # starts clear again on the next call
    V, F = _cavity(2)
    m = MM.Mesh3D(V, F, 0.08, device=DEV)
    P = torch.tensor(_points(300), device=DEV)
    H = torch.full((len(P),), 0.08, dtype=torch.float64, device=DEV)
    ok = m.adjacency(P, H, search=search, layout=layout)
    for bad in (0.08 * (1 + 1e-9), float("nan")):
        Hb = H.clone()
        Hb[137] = bad
        with pytest.raises(ValueError, match="support_max"):
            m.adjacency(P, Hb, search=search, layout=layout)
    m.adjacency(P, H * (1 + 1e-13), search=search, layout=layout)        # within the 1e-12 slack
    again = m.adjacency(P, H, search=search, layout=layout)
    assert again.count == ok.count
    for x, y in zip(again.numpy()[1:], ok.numpy()[1:]):      # per particle (the sorted layout's order inside a
        assert np.array_equal(x, y)                          # cell may differ between calls)


@pytest.mark.parametrize("k", KERNELS)
def test_moments_match_host_and_oracle(k):
    V, F = _cavity(2)
    h = 0.08
    P = _points(120, seed=7)
    H = np.full(len(P), h)
    m = MM.Mesh3D(V, F, h, device=DEV)
    Q, HH = torch.tensor(P, device=DEV), torch.tensor(H, device=DEV)
    got = m.evaluate(Q, HH, k, "fp64").cpu().numpy()
    host = MM.Mesh3D(V, F, h, host_only=True).evaluate_host(P, H, k)
    u = _units(H)
    dh = np.abs(got - host) / u / (1 + np.abs(host) / u)
    assert dh.max() <= 1e-11, (k, dh.max(axis=0))
    w = OM.snapped_winding(V, F, P)
    ref = OM.closed_surface_requests(V, F, P, k, h, list(REQUESTS22), chunk=32, winding=w)
    tol = 1e-11 if k == "w6" else 1e-12
    err = np.abs(got - ref) / u / (1 + np.abs(ref) / u)
    assert err.max() <= tol, (k, err.max())
    mix = m.evaluate(Q, HH, k, "mixed").cpu().numpy()
    sig = np.abs(got / u).max(axis=0) + 1e-300
    assert (np.abs(mix - got) / u <= 1e-6 * sig).all(), (np.abs(mix - got) / u / sig).max()


def test_points_on_the_surface_take_the_exact_sum():
    V, F = MM.box_surface((0, 0, 0), (1, 1, 1))
    # on a face, on an edge, at a vertex, and just off the faces
    P = np.array([[0.0, 0.3, 0.6], [0.0, 0.0, 0.4], [0.0, 0.0, 0.0], [1e-17, 0.3, 0.6], [0.5, 0.5, 1.0], [0.5, 0.5, 0.5]])
    h = 0.1
    m = MM.Mesh3D(V, F, h, device=DEV)
    adj = m.adjacency(torch.tensor(P, device=DEV), torch.full((len(P),), h, dtype=torch.float64, device=DEV))
    chi = adj.chi.cpu().numpy()
    host = MM.Mesh3D(V, F, h, host_only=True).adjacency_host(P, h)[2]
    np.testing.assert_allclose(chi, host, rtol=0, atol=1e-14)
    np.testing.assert_allclose(chi, [0.5, 0.25, 0.125, 1.0, 0.5, 1.0], atol=1e-14)


def test_plane_limit_and_whole_support():
    # a large box: the query deep inside sees the whole ball; near one face it sees the half-space
    from warpsphboundaries_amd import sphere as S
    from warpsphboundaries_amd.plan3d import whole_vector3
    V, F = MM.box_surface((-10, -10, -10), (10, 10, 10))
    m = MM.Mesh3D(V, F, 1.0, device=DEV)
    d = np.array([0.9, 0.5, 0.1, 1e-6, 0.0])
    P = np.zeros((len(d) + 1, 3))
    P[: len(d), 2] = 10.0 - d
    H = torch.ones(len(P), dtype=torch.float64, device=DEV)
    for k in KERNELS:
        out = m.evaluate(torch.tensor(P, device=DEV), H, k, "fp64").cpu().numpy()
        np.testing.assert_allclose(out[-1], whole_vector3(k), atol=1e-13)
        lam = np.array([float(S.half_space_exact(k, -x)) for x in d])
        np.testing.assert_allclose(out[:-1, 0], lam, rtol=0, atol=1e-13)


def test_empty_and_far_particles():
    V, F = _cavity(1)
    m = MM.Mesh3D(V, F, 0.05, device=DEV)
    P = torch.tensor([[5.0, 5.0, 5.0], [0.5, 0.5, 0.5], [0.1, 0.1, 0.1]], dtype=torch.float64, device=DEV)
    H = torch.full((3,), 0.05, dtype=torch.float64, device=DEV)
    out = m.evaluate(P, H, "w2", "fp64").cpu().numpy()
    from warpsphboundaries_amd.plan3d import whole_vector3
    u = _units(np.full(3, 0.05))
    np.testing.assert_allclose(out[0], 0.0, atol=0)
    np.testing.assert_allclose(out[1], 0.0, atol=1e-15)              # inside the cavity (outside the solid)
    np.testing.assert_allclose(out[2] / u[2], whole_vector3("w2"), atol=1e-13)
    e = m.evaluate(P[:0], H[:0], "w2", "fp64")
    assert e.shape == (0, 22)


def test_evaluator_and_element_api():
    c = _fixture_cases()[4]
    V = np.array([[float(x) for x in v] for v in c["vertices"]])
    ev = MM.Evaluator(MM.Geometry([V]), MM.baseline(3)[:22], kernel=c["kernel"])
    got = ev.evaluate([[float(x) for x in c["query"]]], float(c["h"]))[0]
    ref = np.array([float(v) for v in c["values"]])[:22]
    assert (np.abs(got - ref) <= 1e-12 * (1 + np.abs(ref))).all()
    tet = MM.Tetrahedron(V, [float(x) for x in c["query"]])
    assert abs(tet.moment(c["kernel"], (1, 0, 0), float(c["h"])) - ref[1]) <= 1e-12
    g = tet.kernel_gradient(c["kernel"], (0, 1, 0), float(c["h"]))
    assert np.allclose(g, ref[16:19], atol=1e-12)
    with pytest.raises(NotImplementedError):
        tet.moment(c["kernel"], (3, 0, 0))


@pytest.mark.parametrize("n", [3000, 200000])
def test_cell_sorted_schedule_changes_nothing(n):
    """the cell order (wsb_sort3d: 49 chunks of the two-pass sort at 200,000 particles) is a permutation, and the face
    lists, winding numbers and moments served in it are bitwise those in input order"""
    V, F = _cavity(3)
    h = 0.05
    P = _points(n, seed=11)
    m = MM.Mesh3D(V, F, h, device=DEV)
    Q = torch.tensor(P, device=DEV)
    H = torch.full((len(P),), h, dtype=torch.float64, device=DEV)
    order = m.sort(Q)
    assert torch.equal(torch.sort(order.long()).values, torch.arange(len(P), device=DEV))
    a = m.adjacency(Q, H, sort=True)
    b = m.adjacency(Q, H, sort=False)
    assert a.count == b.count
    assert torch.equal(a.offsets, b.offsets) and torch.equal(a.faces[: a.count], b.faces[: b.count])
    assert torch.equal(a.chi, b.chi)
    ma = m.moments(a, Q, H, "w4", "fp64")
    mb = m.moments(b, Q, H, "w4", "fp64")
    assert torch.equal(ma, mb)


def test_heavy_rows_match_host_and_are_batch_independent():
    # a fine sphere inside a box with a support of ~10 face sizes: particles near the sphere have hundreds of faces
    V, F = MM.box_surface((0, 0, 0), (1, 1, 1))
    Vs, Fs = MM.icosphere(4, 0.25, (0.5, 0.5, 0.5))
    V, F = np.vstack([V, Vs]), np.vstack([F, (Fs + 8)[:, ::-1]])
    h = 0.12
    rng = np.random.default_rng(5)
    d = rng.normal(size=(700, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    P = np.vstack([0.5 + d * (0.25 + rng.uniform(0.0, 0.1, 700))[:, None], rng.uniform(0, 1, (300, 3))])
    H = np.full(len(P), h)
    m = MM.Mesh3D(V, F, h, device=DEV)
    Q, HH = torch.tensor(P, device=DEV), torch.tensor(H, device=DEV)
    adj = m.adjacency(Q, HH)
    counts = np.diff(adj.offsets.cpu().numpy())
    assert (counts > 64).sum() > 100 and (counts <= 64).sum() > 50, (counts.max(), (counts > 64).sum())
    u = _units(H)
    host = MM.Mesh3D(V, F, h, host_only=True)
    for k in ("w2", "w6"):
        fp = None
        for form in ("fp64", "mixed"):
            got = m.moments(adj, Q, HH, k, form).cpu().numpy()
            if form == "fp64":
                # the device arithmetic against the same arithmetic on the host, summed in order
                ref = host.evaluate_host(P, H, k, form=form)
                err = np.abs(got - ref) / u / (1 + np.abs(ref) / u)
                assert err.max() <= 1e-11, (k, form, err.max())
                fp = got
            else:
                sig = np.abs(fp / u).max(axis=0) + 1e-300
                assert (np.abs(got - fp) / u <= 1e-6 * sig).all(), (k, (np.abs(got - fp) / u / sig).max())
            # the same rows in another launch (every other particle dropped): bitwise the same values
            sub = np.arange(0, len(P), 2)
            got2 = m.evaluate(Q[sub], HH[sub], k, form).cpu().numpy()
            assert np.array_equal(got2, got[sub])


def test_wide_candidate_cells_bitexact():
    # a fine sphere (5,120 faces) and a support of ~7 face sizes: cells near the sphere list hundreds of candidate
    # faces (scanned one wavefront per particle), cells near the box walls a few (one lane per particle); both kinds
    # share wavefronts when the particles are not sorted by cell.  Likewise for the axis walks of the winding number.
    V, F = MM.box_surface((0, 0, 0), (1, 1, 1))
    Vs, Fs = MM.icosphere(4, 0.25, (0.5, 0.5, 0.5))
    V, F = np.vstack([V, Vs]), np.vstack([F, (Fs + 8)[:, ::-1]])
    h = 0.12
    rng = np.random.default_rng(8)
    d = rng.normal(size=(900, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    P = np.vstack([0.5 + d * (0.25 + rng.uniform(0.0, 0.15, 900))[:, None], rng.uniform(0, 1, (600, 3))])
    P = P[rng.permutation(len(P))]
    H = np.full(len(P), h)
    H[::4] *= 0.5
    m = MM.Mesh3D(V, F, h, device=DEV)
    # candidates of each particle's cell: faces whose box, grown by support_max, meets the cell
    lo, cell = np.array(m.grid_lo), m.cell
    c0 = lo + np.floor((P - lo) / cell) * cell
    T = V[F]
    bl, bh = T.min(1) - h, T.max(1) + h
    cand = ((bl[None] <= c0[:, None] + cell) & (bh[None] >= c0[:, None])).all(-1).sum(1)
    assert (cand > 400).sum() > 200 and (cand < 64).sum() > 50, (cand.min(), cand.max())
    assert m.longest_walk > 200 and m.dirty > 100      # axis walks over more than 128 faces: one wavefront each
    Q, HH = torch.tensor(P, device=DEV), torch.tensor(H, device=DEV)
    hoff, hfaces, hchi = MM.Mesh3D(V, F, h, host_only=True).adjacency_host(P, H)
    cols, offs = OM.face_adjacency(T, P, H)
    assert np.array_equal(hoff, offs) and np.array_equal(hfaces, cols)
    for sort in (True, False):
        adj = m.adjacency(Q, HH, sort=sort)
        q, el, off = adj.numpy()
        assert np.array_equal(off, hoff) and np.array_equal(el, hfaces), sort
        np.testing.assert_allclose(adj.chi.cpu().numpy(), hchi, rtol=0, atol=1e-14)
    short = m.adjacency(Q, HH, capacity=100)              # a face list too short at first: counted, then refilled
    assert short.count == len(hfaces) and np.array_equal(short.faces[: short.count].cpu().numpy(), hfaces)


# ------------------------------------------------------------------------------------------------ Hessian entries
def test_hessian_fixtures_fp64_and_mixed():
    """baseline(3)[22:28] on the device against the 70-digit fixtures: FP64 within 1e-12 (1 + |ref|) (w6: 5e-12, its
    kernel polynomial's coefficients enter the Hessian undivided), mixed within 1e-6 of the signal."""
    worst = 0.0
    for c in _fixture_cases():
        V = np.array([[float(x) for x in v] for v in c["vertices"]])
        h = float(c["h"])
        m = MM.Mesh3D(V, OM.tetrahedron_faces(V), h, device=DEV)
        Q = torch.tensor([[float(x) for x in c["query"]]], dtype=torch.float64, device=DEV)
        H = torch.full((1,), h, dtype=torch.float64, device=DEV)
        adj = m.adjacency(Q, H)
        ref = np.array([float(v) for v in c["values"]])[22:28]
        got = m.hessian(adj, Q, H, c["kernel"], "fp64").cpu().numpy()[0]
        err = np.abs(got - ref) / (1 + np.abs(ref))
        worst = max(worst, err.max())
        assert err.max() <= (5e-12 if c["kernel"] == "w6" else 1e-12), (c.get("label", c["query"]), c["kernel"], err.max())
        mix = m.hessian(adj, Q, H, c["kernel"], "mixed").cpu().numpy()[0]
        assert (np.abs(mix - ref) <= 1e-6 * max(np.abs(ref).max(), 1e-300)).all(), (c.get("label"), c["kernel"])
    print("Hessian fixture worst", worst)


@pytest.mark.parametrize("k", KERNELS)
def test_hessian_tiles_and_heavy_rows_match_host_and_oracle(k):
    """the tile and heavy-row launches of the Hessian against the host arithmetic (summed in order) and the oracle's
    kernel_hessian on a fine sphere in a box (hundreds of faces per particle near the sphere); the trace is the
    oracle's laplacian; physical units (h^-2) against units of h."""
    V, F = MM.box_surface((0, 0, 0), (1, 1, 1))
    Vs, Fs = MM.icosphere(3, 0.25, (0.5, 0.5, 0.5))
    V, F = np.vstack([V, Vs]), np.vstack([F, (Fs + 8)[:, ::-1]])
    h = 0.12
    rng = np.random.default_rng(9)
    d = rng.normal(size=(150, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    P = np.vstack([0.5 + d * (0.25 + rng.uniform(-0.05, 0.1, 150))[:, None], rng.uniform(-0.05, 1.05, (90, 3))])
    H = np.full(len(P), h)
    H[::4] *= 0.7
    m = MM.Mesh3D(V, F, h, device=DEV)
    Q, HH = torch.tensor(P, device=DEV), torch.tensor(H, device=DEV)
    adj = m.adjacency(Q, HH)
    counts = np.diff(adj.offsets.cpu().numpy())
    assert (counts > 64).sum() > 20 and (counts <= 64).sum() > 20
    got = m.hessian(adj, Q, HH, k, "fp64").cpu().numpy()
    host = MM.Mesh3D(V, F, h, host_only=True).hessian_host(P, H, k)
    scale = np.abs(host).max()
    # the host build contracts other products than the device: w6's undivided coefficients amplify that fivefold
    assert np.abs(got - host).max() <= (5e-12 if k == "w6" else 1e-12) * scale, (k, np.abs(got - host).max() / scale)
    units = m.hessian(adj, Q, HH, k, "fp64", physical=False).cpu().numpy()
    np.testing.assert_allclose(units / (H * H)[:, None], got, rtol=1e-14, atol=1e-14 * scale)
    cs = OM.ClosedSurfacePairs(V, F, P)
    Hs = [cs.kernel_hessian(k, (0, 0, 0), hh) for hh in sorted(set(H))]
    ref = np.zeros_like(got)
    for hh, Hm in zip(sorted(set(H)), Hs):
        sel = H == hh
        ref[sel] = np.stack([Hm[i][j] for i, j in MM._HESSIAN], 1)[sel]
    assert np.abs(got - ref).max() <= (1e-11 if k == "w6" else 2e-12) * scale, (k, np.abs(got - ref).max() / scale)
    lap = np.zeros(len(P))
    for hh in sorted(set(H)):
        sel = H == hh
        lap[sel] = cs.laplacian(k, (0, 0, 0), hh)[sel]
    np.testing.assert_allclose(got[:, 0] + got[:, 3] + got[:, 5], lap, rtol=0, atol=1e-11 * scale)
    # accumulate adds
    acc = m.hessian(adj, Q, HH, k, "fp64", out=torch.tensor(got, device=DEV), accumulate=True).cpu().numpy()
    assert np.array_equal(acc, 2 * got)


def test_hessian_element_api_and_full_baseline_evaluator():
    c = _fixture_cases()[4]
    V = np.array([[float(x) for x in v] for v in c["vertices"]])
    q = [float(x) for x in c["query"]]
    h = float(c["h"])
    ref = np.array([float(v) for v in c["values"]])
    ev = MM.Evaluator(MM.Geometry([V]), kernel=c["kernel"])
    assert len(ev.requests) == 28
    got = ev.evaluate([q], h)[0]
    gate = 5e-12 if c["kernel"] == "w6" else 1e-12
    assert (np.abs(got - ref[:28]) <= gate * (1 + np.abs(ref[:28]))).all()
    tet = MM.Tetrahedron(V, q)
    Hm = np.array(tet.kernel_hessian(c["kernel"], (0, 0, 0), h))
    assert np.allclose(Hm, Hm.T, rtol=0, atol=0)
    assert np.allclose([Hm[i, j] for i, j in MM._HESSIAN], ref[22:28], rtol=0, atol=gate * (1 + np.abs(ref[22:28]).max()))
    assert abs(tet.laplacian(c["kernel"], (0, 0, 0), h) - np.trace(Hm)) <= 1e-13 * (1 + abs(np.trace(Hm)))
    with pytest.raises(NotImplementedError):
        tet.kernel_hessian(c["kernel"], (1, 0, 0), h)


# ------------------------------------------------------------------------------------------------ face-tree search
@pytest.mark.parametrize("case", ["cavity", "heavy", "config3"])
def test_face_tree_search_equals_the_cell_lists(case):
    """wsb_adjacency3d_bvh: the same offsets, ascending face lists and winding numbers as the cell lists, on the
    cavity test points, on a fine sphere whose particles have hundreds of faces (the rows sorted through the bitmap),
    and on the level-6 cavity; a short face list is retried as for the cell lists."""
    if case == "cavity":
        V, F = _cavity(2)
        h = 0.08
        P = _points(400)
    elif case == "heavy":
        V, F = MM.box_surface((0, 0, 0), (1, 1, 1))
        Vs, Fs = MM.icosphere(4, 0.25, (0.5, 0.5, 0.5))
        V, F = np.vstack([V, Vs]), np.vstack([F, (Fs + 8)[:, ::-1]])
        h = 0.12
        rng = np.random.default_rng(5)
        d = rng.normal(size=(700, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        P = np.vstack([0.5 + d * (0.25 + rng.uniform(0.0, 0.1, 700))[:, None], rng.uniform(0, 1, (300, 3))])
    else:
        from warpsphboundaries_amd import workloads as WL
        w = WL.config3()
        V, F, h = w.vertices, w.faces, w.h
        P = w.positions[np.random.default_rng(3).choice(len(w.positions), 20000, replace=False)]
    H = np.full(len(P), h)
    H[::3] *= 0.7
    m = MM.Mesh3D(V, F, h, device=DEV)
    Q, HH = torch.tensor(P, device=DEV), torch.tensor(H, device=DEV)
    a = m.adjacency(Q, HH, search="grid")
    b = m.adjacency(Q, HH, search="bvh")
    assert a.count == b.count
    assert torch.equal(a.offsets, b.offsets)
    assert torch.equal(a.faces[: a.count], b.faces[: b.count])
    assert torch.equal(a.chi, b.chi)
    counts = np.diff(a.offsets.cpu().numpy())
    if case == "heavy":
        assert (counts > 32).sum() > 100
    c = m.adjacency(Q, HH, capacity=3, search="bvh")
    assert c.count == b.count and torch.equal(c.faces[: c.count], b.faces[: b.count])


@pytest.mark.parametrize("case", ["cavity-300k", "fine-sphere-heavy"])
def test_sorted_layout_equals_rows_layout(case):
    """wsb_adjacency3d_sorted / wsb_moments3d_sorted: the entry at position i of the cell order is particle order[i]'s
    face list, winding number and (x, y, z, h), bitwise those of wsb_adjacency3d, and the 22 moments of both forms,
    stored at the particles, bitwise those of wsb_moments3d.  The cavity at 300,000 particles keeps one lane per
    particle; the fine sphere (1,000 particles, hundreds of faces each) takes the wide-candidate, long-walk and
    heavy-row kernels."""
    if case == "cavity-300k":
        V, F = _cavity(3)
        h = 0.05
        P = _points(100000, seed=5)
    else:
        V, F = MM.box_surface((0, 0, 0), (1, 1, 1))
        Vs, Fs = MM.icosphere(4, 0.25, (0.5, 0.5, 0.5))
        V, F = np.vstack([V, Vs]), np.vstack([F, (Fs + 8)[:, ::-1]])
        h = 0.12
        rng = np.random.default_rng(6)
        d = rng.normal(size=(700, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        P = np.vstack([0.5 + d * (0.25 + rng.uniform(0.0, 0.1, 700))[:, None], rng.uniform(0, 1, (300, 3))])
    m = MM.Mesh3D(V, F, h, device=DEV)
    Q = torch.tensor(P, device=DEV)
    H = torch.full((len(P),), h, dtype=torch.float64, device=DEV)
    a = m.adjacency(Q, H)
    b = m.adjacency(Q, H, layout="sorted")
    assert b.count == a.count and b.prow is not None
    if case != "cavity-300k":
        assert (np.diff(a.offsets.cpu().numpy()) > 64).sum() > 100
    off, faces, chi = b.by_particle()
    assert torch.equal(off, a.offsets) and torch.equal(faces, a.faces[: a.count])
    assert torch.equal(chi, a.chi)
    o = b.order.long()
    assert torch.equal(b.prow[:, :3], Q[o]) and torch.equal(b.prow[:, 3], H[o])
    for x, y in zip(a.numpy(), b.numpy()):
        assert np.array_equal(x, y)
    for form in ("fp64", "mixed"):
        assert torch.equal(m.moments(a, Q, H, "w4", form), m.moments(b, Q, H, "w4", form)), form
    with pytest.raises(NotImplementedError):
        m.hessian(b, Q, H, "w4")
