"""Product surfaces beside the fused pass, each against the oracle or the reference's own fixtures and tests.

* the ``edge`` element API (value, gradient, moment, moment_gradient) on every golden element, moments up to order 4
  (the pattern and gates of the reference's tests/edge/test_warp2d.py:14-26);
* ``edge_channels`` / ``second_channels`` layout and options against the oracle's edge-local channels;
* the ``a1g1`` output kind of ``FusedWall.evaluate`` against the oracle (sum_d a1_d cov_dj, world frame);
* ``FusedWall`` captured in a HIP graph: a replay after ``copy_from`` equals a fresh evaluation, a warm build and
  evaluation make no host synchronisation, and a graph of the whole build replays the eager result after the positions
  move (reference tests/scene/test_complete_surface_moments.py:86-110, tests/scene/test_fixed_adjacency.py:66-78);
* the broadphase adjacency (``Scene.adjacency``, ``aggregate(fixedAdjacency=False)``) against the fixed-capacity one;
* ``BoundaryServer.cancel`` of a pending and of an in-flight request.
"""
import math

import numpy as np
import pytest
import torch
from fractions import Fraction as F

from oracle import adjacency as oadj
from oracle import edge2d
from tests.golden_data import ELEMENTS, KERNELS
from warpsphboundaries_amd import edge as EDGE
from warpsphboundaries_amd import workloads as WL
from warpsphboundaries_amd.fused import FusedGroup, FusedWall, WallOutput
from warpsphboundaries_amd.provider import AnalyticBoundary
from warpsphboundaries_amd.scene import Body, ParticleState, Scene, SurfaceRep, fixed_adjacency

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
NOTCH = np.array([[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.52, 0.9], [0.5, 0.3], [0.48, 0.9], [0.1, 0.9]])
OUTS = tuple(WallOutput(k, 0, k) for k in ["lam", "m1", "g0", "cov", "m2"])


def _fixture_geometry(recs):
    """elements of one vertex count: (v - x) exact, rounded once to float64 (physical units), the query at the origin;
    h per element."""
    V, H = [], []
    for c in recs:
        x = (F(c["x"][0]), F(c["x"][1]))
        V.append([[float(F(a) - x[0]), float(F(b) - x[1])] for a, b in c["polygon"]])
        H.append(float(F(c["h"])))
    return np.array(V), np.zeros((len(recs), 2)), np.array(H)


def _by_size():
    sizes = sorted({len(c["polygon"]) for c in ELEMENTS})
    return [[c for c in ELEMENTS if len(c["polygon"]) == n] for n in sizes]


def _fx(c, k, what, key=None):
    e = c["kernels"][k]
    num = lambda s: float(F(s))
    if what == "value":
        return num(e["value"])
    if what == "grad":
        return np.array([num(s) for s in e["grad"]])
    if what == "m":
        return num(e["moments"][key])
    return np.array([num(s) for s in e["moment_grad"][key]])


@pytest.mark.parametrize("k", KERNELS)
def test_edge_api_on_every_golden_element(k):
    keys = sorted(ELEMENTS[0]["kernels"][k]["moments"])                   # every alpha of order 1 .. 4
    assert len(keys) == 14
    seen = 0
    for recs in _by_size():                                               # one batch per vertex count
        V, X, H = _fixture_geometry(recs)
        seen += len(recs)
        val = EDGE.value(V, X, k, H).cpu().numpy()
        gr = EDGE.gradient(V, X, k, H).cpu().numpy()
        for i, c in enumerate(recs):
            assert abs(val[i] - _fx(c, k, "value")) < 1e-12, c["id"]
            assert np.abs(gr[i] - _fx(c, k, "grad")).max() * H[i] < 1e-11, c["id"]
        for key in keys:
            al = tuple(int(v) for v in key.split(","))
            m = EDGE.moment(V, X, k, al, H).cpu().numpy()
            for i, c in enumerate(recs):
                assert abs(m[i] - _fx(c, k, "m", key)) / H[i] ** sum(al) < 1e-12, (c["id"], al, k)
        for key in sorted(ELEMENTS[0]["kernels"][k]["moment_grad"]):      # every alpha of order 0 .. 2
            al = tuple(int(v) for v in key.split(","))
            g = EDGE.moment_gradient(V, X, k, al, H).cpu().numpy()
            for i, c in enumerate(recs):
                err = np.abs(g[i] - _fx(c, k, "mg", key)).max() * H[i] ** (1 - sum(al))
                assert err < 1e-11, (c["id"], al, k, err)
    assert seen == len(ELEMENTS) == 87


@pytest.mark.parametrize("form", ["plain", "fp64", "mixed"])
def test_edge_channels_layout_and_options(form):
    g = WL.rng(5)
    star = WL.star_polygon(96, g=g)
    n = len(star)
    edges = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
    q = g.uniform(0.15, 0.85, (300, 2))
    sup = g.uniform(0.02, 0.06, 300)
    pq = np.repeat(np.arange(300), n)
    pe = np.tile(np.arange(n), 300)
    for k in KERNELS:
        c9 = EDGE.edge_channels(pq, pe, q, sup, star, edges, k, form=form).cpu().numpy()
        c3 = EDGE.second_channels(pq, pe, q, sup, star, edges, k, form=form).cpu().numpy()
        P = (star[edges[pe, 0]] - q[pq]) / sup[pq, None]
        Q = (star[edges[pe, 1]] - q[pq]) / sup[pq, None]
        ref = edge2d.edge_channels12(P, Q, k)
        got = np.concatenate([c9, c3], 1)
        err = np.abs(got - ref)
        if form == "mixed":
            assert (err.max(0) / np.maximum(np.abs(ref).max(0), 1e-300)).max() <= 1e-6, k
        else:
            assert err.max() <= (5e-12 if k == "w6" else 1e-12), (k, err.max())
        sub = EDGE.edge_channels(pq, pe, q, sup, star, edges, k, form=form, channels=(0, 3, 4), as_torch=False)
        assert isinstance(sub, np.ndarray)
        keep = np.zeros(9, dtype=bool)
        keep[[0, 3, 4]] = True
        assert np.array_equal(sub[:, keep], c9[:, keep]) and not sub[:, ~keep].any()
    np.testing.assert_array_equal(EDGE.indicator_vector("w4"), [1, 0, 0, 0, 0, 1, 0, 0, 1])
    with pytest.raises(NotImplementedError):
        EDGE.edge_channels(pq, pe, q, sup, star, edges, "w2", stable=(8, 6))


def _oracle_world(sc, rep, ps, h, kname, smax):
    """world-frame physical channels of a one-body scene from the oracle."""
    body = sc.bodies[0]
    lp = body.toLocal(ps.positions).cpu().numpy()
    st = oadj.Static(rep.vertices.cpu().numpy(), rep.edges.cpu().numpy(), rep.background, smax)
    e, c, ind = oadj.fixed_topology(st, lp, h)
    K = st.K
    rows = np.repeat(np.arange(len(lp)), c)
    sl = e.reshape(-1, K)[np.arange(K)[None, :] < c[:, None]]
    P = (st.verts[st.edges[sl, 0]] - lp[rows]) / h[rows, None]
    Q = (st.verts[st.edges[sl, 1]] - lp[rows]) / h[rows, None]
    ref = np.zeros((len(lp), 12))
    np.add.at(ref, rows, edge2d.edge_channels12(P, Q, kname))
    ref += ind[:, None] * edge2d.plane_vector(kname)[None]
    R = body.pose.R.cpu().numpy()
    cov = R @ ref[:, 5:9].reshape(-1, 2, 2) @ R.T
    return ref, cov


def test_a1g1_against_the_oracle():
    angle, center = 0.43, (0.7, -0.5)
    rep = SurfaceRep.polygon(NOTCH, "inside")
    sc = Scene([Body(bodyId=0, reps=[rep], center=center, angle=angle)], DEV)
    g = WL.rng(6)
    local = g.uniform(0.0, 1.0, (2500, 2))
    R = np.array([[math.cos(angle), -math.sin(angle)], [math.sin(angle), math.cos(angle)]])
    world = local @ R.T + np.array(center)
    h = np.where(g.uniform(0, 1, len(local)) < 0.5, 0.05, 0.03)
    ps = ParticleState(torch.tensor(world, device=DEV), torch.tensor(h, device=DEV))
    wall = FusedWall(sc, fixed_adjacency(sc, ps, None, 0.05), (FusedGroup("w4", None, "fp64"),))
    a1 = g.normal(size=(1, len(local), 2))
    got = wall.evaluate((WallOutput("A", 0, "a1g1"),), a1=torch.tensor(a1, device=DEV))["A"].cpu().numpy()
    _, cov = _oracle_world(sc, rep, ps, h, "w4", 0.05)
    ref = np.einsum("nd,ndj->nj", a1[0], cov)
    np.testing.assert_allclose(got[0], ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()))
    with pytest.raises(ValueError):
        wall.evaluate((WallOutput("A", 0, "a1g1"),))


def test_fused_wall_graph_replay_copy_from_and_no_host_sync():
    rep = SurfaceRep.polygon(NOTCH, "inside")
    sc = Scene([Body(bodyId=0, reps=[rep], center=(0.1, 0.2), angle=0.3)], DEV)
    g = WL.rng(7)
    pos = torch.tensor(g.uniform(0.0, 1.2, (3000, 2)), device=DEV)
    sup = torch.full((3000,), 0.04, dtype=torch.float64, device=DEV)
    ps = ParticleState(pos, sup)

    def build(state):
        return FusedWall(sc, fixed_adjacency(sc, state, None, 0.04), (FusedGroup("w4", None, "mixed"),))

    current = build(ps)
    current.evaluate(OUTS)
    current.reset_derived()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        with torch.cuda.graph(graph, stream=stream):
            out = current.evaluate(OUTS)
    torch.cuda.current_stream().wait_stream(stream)
    ps.positions.add_(torch.tensor([0.02, 0.03], device=DEV))
    updated = build(ps)
    expected = updated.evaluate(OUTS)
    current.copy_from(updated)
    graph.replay()
    torch.cuda.synchronize()
    for key in expected:
        assert torch.equal(out[key], expected[key]), key
    # warm: the build and the evaluation make no host synchronisation
    torch.cuda.set_sync_debug_mode("error")
    try:
        another = build(ps)
        another.evaluate(OUTS)
    finally:
        torch.cuda.set_sync_debug_mode("default")
    # the whole build captured once; replays after the positions move, eager work between them
    full = torch.cuda.CUDAGraph()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        with torch.cuda.graph(full, stream=stream):
            captured = build(ps).evaluate(OUTS)
    torch.cuda.current_stream().wait_stream(stream)
    for step in range(3):
        ps.positions.add_(torch.tensor([-0.01, 0.04], device=DEV) * (step + 1))
        eager = build(ps).evaluate(OUTS)
        torch.cuda.synchronize()
        full.replay()
        torch.cuda.synchronize()
        for key in eager:
            assert torch.equal(captured[key], eager[key]), (step, key)


def test_broadphase_adjacency_route_equals_the_fixed_capacity_route():
    """two bodies, one unbounded; particles partly outside the broadphase of the obstacle and some not fluid."""
    tank = SurfaceRep.box((0.0, 0.0), (2.0, 1.0), solid="outside")
    obstacle = SurfaceRep.polygon(NOTCH * 0.4 + np.array([0.7, 0.3]), "inside")
    sc = Scene([Body(bodyId=0, reps=[tank]), Body(bodyId=1, reps=[obstacle], center=(0.05, 0.0), angle=0.2)], DEV)
    g = WL.rng(8)
    n = 4000
    pos = torch.tensor(g.uniform(0.0, [2.0, 1.0], (n, 2)), device=DEV)
    sup = torch.tensor(np.where(g.uniform(0, 1, n) < 0.5, 0.05, 0.03), device=DEV)
    kinds = torch.tensor((g.uniform(0, 1, n) < 0.05).astype(np.int32), device=DEV)
    ps = ParticleState(pos, sup, kinds=kinds)
    ab = AnalyticBoundary(sc, form="fp64")
    a = ab.aggregate(ps, 0.05, kernel="w2", laplacian=True)
    b = ab.aggregate(ps, 0.05, kernel="w2", laplacian=True, fixedAdjacency=False)
    for key in a.out:
        assert torch.equal(a.out[key], b.out[key]), key
    adj = sc.adjacency(ps, None, 0.05)
    assert adj.stats["candidates"][1] < n
    axes = torch.tensor(g.normal(size=(n, 2)), device=DEV)
    assert torch.equal(a.cone_area(axes, math.pi / 5), b.cone_area(axes, math.pi / 5))
    vecs = torch.tensor(g.normal(size=(2, n, 2)), device=DEV)
    assert torch.equal(a.dir_extreme(vecs), b.dir_extreme(vecs))
    # the precompute format over the broadphase rows: per-query sums equal the fused channels
    pm = sc.precompute(adj, "w2", form="fp64")
    for bi, ent in enumerate(pm.entries):
        tot = torch.zeros((n,), dtype=torch.float64, device=DEV)
        for rm in ent["surface"]:
            tot.index_add_(0, rm.q, rm.lam)
        assert torch.equal(tot, a.channels[0, bi, :, 0]), bi


def test_server_cancel_pending_and_in_flight():
    from warpsphboundaries_amd.serve import BoundaryServer
    scenes = WL.config5(n_scenes=4, seed=40, n_min=2000, n_max=5000, e_min=16, e_max=256, frac3d=0.0)
    srv = BoundaryServer(streams=1)
    for s in scenes:
        srv.add_scene(s.name, s.loops, s.h)
    reqs = [srv.submit(s.name, s.positions, s.h, s.kernel) for s in scenes for _ in range(2)]
    srv.cancel(reqs[-1])                                  # pending: never launched
    assert reqs[-1] not in srv.pending
    srv.poll()                                            # launches the first batch
    launched = [r for r in reqs if r not in srv.pending and not r.cancelled]
    assert launched
    srv.cancel(launched[0])                               # in flight: its result is dropped
    srv.run()
    assert reqs[-1].result is None and reqs[-1].done is None
    assert launched[0].result is None and launched[0].done is not None
    alone = BoundaryServer(streams=1)
    for s in scenes:
        alone.add_scene(s.name, s.loops, s.h)
    for r in reqs:
        if r.cancelled:
            continue
        s = next(x for x in scenes if x.name == r.scene)
        ref = alone.submit(s.name, s.positions, s.h, s.kernel)
        alone.run()
        assert np.array_equal(r.result, ref.result), r.scene
