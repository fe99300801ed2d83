"""Oracle parity on the workloads of the benchmark configurations themselves, at sizes the oracle finishes in seconds.

* box + 64-gon (two surface representations of one body, 16,384 particles, w4, h = 0.02), in full: slots, counts and
  indicator of both representations bit-exact with the oracle, the twelve channels of all three forms, and the signed
  distance with its normal against the reference's rule (nearest edge, sign from the winding number at supportMax,
  normal = the edge's fluid normal, or the direction from the closest point outside the solid);
* the 4,096-edge star at h = 0.004 on a stratified 65,536-particle subsample (the 30 % near-wall fraction kept): the
  adjacency bit-exact and the twelve channels of cubic, w2, w4, w6 in every form (the fused multi-kernel pass in cell
  order and in count order);
* the level-6 icosphere cavity in the unit box (81,932 faces, w4, h = 0.01) on 4,096 of its particles (2,048 of them
  within 1.5 h of the sphere, 2,048 farther): face lists bit-exact with a brute-force scan, winding number 1, the 22 stored moments of
  both forms against the FP64 oracle; the solid-sphere closed form on all 65,536 (R, d) pairs of the configuration
  against the exact rational evaluation;
* the served mix at its stated ranges (up to 2e5 particles, polygons of up to 8,192 edges, meshes of up to 80k
  triangles): a row subsample of every scene's served result against the oracle (the signal of a channel: its largest
  value over the scene).

Gates (as the other parity tests): FP64 forms 1e-12 in units of h (5e-12 for w6, the reference's own FP64 gate for its
monomial plan), mixed form 1e-6 signal-relative per channel; 3-D 1e-12 (1 + |v|) per channel in units of h.
"""
import numpy as np
import pytest
import torch

from oracle import adjacency as oadj
from oracle import arith as A
from oracle import edge2d
from oracle import mesh3d as OM
from tests import config_oracles as CO
from warpsphboundaries_amd import workloads as WL
from warpsphboundaries_amd.fused import FusedGroup, FusedWall
from warpsphboundaries_amd.scene import Body, ParticleState, Scene, SurfaceRep, fixed_adjacency

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FORMS = ["plain", "fp64", "mixed"]
UNIT_EXP = np.array([0, 1, 1, -1, -1, 0, 0, 0, 0, 2, 2, 2])


def _oracle_rows(rep, h_max, pts, sup, kname, topo=None):
    """(e, counts, ind, channels [N,12] in units of h) of one surface representation (``topo``: a previous result's
    (static, e, counts, ind) to reuse)."""
    if topo is None:
        st = oadj.Static(rep.vertices.cpu().numpy(), rep.edges.cpu().numpy(), rep.background, h_max)
        e, c, ind = oadj.fixed_topology(st, pts, sup)
    else:
        st, e, c, ind = topo
    K = st.K
    rows = np.repeat(np.arange(len(pts)), c)
    sl = e.reshape(-1, K)[np.arange(K)[None, :] < c[:, None]] if len(pts) else np.zeros(0, np.int64)
    out = np.zeros((len(pts), 12))
    if len(rows):
        P = (st.verts[st.edges[sl, 0]] - pts[rows]) / sup[rows, None]
        Q = (st.verts[st.edges[sl, 1]] - pts[rows]) / sup[rows, None]
        # the edge channels in x86 long double (oracle.arith), so that the reference carries no FP64 rounding of its own
        with A.precision("ld"):
            ch = A.to_float(edge2d.edge_channels12(A.asarray(P), A.asarray(Q), kname))
        np.add.at(out, rows, ch)
    return e, c, ind, out + ind[:, None] * edge2d.plane_vector(kname)[None]


def _gate(got, ref, form, kname, what):
    err = np.abs(got - ref)
    sig = err.max(0) / np.maximum(np.abs(ref).max(0), 1e-300)
    if form == "mixed":
        print(what, form, kname, "signal-relative", sig.max())
        assert sig.max() <= 1e-6, (what, kname, sig)
    else:
        print(what, form, kname, "max abs (units of h)", err.max(), "signal-relative", sig.max())
        assert err.max() <= (5e-12 if kname == "w6" else 1e-12), (what, kname, err.max(0))


def _check_indicator(got, ref):
    # the boundary angle sums go through atan2 (device vs host libm): equal to rounding, every other row bit-exact
    frac = ref != np.round(ref)
    assert np.array_equal(got[~frac], ref[~frac])
    np.testing.assert_allclose(got[frac], ref[frac], rtol=0, atol=1e-14)


def _signed_reference(reps, pts, smax):
    """the reference's rule (scene/scene.py:1320-1368) restated on the host: per representation the nearest edge
    (lowest id among ties), d < 0 where the winding indicator at supportMax exceeds 1/2, normal = the edge's right
    normal, or (x - cp)/|x - cp| outside the solid when that distance exceeds 1e-12; the smallest d over the
    representations, the first one on a tie."""
    best = np.full(len(pts), np.inf)
    nrm = np.zeros((len(pts), 2))
    for rep in reps:
        V, E = rep.vertices.cpu().numpy(), rep.edges.cpu().numpy().astype(np.int64)
        a, b = V[E[:, 0]], V[E[:, 1]]
        e = b - a
        t = np.clip(((pts[:, None] - a[None]) * e[None]).sum(2) / (e * e).sum(1)[None], 0, 1)
        cp_all = a[None] + t[..., None] * e[None]
        dist = np.linalg.norm(pts[:, None] - cp_all, axis=2)
        k = dist.argmin(1)
        dmin = dist[np.arange(len(pts)), k]
        cp = cp_all[np.arange(len(pts)), k]
        st = oadj.Static(V, E, rep.background, smax)
        ins = oadj.fixed_topology(st, pts, np.full(len(pts), smax))[2] > 0.5
        d = np.where(ins, -dmin, dmin)
        ek = e[k] / np.linalg.norm(e[k], axis=1)[:, None]
        n = np.stack([ek[:, 1], -ek[:, 0]], 1)
        dirv = pts - cp
        dn = np.linalg.norm(dirv, axis=1)
        out = (dn > 1e-12) & ~ins
        n[out] = dirv[out] / dn[out, None]
        better = d < best
        best = np.where(better, d, best)
        nrm[better] = n[better]
    return best, nrm


# ------------------------------------------------------------------------------------------------ box + 64-gon
@pytest.mark.parametrize("form", FORMS)
def test_config1_box_and_64gon_scene_in_full(form):
    w = WL.config1()
    reps = [SurfaceRep.polygon(loop, solid) for loop, solid in w.loops]
    sc = Scene([Body(bodyId=0, reps=reps)], DEV)
    ps = ParticleState(torch.tensor(w.positions, device=DEV), torch.tensor(w.supports, device=DEV))
    adj = fixed_adjacency(sc, ps, None, w.h)
    ref = np.zeros((len(w.positions), 12))
    for rep, topo in zip(reps, adj.bodies[0].reps):
        e, c, ind, rows = _oracle_rows(rep, w.h, w.positions, w.supports, "w4")
        assert np.array_equal(topo.e.cpu().numpy(), e)
        assert np.array_equal(topo.counts.cpu().numpy(), c)
        _check_indicator(topo.ind.cpu().numpy(), ind)
        ref += rows
    wall = FusedWall(sc, adj, (FusedGroup("w4", None, form),))
    got = wall.channels[0, 0].cpu().numpy() / w.supports[:, None] ** UNIT_EXP[None]
    _gate(got, ref, form, "w4", "box+64gon")
    if form == "fp64":
        d, n, hit = sc.signed_distance(ps.positions, supportMax=w.h)
        d_ref, n_ref = _signed_reference(reps, w.positions, w.h)
        assert hit.all()
        np.testing.assert_allclose(d.cpu().numpy(), d_ref, rtol=0, atol=1e-14)
        np.testing.assert_allclose(n.cpu().numpy(), n_ref, rtol=0, atol=1e-12)


# ------------------------------------------------------------------------------------------------ star, 4,096 edges
@pytest.fixture(scope="module")
def star_subsample():
    w = WL.subsample(WL.config2(), 65536, seed=9)
    rep = SurfaceRep.polygon(*w.loops[0])
    st = oadj.Static(rep.vertices.numpy(), rep.edges.numpy(), rep.background, w.h)
    topo = (st,) + oadj.fixed_topology(st, w.positions, w.supports)
    oracle = {k: _oracle_rows(rep, w.h, w.positions, w.supports, k, topo) for k in w.kernels}
    return w, rep, oracle


@pytest.mark.parametrize("form", FORMS)
def test_config2_star_subsample_all_kernels(star_subsample, form):
    w, rep, oracle = star_subsample
    geo = rep.geom(w.h, DEV)
    lp = torch.tensor(w.positions, device=DEV)
    ls = torch.tensor(w.supports, device=DEV)
    order = geo.sort(lp)
    slots, counts, ind = geo.adjacency(lp, ls, pad=True, order=order)
    e, c, i, _ = oracle["w4"]
    assert np.array_equal(slots.cpu().numpy(), e) and np.array_equal(counts.cpu().numpy(), c)
    _check_indicator(ind.cpu().numpy(), i)
    assert int(c.sum()) > 100000 and (c[w.near] > 0).all()
    if form == "plain":
        got = {k: geo.moments(k, slots, counts, ind, lp, ls, form=form).cpu().numpy() for k in w.kernels}
    else:
        got = {k: v.cpu().numpy() for k, v in geo.moments_multi(w.kernels, slots, counts, ind, lp, ls, form=form,
                                                                  order=order).items()}
        corder, nwork = geo.count_order(counts, base=order)
        by_count = geo.moments_multi(w.kernels, slots, counts, ind, lp, ls, form=form, order=corder, nwork=nwork)
        for k in w.kernels:
            if form == "mixed":     # the same row-per-thread kernel in both orders
                assert np.array_equal(by_count[k].cpu().numpy(), got[k])
            else:                   # the FP64 form's count order runs the row-per-thread kernel, cell order the tile
                _gate(by_count[k].cpu().numpy(), oracle[k][3], form, k, "star4096 subsample, count order")
    for k in w.kernels:
        _gate(got[k], oracle[k][3], form, k, "star4096 65,536-particle subsample")


def _nearest_edges(V, E, pts, chunk=1024):
    """(distance, edge) of the nearest edge of every point by a scan over all edges (float64, lowest id on a tie),
    and the distance to the second nearest edge."""
    a, b = V[E[:, 0]], V[E[:, 1]]
    e = b - a
    ee = (e * e).sum(1)
    dist = np.empty(len(pts))
    arg = np.empty(len(pts), dtype=np.int64)
    second = np.empty(len(pts))
    for s in range(0, len(pts), chunk):
        p = pts[s:s + chunk]
        t = np.clip(((p[:, None] - a[None]) * e[None]).sum(2) / ee[None], 0.0, 1.0)
        dd = np.linalg.norm(p[:, None] - (a[None] + t[..., None] * e[None]), axis=2)
        k = dd.argmin(1)
        dist[s:s + chunk] = dd[np.arange(len(p)), k]
        arg[s:s + chunk] = k
        dd[np.arange(len(p)), k] = np.inf
        second[s:s + chunk] = dd.min(1)
    return dist, arg, second


def test_config2_signed_distance_edge_tree(star_subsample):
    """signed distance on the 4,096-edge star at h = 0.004 for the 65,536-particle subsample (70 % of it farther than
    one support from every edge, where the nearest-edge candidate grid finds the nearest edge), 8,192 points over a
    box twice the domain (inside the solid, and outside the cell grid) and 2,048 points over a box seven times the
    domain (most of them outside the candidate grid: the edge tree): the distance within 1e-14 of a scan over all
    edges, the scan's edge wherever its nearest edge is unique by more than 1e-12, the sign from the even-odd rule,
    and the normal of the reference's rule."""
    w, rep, _ = star_subsample
    g = WL.rng(31)
    pts = np.concatenate([w.positions, g.uniform(-0.5, 1.5, (8192, 2)), g.uniform(-3.0, 4.0, (2048, 2))])
    geo = rep.geom(w.h, DEV)
    d, n, edge = geo.signed_distance(torch.tensor(pts, device=DEV), w.h)
    d, n, edge = d.cpu().numpy(), n.cpu().numpy(), edge.cpu().numpy()
    V, E = rep.vertices.numpy(), rep.edges.numpy().astype(np.int64)
    dist, arg, second = _nearest_edges(V, E, pts)
    far = dist > w.h
    assert far.sum() > 40000
    np.testing.assert_allclose(np.abs(d), dist, rtol=0, atol=1e-14)
    unique = second - dist > 1e-12           # the others are nearest to a vertex shared by two edges
    assert unique.sum() > 20000
    assert np.array_equal(edge[unique], arg[unique])
    # the device's edge is a nearest edge wherever the scan's is not unique
    a, b = V[E[edge, 0]], V[E[edge, 1]]
    ev = b - a
    t = np.clip(((pts - a) * ev).sum(1) / (ev * ev).sum(1), 0.0, 1.0)
    cp = a + t[:, None] * ev
    np.testing.assert_allclose(np.linalg.norm(pts - cp, axis=1), dist, rtol=0, atol=1e-14)
    clear = np.abs(dist) > 1e-9
    inside = WL.inside_polygon(pts, V)
    assert np.array_equal(d[clear] < 0, inside[clear])
    # the normal: the edge's right normal inside the solid, else the direction from the closest point
    el = ev / np.linalg.norm(ev, axis=1)[:, None]
    nref = np.stack([el[:, 1], -el[:, 0]], 1)
    dirv = pts - cp
    dn = np.linalg.norm(dirv, axis=1)
    out = (dn > 1e-12) & (d > 0)
    nref[out] = dirv[out] / dn[out, None]
    np.testing.assert_allclose(n, nref, rtol=0, atol=1e-12)


# ------------------------------------------------------------------------------------------------ icosphere cavity
def _config3_particles(n=4096, seed=12):
    w = WL.config3()
    g = WL.rng(seed)
    r = np.linalg.norm(w.positions - 0.5, axis=1)
    near = r < 0.25 + 1.5 * w.h
    idx = np.concatenate([g.choice(np.nonzero(near)[0], n // 2, replace=False),
                          g.choice(np.nonzero(~near)[0], n - n // 2, replace=False)])
    return w, w.positions[np.sort(idx)]


def test_config3_level6_cavity_subsample():
    from warpsphboundaries_amd import mesh3d as MM
    from warpsphboundaries_amd.plan3d import H_POWER3, whole_vector3
    w, P = _config3_particles()
    V, F = w.vertices, w.faces
    assert len(F) == 81932
    h = w.h
    H = np.full(len(P), h)
    m = MM.Mesh3D(V, F, h, device=DEV)
    Q, HH = torch.tensor(P, device=DEV), torch.tensor(H, device=DEV)
    adj = m.adjacency(Q, HH)
    _, faces, off = adj.numpy()
    cols, offs = OM.face_adjacency(V[F], P, H)
    assert np.array_equal(off, offs) and np.array_equal(faces, cols)
    counts = np.diff(offs)
    assert (counts > 0).sum() > 1000
    # every particle lies in the fluid (inside the box, outside the faceted sphere inscribed in the radius-0.25 ball)
    assert np.array_equal(adj.chi.cpu().numpy(), np.ones(len(P)))
    u = P[:, :1] * 0 + h ** np.array(H_POWER3, dtype=np.float64)[None]
    ref = np.broadcast_to(np.asarray(whole_vector3("w4"))[None] * u[:1], (len(P), 22)).copy()
    # the compact form over each particle's own face list plus the winding number times the whole-ball vector
    work = np.nonzero(counts > 0)[0]
    parts = np.array_split(work, 64)
    jobs = [(V, F, P[p], [cols[offs[r]:offs[r + 1]] for r in p], "w4", h, np.ones(len(p))) for p in parts if len(p)]
    for p, rows in zip([p for p in parts if len(p)], CO.pool_map(CO.compact_rows, jobs)):
        ref[p] = rows
    fp = m.moments(adj, Q, HH, "w4", "fp64").cpu().numpy()
    err = np.abs(fp - ref) / u / (1 + np.abs(ref) / u)
    print("config3 fp64 vs oracle", err.max())
    assert err.max() <= 1e-12, err.max(0)
    mix = m.moments(adj, Q, HH, "w4", "mixed").cpu().numpy()
    sig = np.abs(ref / u).max(axis=0) + 1e-300
    rel = (np.abs(mix - ref) / u / sig).max()
    print("config3 mixed signal-relative", rel)
    assert rel <= 1e-6


def test_config3_sphere_pairs_all_65536_exact():
    from warpsphboundaries_amd import sphere as S
    R, d = S.sample_pairs(65536, seed=3)
    got = S.closed_form(torch.tensor(R, device=DEV), torch.tensor(d, device=DEV), "w4").cpu().numpy()
    parts = np.array_split(np.arange(len(R)), 64)
    ref = np.concatenate(CO.pool_map(CO.sphere_exact, [("w4", R[p], d[p]) for p in parts]))
    err = np.abs(got - ref)
    print("sphere pairs max abs", err.max(), "max rel", (err / np.maximum(np.abs(ref), 1e-300)).max())
    assert (err <= 2.3e-16 * np.abs(ref) + 1e-26).all(), err.max()


# ------------------------------------------------------------------------------------------------ served mix
def test_config5_served_mix_at_stated_ranges_against_the_oracle():
    from warpsphboundaries_amd.plan3d import H_POWER3
    from warpsphboundaries_amd.serve import BoundaryServer
    # the served workload (256 scenes, its default ranges): its first 24 scenes and the extremes of the ranges -- the
    # polygon with the most edges, the 2-D and 3-D scenes with the most particles, the mesh with the most faces
    every = WL.config5(seed=5)
    two = [s for s in every if s.dim == 2]
    three = [s for s in every if s.dim == 3]
    extremes = {max(two, key=lambda s: sum(len(l[0]) for l in s.loops)).name, max(two, key=lambda s: len(s.positions)).name,
                max(three, key=lambda s: len(s.mesh[1])).name, max(three, key=lambda s: len(s.positions)).name}
    scenes = every[:24] + [s for s in every[24:] if s.name in extremes]
    assert max(sum(len(l[0]) for l in s.loops) for s in scenes if s.dim == 2) > 7000
    assert max(len(s.mesh[1]) for s in scenes if s.dim == 3) > 70000 and max(len(s.positions) for s in scenes) > 190000
    srv = BoundaryServer(streams=4, form="mixed")
    for s in scenes:
        srv.add_scene(s.name, s.loops, s.h, mesh=s.mesh)
    pos = {s.name: WL.jitter(s.positions, s.h, 7) for s in scenes}
    reqs = {s.name: srv.submit(s.name, pos[s.name], s.h, s.kernel) for s in scenes}
    srv.run()
    g = WL.rng(31)
    for s in scenes:
        r = reqs[s.name]
        assert r.error is None, (s.name, r.error)
        p = pos[s.name]
        if s.dim == 2:
            reps = [SurfaceRep.polygon(l, solid) for l, solid in s.loops]
            # half the subsample among the rows with a slot (the device's counts only choose which rows are checked)
            near = np.zeros(len(p), dtype=bool)
            for rep in reps:
                lp = torch.tensor(p, device=DEV)
                near |= rep.geom(s.h, DEV).adjacency(lp, torch.full((len(p),), s.h, dtype=torch.float64,
                                                                    device=DEV))[1].cpu().numpy() > 0
            idx = np.unique(np.concatenate([g.choice(np.nonzero(near)[0], min(256, near.sum()), replace=False),
                                            g.choice(len(p), 256, replace=False)]))
            sup = np.full(len(idx), s.h)
            ref = sum(_oracle_rows(rep, s.h, p[idx], sup, s.kernel)[3] for rep in reps)
            full = np.asarray(r.result) / s.h ** UNIT_EXP[None]
            got = full[idx]
            # the signal of the scene: the larger of the subsample's oracle values and the whole served result
            sig = np.maximum(np.abs(ref).max(0), np.abs(full).max(0)) + 1e-300
            rel = (np.abs(got - ref).max(0) / sig).max()
            print(s.name, "2-D mixed signal-relative", rel, "edges", sum(len(l) for l, _ in s.loops), "rows", len(p))
            assert rel <= 1e-6, (s.name, rel)
        else:
            V, F = s.mesh
            idx = g.choice(len(p), 128, replace=False)
            X = p[idx]
            cols, offs = OM.face_adjacency(V[F], X, np.full(len(X), s.h))
            wnd = OM.snapped_winding(V, F, X)
            u = s.h ** np.array(H_POWER3, dtype=np.float64)[None]
            parts = np.array_split(np.arange(len(X)), 16)
            jobs = [(V, F, X[p], [cols[offs[r]:offs[r + 1]] for r in p], s.kernel, s.h, wnd[p]) for p in parts]
            ref = np.concatenate(CO.pool_map(CO.compact_rows, jobs))
            full = np.asarray(r.result)
            got = full[idx]
            # the signal of the scene: the larger of the subsample's oracle values and the whole served result
            sig = np.maximum(np.abs(ref / u).max(axis=0), np.abs(full / u).max(axis=0)) + 1e-300
            rel = (np.abs(got - ref) / u / sig).max()
            print(s.name, "3-D mixed signal-relative", rel, "faces", len(F), "rows", len(p))
            assert rel <= 1e-6, (s.name, rel)
