"""Parity of the HIP 2-D path against the CPU oracle (adjacency bit-exact; moments 1e-12 for the FP64 forms,
1e-6 signal-relative for the mixed form) and against the reference's golden fixtures."""
import numpy as np
import pytest
import torch

from oracle import adjacency as oadj
from oracle import edge2d
from tests.golden_data import ELEMENTS, KERNELS, rel_units_h, twelve
from warpsphboundaries_amd import workloads as WL
from warpsphboundaries_amd.device import count_order, edge_channels12
from warpsphboundaries_amd.scene import SurfaceRep

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FORMS = ["plain", "fp64", "mixed"]


def _rep(loop, solid):
    return SurfaceRep.polygon(loop, solid)


def _geometries():
    g = WL.rng(11)
    star = WL.star_polygon(256, g=g)
    notch = np.array([[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.52, 0.9], [0.5, 0.3], [0.48, 0.9], [0.1, 0.9]])
    return {
        "box-outside": (np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]), "outside"),
        "64gon": (WL.regular_polygon((0.5, 0.5), 0.15, 64), "inside"),
        "star256": (star, "inside"),
        "notch": (notch, "inside"),
    }


def _particles(loop, h, n, seed):
    """uniform points plus adversarial ones: on edges, at vertices, on cell and bin boundaries, near endpoints."""
    g = WL.rng(seed)
    lo, hi = loop.min(0) - 2 * h, loop.max(0) + 2 * h
    pts = [g.uniform(lo, hi, (n, 2))]
    p, q = loop, np.roll(loop, -1, 0)
    k = g.integers(0, len(loop), n // 4)
    t = g.uniform(0, 1, (n // 4, 1))
    pts.append(p[k] + t * (q[k] - p[k]))                                   # on the edges (up to rounding)
    pts.append(loop.copy())                                                # exactly at the vertices
    pts.append(p[k] + g.normal(0, 1e-9, (n // 4, 2)))                     # near the endpoints
    mid = 0.5 * (p + q)
    pts.append(mid)                                                        # edge midpoints
    return np.concatenate(pts)


def _oracle_rows(st, kname, lpos, lsup, e, cnt, ind, plane):
    K = st.K
    rows = np.repeat(np.arange(len(lpos)), cnt)
    slots = np.concatenate([e[r * K:r * K + c] for r, c in enumerate(cnt)]) if len(rows) else np.zeros(0, int)
    P = (st.verts[st.edges[slots, 0]] - lpos[rows]) / lsup[rows, None]
    Q = (st.verts[st.edges[slots, 1]] - lpos[rows]) / lsup[rows, None]
    c = edge2d.edge_channels12(P, Q, kname) if len(rows) else np.zeros((0, 12))
    out = np.zeros((len(lpos), 12))
    np.add.at(out, rows, c)
    return out + ind[:, None] * plane[None]


@pytest.mark.parametrize("name", ["box-outside", "64gon", "star256", "notch"])
def test_adjacency_bit_exact(name):
    loop, solid = _geometries()[name]
    rep = _rep(loop, solid)
    h = 0.02 if name != "star256" else 0.01
    pts = _particles(rep.vertices.numpy(), h, 3000, 5)
    g = WL.rng(6)
    sup = np.where(g.uniform(0, 1, len(pts)) < 0.3, 0.6 * h, h)
    smax = h
    # put a few particles exactly on cell corners of the static grid
    st = oadj.Static(rep.vertices.numpy(), rep.edges.numpy(), rep.background, smax)
    corners = st.cl.lo + st.cl.cell * g.integers(0, min(st.cl.dims), (200, 2)).astype(np.float64)
    pts = np.concatenate([pts, corners])
    sup = np.concatenate([sup, np.full(len(corners), h)])
    geo = rep.geom(smax, DEV)
    assert geo.K == st.K
    lp = torch.tensor(pts, dtype=torch.float64, device=DEV)
    ls = torch.tensor(sup, dtype=torch.float64, device=DEV)
    valid = torch.tensor(g.uniform(0, 1, len(pts)) < 0.95, device=DEV)
    slots, counts, ind = geo.adjacency(lp, ls, valid=valid, pad=True)
    e_ref, c_ref, i_ref = oadj.fixed_topology(st, pts, sup, valid.cpu().numpy())
    print(name, "rows", len(pts), "K", st.K, "pairs", int(c_ref.sum()))
    assert np.array_equal(slots.cpu().numpy(), e_ref)
    assert np.array_equal(counts.cpu().numpy(), c_ref)
    # rows served in cell order give the same slots and indicator
    order = geo.sort(lp)
    assert sorted(order.cpu().numpy().tolist()) == list(range(len(pts)))
    # grouped by cell key in ascending order, the rows outside the grid last
    nx, ny = st.cl.dims
    fx = np.floor((pts[:, 0] - st.cl.lo[0]) / st.cl.cell)
    fy = np.floor((pts[:, 1] - st.cl.lo[1]) / st.cl.cell)
    inside = (fx >= 0) & (fx < nx) & (fy >= 0) & (fy < ny)
    key = np.where(inside, fx * ny + fy, nx * ny)
    assert np.array_equal(key[order.cpu().numpy()], np.sort(key))
    s2, c2, i2 = geo.adjacency(lp, ls, valid=valid, pad=True, order=order)
    assert torch.equal(s2, slots) and torch.equal(c2, counts) and torch.equal(i2, ind)
    got = ind.cpu().numpy()
    # the boundary angle sums go through atan2 (device vs host libm): equal to rounding, every other row bit-exact
    frac = (i_ref != np.round(i_ref))
    assert np.array_equal(got[~frac], i_ref[~frac])
    np.testing.assert_allclose(got[frac], i_ref[frac], rtol=0, atol=1e-14)


@pytest.mark.parametrize("case", ["star-two-pass", "box-fine-grid-one-pass"])
def test_cell_sort_groups_rows_by_cell_at_scale(case):
    """wsb_sort2d over 300,000 rows (73 chunks of the two-pass sort), a large share outside the grid: a permutation of
    the rows whose cell keys ascend, the rows outside the grid last.  The 4-edge box at support 4e-4 has a grid of
    2,502 x 2,502 cells, too fine for the two-pass sort's coarse bins: the one-pass sort serves it."""
    if case.startswith("star"):
        loop, solid, smax = WL.star_polygon(512, g=WL.rng(21)), "inside", 0.004
    else:
        loop, solid, smax = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]]), "outside", 4e-4
    rep = _rep(loop, solid)
    pts = WL.rng(22).uniform(-0.1, 1.1, (300000, 2))
    pts[:64] = pts[64]                                    # one cell crowded by a whole wavefront of rows
    cl = oadj.edge_cell_list(rep.vertices.numpy(), rep.edges.numpy().astype(np.int64), smax)
    nx, ny = cl.dims
    geo = rep.geom(smax, DEV)
    assert (geo.info["nx"], geo.info["ny"]) == (nx, ny)
    order = geo.sort(torch.tensor(pts, dtype=torch.float64, device=DEV)).cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(len(pts)))
    fx = np.floor((pts[:, 0] - cl.lo[0]) / cl.cell)
    fy = np.floor((pts[:, 1] - cl.lo[1]) / cl.cell)
    inside = (fx >= 0) & (fx < nx) & (fy >= 0) & (fy < ny)
    key = np.where(inside, fx * ny + fy, nx * ny)
    print(case, "cells", nx * ny, "rows outside the grid", int((~inside).sum()))
    assert 0 < (~inside).sum() < len(pts)
    assert np.array_equal(key[order], np.sort(key))


@pytest.mark.parametrize("name", ["star256", "notch"])
def test_sorted_layout_equals_row_layout(name):
    """wsb_adjacency2d_sorted / wsb_moments2d_multi_sorted: position i holds row order[i]'s slots, count and (x, y, h,
    indicator), bitwise those of wsb_adjacency2d (with valid rows masked, padded and not), and the fused moments of
    both forms, written at the rows, bitwise those of wsb_moments2d_multi; the slots also against the oracle."""
    loop, solid = _geometries()[name]
    rep = _rep(loop, solid)
    h = 0.01
    pts = _particles(rep.vertices.numpy(), h, 3000, 31)
    g = WL.rng(32)
    sup = np.where(g.uniform(0, 1, len(pts)) < 0.3, 0.7 * h, h)
    st = oadj.Static(rep.vertices.numpy(), rep.edges.numpy(), rep.background, h)
    geo = rep.geom(h, DEV)
    K = geo.K
    lp = torch.tensor(pts, dtype=torch.float64, device=DEV)
    ls = torch.tensor(sup, dtype=torch.float64, device=DEV)
    valid = torch.tensor(g.uniform(0, 1, len(pts)) < 0.9, device=DEV)
    e_ref, c_ref, i_ref = oadj.fixed_topology(st, pts, sup, valid.cpu().numpy())
    slots, counts, ind = geo.adjacency(lp, ls, valid=valid, pad=True)
    order = geo.sort(lp)
    o = order.long()
    for pad in (False, True):
        sp, cp, prow = geo.adjacency_sorted(lp, ls, order, valid=valid, pad=pad)
        assert torch.equal(cp, counts[o])
        assert torch.equal(prow[:, 3], ind[o]) and torch.equal(prow[:, :2], lp[o]) and torch.equal(prow[:, 2], ls[o])
        got = sp.view(-1, K)
        want = slots.view(-1, K)[o]
        if pad:
            assert torch.equal(got, want)
        else:
            live = torch.arange(K, device=DEV)[None, :] < cp[:, None]
            assert torch.equal(torch.where(live, got, 0), torch.where(live, want, 0))
    inv = np.empty(len(pts), dtype=np.int64)
    inv[o.cpu().numpy()] = np.arange(len(pts))
    assert np.array_equal(sp.view(-1, K).cpu().numpy()[inv].reshape(-1), e_ref)
    kset = ("cubic", "w2", "w4", "w6")
    for form in ("fp64", "mixed"):
        a = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, order=order)
        b = geo.moments_multi_sorted(kset, order, sp, cp, prow, form=form)
        for k in kset:
            assert torch.equal(a[k], b[k]), (form, k)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("kname", KERNELS)
def test_moments_match_oracle(form, kname):
    loop, solid = _geometries()["star256"]
    rep = _rep(loop, solid)
    h = 0.01
    pts = _particles(rep.vertices.numpy(), h, 2000, 9)
    sup = np.full(len(pts), h)
    st = oadj.Static(rep.vertices.numpy(), rep.edges.numpy(), rep.background, h)
    e_ref, c_ref, i_ref = oadj.fixed_topology(st, pts, sup)
    ref = _oracle_rows(st, kname, pts, sup, e_ref, c_ref, i_ref, edge2d.plane_vector(kname))
    geo = rep.geom(h, DEV)
    lp = torch.tensor(pts, dtype=torch.float64, device=DEV)
    ls = torch.tensor(sup, dtype=torch.float64, device=DEV)
    slots, counts, ind = geo.adjacency(lp, ls)
    out = geo.moments(kname, slots, counts, ind, lp, ls, form=form).cpu().numpy()
    err = np.abs(out - ref)
    sig = err.max(0) / np.maximum(np.abs(ref).max(0), 1e-300)
    print(form, kname, "max abs", err.max(), "signal-relative per channel", sig.max())
    if form == "mixed":
        assert sig.max() <= 1e-6
    else:
        # the monomial plan of w6 carries coefficients up to 1.3e4: the reference's own FP64 gate for it is 5e-12
        assert err.max() <= (5e-12 if kname == "w6" else 1e-12)
    # relaunch is bitwise identical, and so is the launch over rows grouped by cell
    again = geo.moments(kname, slots, counts, ind, lp, ls, form=form).cpu().numpy()
    assert np.array_equal(out, again)
    order = geo.sort(lp)
    grouped = geo.moments(kname, slots, counts, ind, lp, ls, form=form, order=order).cpu().numpy()
    assert np.array_equal(out, grouped)


@pytest.mark.parametrize("form", FORMS)
def test_pair_channels_match_oracle_on_degenerate_edges(form):
    g = WL.rng(3)
    n = 3000
    p = g.uniform(-1.2, 1.2, (n, 2))
    q = p + g.normal(0, 0.5, (n, 2))
    q[:50] = p[:50] * g.uniform(-2, -0.1, (50, 1))
    p[50:100] = g.normal(0, 1e-9, (50, 2))
    q[100:150] = p[100:150] + g.normal(0, 1e-6, (50, 2))
    verts = np.concatenate([p, q])
    edges = np.stack([np.arange(n), np.arange(n) + n], 1)
    pos = np.zeros((1, 2))
    sup = np.ones(1)
    for kname in KERNELS:
        ref = edge2d.edge_channels12(p, q, kname)
        t = lambda a, dt: torch.tensor(a, dtype=dt, device=DEV)
        got = edge_channels12(t(np.zeros(n), torch.int32), t(np.arange(n), torch.int32), t(pos, torch.float64),
                              t(sup, torch.float64), t(verts, torch.float64), t(edges, torch.int32), kname, form).cpu().numpy()
        err = np.abs(got - ref)
        tol = 1e-6 * np.abs(ref).max() if form == "mixed" else (5e-12 if kname == "w6" else 1e-12)
        assert err.max() <= tol, (kname, form, err.max())


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("kname", KERNELS)
def test_golden_fixture_elements_on_device(form, kname):
    """every golden element: device edge channels summed over the polygon + indicator x plane vector."""
    verts, edges, owner, inds, refs = [], [], [], [], []
    base = 0
    for i, rec in enumerate(ELEMENTS):
        rel = rel_units_h(rec["polygon"], rec["x"], rec["h"])
        n = len(rel)
        verts.append(rel)
        edges.append(np.stack([np.arange(n), (np.arange(n) + 1) % n], 1) + base)
        owner.append(np.full(n, i))
        inds.append(oadj.indicator_points(rel, np.stack([np.arange(n), (np.arange(n) + 1) % n], 1), 0, np.zeros((1, 2)),
                                          np.ones(1))[0])
        refs.append(twelve(rec, kname))
        base += n
    V = np.concatenate(verts)
    E = np.concatenate(edges)
    t = lambda a, dt: torch.tensor(a, dtype=dt, device=DEV)
    P = len(E)
    c = edge_channels12(t(np.zeros(P), torch.int32), t(np.arange(P), torch.int32), t(np.zeros((1, 2)), torch.float64),
                        t(np.ones(1), torch.float64), t(V, torch.float64), t(E, torch.int32), kname, form).cpu().numpy()
    tot = np.zeros((len(ELEMENTS), 12))
    np.add.at(tot, np.concatenate(owner), c)
    tot += np.array(inds)[:, None] * edge2d.plane_vector(kname)[None]
    ref = np.array(refs)
    err = np.abs(tot - ref)
    if form == "mixed":
        assert (err.max(0) / np.abs(ref).max(0)).max() <= 1e-6
    else:
        tol = np.array([1e-12] * 3 + [1e-11] * 6 + [1e-12] * 3) * (5 if kname == "w6" else 1)
        assert (err <= tol).all(), err.max(0)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("kset", [("cubic", "w2", "w4", "w6"), ("w2",), ("cubic", "w6"), ("w4", "cubic", "w2")])
def test_moments_multi_matches_single_launches_and_oracle(form, kset):
    """wsb_moments2d_multi (several kernels in one pass sharing each pair's chord primitives) against the oracle
    (the single-launch gates) and against one launch per kernel (the same closed form, summed in the same slot
    order; within 1e-13 of the signal, 1e-12 for w6, whose oracle gate is five times the others'); rows in cell order
    and in input order bitwise equal, in count order too for the mixed form (the FP64 form's count order runs another
    kernel: within the multi-vs-single gate); WSB_OUT_ACCUMULATE doubles; physical units and a world rotation as the single
    launch."""
    from warpsphboundaries_amd import _lib as L
    loop, solid = _geometries()["star256"]
    rep = _rep(loop, solid)
    h = 0.01
    pts = _particles(rep.vertices.numpy(), h, 2000, 13)
    sup = np.where(WL.rng(14).uniform(0, 1, len(pts)) < 0.3, 0.7 * h, h)
    st = oadj.Static(rep.vertices.numpy(), rep.edges.numpy(), rep.background, h)
    e_ref, c_ref, i_ref = oadj.fixed_topology(st, pts, sup)
    geo = rep.geom(h, DEV)
    lp = torch.tensor(pts, dtype=torch.float64, device=DEV)
    ls = torch.tensor(sup, dtype=torch.float64, device=DEV)
    slots, counts, ind = geo.adjacency(lp, ls)
    multi = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form)
    order = geo.sort(lp)
    grouped = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, order=order)
    corder, nwork = geo.count_order(counts, base=order)
    by_count = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, order=corder, nwork=nwork)
    acc_count = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, order=corder, nwork=nwork,
                                  flags=L.OUT_ACCUMULATE)
    acc_count = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, order=corder, nwork=nwork,
                                  flags=L.OUT_ACCUMULATE, outs=acc_count)
    acc = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, flags=L.OUT_ACCUMULATE)
    acc = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, flags=L.OUT_ACCUMULATE, outs=acc)
    a = 0.83
    rot4 = torch.tensor([np.cos(a), -np.sin(a), np.sin(a), np.cos(a)], dtype=torch.float64, device=DEV)
    fl = L.OUT_PHYSICAL | L.OUT_WORLD
    world = geo.moments_multi(kset, slots, counts, ind, lp, ls, form=form, rot4=rot4, flags=fl)
    assert set(multi) == set(kset)
    for k in kset:
        got = multi[k].cpu().numpy()
        ref = _oracle_rows(st, k, pts, sup, e_ref, c_ref, i_ref, edge2d.plane_vector(k))
        err = np.abs(got - ref)
        sig = err.max(0) / np.maximum(np.abs(ref).max(0), 1e-300)
        if form == "mixed":
            assert sig.max() <= 1e-6, (k, sig.max())
        else:
            assert err.max() <= (5e-12 if k == "w6" else 1e-12), (k, err.max())
        single = geo.moments(k, slots, counts, ind, lp, ls, form=form).cpu().numpy()
        scale = np.maximum(np.abs(single).max(0), 1.0)
        d = (np.abs(got - single) / scale).max()
        print(form, kset, k, "oracle", err.max(), "multi vs single", d)
        # which product of a difference the compiler fuses is its own choice; w6's coefficients (to 1.3e4) amplify
        # that rounding tenfold over the other kernels (its oracle gate is 5e-12 against 1e-12)
        assert d <= (1e-12 if k == "w6" else 1e-13), (k, d)
        assert torch.equal(grouped[k], multi[k])
        if form == "fp64":
            # cell and input order run the FP64 form through the workgroup tile kernel, count order through the
            # row-per-thread kernel: the same slot-order sums, each with the contractions its compiler chose
            dc = ((by_count[k] - multi[k]).abs().amax(0) / multi[k].abs().amax(0).clamp(min=1.0)).max().item()
            assert dc <= (1e-12 if k == "w6" else 1e-13), (k, dc)
        else:
            assert torch.equal(by_count[k], multi[k])
        assert torch.equal(acc_count[k], 2.0 * by_count[k])
        assert torch.equal(acc[k], 2.0 * multi[k])
        sw = geo.moments(k, slots, counts, ind, lp, ls, form=form, rot4=rot4, flags=fl)
        dw = ((world[k] - sw).abs().amax(0) / sw.abs().amax(0).clamp(min=1.0)).max().item()
        assert dw <= (1e-12 if k == "w6" else 1e-13), (k, dw)


def test_moments_multi_rejects_what_it_does_not_fuse():
    loop, solid = _geometries()["64gon"]
    rep = _rep(loop, solid)
    geo = rep.geom(0.02, DEV)
    lp = torch.tensor([[0.5, 0.36]], dtype=torch.float64, device=DEV)
    ls = torch.full((1,), 0.02, dtype=torch.float64, device=DEV)
    slots, counts, ind = geo.adjacency(lp, ls)
    with pytest.raises(NotImplementedError):
        geo.moments_multi(("w2", "lw2"), slots, counts, ind, lp, ls)
    with pytest.raises(NotImplementedError):
        geo.moments_multi(("w2",), slots, counts, ind, lp, ls, form="plain")


def _line_points(loop, h, n, seed):
    """points on the lines of the edges beyond their ends, at the vertices' heights, and within 1e-12 of vertices:
    where a ray test or a segment-crossing test is closest to its rounding bound."""
    g = WL.rng(seed)
    lo, hi = loop.min(0) - 2 * h, loop.max(0) + 2 * h
    p, q = loop, np.roll(loop, -1, 0)
    k = g.integers(0, len(loop), n)
    ext = p[k] + (1.0 + g.uniform(0.0, 3.0, (n, 1))) * (q[k] - p[k])
    level = np.stack([g.uniform(lo[0], hi[0], n), p[k][:, 1]], 1)
    near = p[k] + g.normal(0.0, 1e-12, (n, 2))
    return np.concatenate([ext, level, near, g.uniform(lo, hi, (n, 2))])


@pytest.mark.parametrize("name", ["star256", "notch", "box-outside", "nearly-horizontal"])
def test_local_winding_indicator_bit_exact(name):
    """the indicator of rows in cells with edges is the cell reference point's plus the crossings of the segment to
    the row (geom2d_device.h local_winding_edge), with the reference's ray where a decision is uncertain: equal to
    the oracle's ray cast on points placed on edge lines, at vertex heights and next to vertices.  A geometry with
    an edge 1e-15 off horizontal takes the ray for every row (info local_winding 0)."""
    if name == "nearly-horizontal":
        loop, solid = np.array([[0.1, 0.2], [0.9, 0.2 + 1e-15], [0.7, 0.8], [0.2, 0.7]]), "inside"
    else:
        loop, solid = _geometries()[name]
    rep = _rep(loop, solid)
    h = 0.02
    pts = np.concatenate([_line_points(rep.vertices.numpy(), h, 4000, 21), _particles(rep.vertices.numpy(), h, 2000, 22)])
    sup = np.full(len(pts), h)
    st = oadj.Static(rep.vertices.numpy(), rep.edges.numpy(), rep.background, h)
    geo = rep.geom(h, DEV)
    assert geo.info["local_winding"] == (0 if name == "nearly-horizontal" else 1)
    assert geo.info["cells_without_reference"] == 0
    lp = torch.tensor(pts, dtype=torch.float64, device=DEV)
    ls = torch.tensor(sup, dtype=torch.float64, device=DEV)
    slots, counts, ind = geo.adjacency(lp, ls, pad=True)
    e_ref, c_ref, i_ref = oadj.fixed_topology(st, pts, sup)
    assert np.array_equal(slots.cpu().numpy(), e_ref) and np.array_equal(counts.cpu().numpy(), c_ref)
    got = ind.cpu().numpy()
    frac = i_ref != np.round(i_ref)
    print(name, "rows", len(pts), "boundary rows", int(frac.sum()))
    assert np.array_equal(got[~frac], i_ref[~frac])
    np.testing.assert_allclose(got[frac], i_ref[frac], rtol=0, atol=1e-14)


@pytest.mark.parametrize("with_base", [False, True])
def test_count_order_groups_rows_by_slot_count(with_base):
    """wsb_count_order2d: a permutation of the base order (of 0..N-1 without one) along which the slot counts never
    increase (rows with 1023 or more slots share the last bin), and nwork = the number of rows with a slot.  Checked on
    the star workload's counts, a synthetic count vector with every count up to 1,500, sizes that are not multiples of
    a workgroup and a vector with no slot at all."""
    loop, solid = _geometries()["star256"]
    rep = _rep(loop, solid)
    h = 0.01
    pts = _particles(rep.vertices.numpy(), h, 20000, 21)
    geo = rep.geom(h, DEV)
    lp = torch.tensor(pts, dtype=torch.float64, device=DEV)
    ls = torch.full((len(pts),), h, dtype=torch.float64, device=DEV)
    _, counts, _ = geo.adjacency(lp, ls)
    g = WL.rng(22)
    synth = torch.tensor(g.integers(0, 1500, 70000), dtype=torch.int32, device=DEV)
    small = torch.tensor(np.where(g.uniform(0, 1, 1001) < 0.2, g.integers(1, 9, 1001), 0), dtype=torch.int32, device=DEV)
    for cnt, K in ((counts, geo.K), (synth, 4000), (small, 8), (small[:63], 8), (torch.zeros(200, dtype=torch.int32,
                                                                                                   device=DEV), 8)):
        n = cnt.numel()
        base = torch.tensor(g.permutation(n), dtype=torch.int32, device=DEV) if with_base else None
        order, nwork = count_order(cnt, K, base=base)
        o = order.cpu().numpy()
        want = (base.cpu().numpy() if base is not None else np.arange(n))
        assert np.array_equal(np.sort(o), np.sort(want))
        c = np.minimum(np.minimum(cnt.cpu().numpy(), K), 1023)[o]
        assert (np.diff(c) <= 0).all()
        assert int(nwork.item()) == int((c > 0).sum())
