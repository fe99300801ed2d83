"""The disk element on the device (reference scene/scene.py:808-888 DiskArrayRep, scene/fixedadj.py:199-246,
edge/warpdisk.py, and its tests/scene/test_disk_array.py):

(a) the slots equal the oracle's slot fill bit for bit (cell list over the centres, 3 x 3 cells, fixed order), the
    counts are the filled slots, rows left out by `valid` have none;
(b) the twelve channels of every row (units of h, body frame) equal the oracle's circle identity summed over the
    row's slots in slot order, every kernel family, to 1e-12 of the scale of each channel's vector or tensor;
(c) physical units and the world rotation of a rotated, translated bundle through FusedWall equal the oracle's
    channels scaled and rotated on the host (1e-12);
(d) a bundle of disks with radii 0.1 h ... 6 h equals the sum of the same disks as fine area-preserving polygons
    (edges <= h/16, one body each) in lam, G, Cov, m1, m2 and the aggregate layout (cover, wall Laplacian, tensile) to
    5e-6 of the scale (the polygons' own error), as the reference's own gate;
(e) a solid ImplicitRep disk is the one-disk DiskArrayRep (bitwise); signed distance and inside of the bundle are the
    exact ones; the adjacency and moments entry points replay bitwise from a captured HIP graph.
"""
import math

import numpy as np
import pytest
import torch

from oracle import disk2d as OD
from warpsphboundaries_amd import _lib as L
from warpsphboundaries_amd.fused import FusedGroup, FusedWall, WallOutput
from warpsphboundaries_amd.provider import AnalyticBoundary
from warpsphboundaries_amd.scene import (Body, DiskArrayRep, DiskBody, ImplicitRep, ParticleState, Scene, SurfaceRep,
                                         fixed_adjacency)

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
F64 = torch.float64
H = 0.3
BUNDLE = ([(0.0, 0.0), (0.9, 0.2), (-0.7, 0.6), (0.3, -0.8)], [0.1 * H, 0.6 * H, 2.0 * H, 6.0 * H])
KERNELS = ["cubic", "w2", "w4", "w6", "cone", "lw2", "lw4", "w2p5", "w4p5"]


def _ps(pos, sup=None):
    pos = torch.as_tensor(pos, dtype=F64, device=DEV)
    n = len(pos)
    sup = torch.full((n,), H, dtype=F64, device=DEV) if sup is None else torch.as_tensor(sup, dtype=F64, device=DEV)
    return ParticleState(pos, sup, kinds=torch.zeros(n, dtype=torch.int32, device=DEV))


def _queries(n=500, seed=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.6, 1.6, (n, 2))


FAMILIES = ([0], [1, 2], [3, 4], [5, 6, 7, 8], [9, 10, 11])     # lam, m1, g0, g1, m2


def _family_scale(ref, floor):
    """per channel the largest magnitude over its family (a rotation mixes the components of a vector or tensor)."""
    scale = np.empty(12)
    for f in FAMILIES:
        scale[f] = max(float(np.abs(ref[:, f]).max()), floor)
    return scale


def _oracle_rows(name, centres, radii, lpos, lsup, slots):
    """[N, 12] body-frame channels, units of h: the oracle's pair channels summed in slot order."""
    c = np.asarray(centres, dtype=float)
    r = np.asarray(radii, dtype=float)
    N, K = slots.shape
    out = np.zeros((N, 12))
    for s in range(K):
        rows = np.nonzero(slots[:, s] >= 0)[0]
        if len(rows) == 0:
            continue
        m = slots[rows, s]
        y = (c[m] - lpos[rows]) / lsup[rows, None]
        D = np.sqrt(y[:, 0] * y[:, 0] + y[:, 1] * y[:, 1])
        ch = OD.circle_channels(name, D, r[m] / lsup[rows])
        out[rows] += OD.assemble(ch, y)
    return out


def test_slots_bit_exact_with_the_oracle():
    cs, rs = BUNDLE
    rng = np.random.default_rng(3)
    pos = _queries(2000, seed=3)
    sup = H * rng.uniform(0.4, 1.0, len(pos))
    valid = rng.uniform(size=len(pos)) > 0.1
    rep = DiskArrayRep(cs, rs).to(DEV)
    g = rep.geom(H, DEV)
    print("disk slots: n", len(pos), "K", g.K, flush=True)
    lp = torch.as_tensor(pos, dtype=F64, device=DEV)
    ls = torch.as_tensor(sup, dtype=F64, device=DEV)
    slots, counts = g.adjacency(lp, ls, valid=torch.as_tensor(valid, device=DEV))
    got = slots.reshape(-1, g.K).cpu().numpy()
    cl = OD.cells(cs, rs, H)
    assert cl[-1] == g.K and (cl[2], cl[3]) == (g.info["nx"], g.info["ny"])
    ref = OD.disk_slots(cl, cs, rs, pos, sup, valid)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(counts.cpu().numpy(), (ref >= 0).sum(1))
    assert (got[~valid] == -1).all() and (ref >= 0).sum() > len(pos) // 2


@pytest.mark.parametrize("name", KERNELS)
def test_channels_match_the_oracle(name):
    cs, rs = BUNDLE
    rng = np.random.default_rng(5)
    # queries crowding the circles: on them, just inside / outside, near the centres, and uniform
    k = rng.integers(0, len(rs), 600)
    ang = rng.uniform(0, 2 * math.pi, 600)
    off = np.concatenate([np.zeros(60), 10.0 ** rng.uniform(-12, -1, 240) * rng.choice([-1, 1], 240),
                          rng.uniform(-1.2, 1.2, 300)]) * H
    rad = np.maximum(np.asarray(rs)[k] + off, 0.0)
    pos = np.asarray(cs)[k] + rad[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    pos = np.concatenate([pos, _queries(400, seed=6), np.asarray(cs) + 1e-10])
    sup = H * rng.uniform(0.5, 1.0, len(pos))
    rep = DiskArrayRep(cs, rs).to(DEV)
    g = rep.geom(H, DEV)
    lp = torch.as_tensor(pos, dtype=F64, device=DEV)
    ls = torch.as_tensor(sup, dtype=F64, device=DEV)
    print("disk channels:", name, "n", len(pos), flush=True)
    slots, counts = g.adjacency(lp, ls)
    got = g.moments(name, slots, counts, lp, ls).cpu().numpy()
    ref = _oracle_rows(name, cs, rs, pos, sup, slots.reshape(-1, g.K).cpu().numpy())
    err = (np.abs(got - ref) / _family_scale(ref, 1e-3)).max(0)
    assert err.max() < 1e-12, (name, err)


@pytest.mark.parametrize("kernels", [("cubic", "w2", "w4", "w6"), ("w2", "w4", "w6"), ("w4",), ("cubic", "w6"),
                                     ("w2", "w6")])
def test_several_kernels_in_one_pass(kernels):
    # one pass over the scheme kernels: the single-kernel values to rounding (with the cubic the others on the cubic's
    # finer panels), and the oracle at the single-kernel tolerance
    cs, rs = BUNDLE
    rng = np.random.default_rng(9)
    k = rng.integers(0, len(rs), 500)
    ang = rng.uniform(0, 2 * math.pi, 500)
    off = np.concatenate([10.0 ** rng.uniform(-12, -1, 200) * rng.choice([-1, 1], 200), rng.uniform(-1.2, 1.2, 300)]) * H
    pos = np.asarray(cs)[k] + np.maximum(np.asarray(rs)[k] + off, 0.0)[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    pos = np.concatenate([pos, _queries(300, seed=10), np.asarray(cs) + 1e-10])
    sup = H * rng.uniform(0.5, 1.0, len(pos))
    rep = DiskArrayRep(cs, rs).to(DEV)
    g = rep.geom(H, DEV)
    lp = torch.as_tensor(pos, dtype=F64, device=DEV)
    ls = torch.as_tensor(sup, dtype=F64, device=DEV)
    print("disk multi:", kernels, "n", len(pos), flush=True)
    slots, counts = g.adjacency(lp, ls)
    for flags in (0, L.OUT_PHYSICAL):
        multi = g.moments_multi(kernels, slots, counts, lp, ls, flags=flags)
        for name in kernels:
            single = g.moments(name, slots, counts, lp, ls, flags=flags).cpu().numpy()
            d = (np.abs(multi[name].cpu().numpy() - single) / _family_scale(single, 1e-3)).max()
            assert d <= (1e-13 if "cubic" in kernels else 1e-14), (name, d)
    multi = g.moments_multi(kernels, slots, counts, lp, ls)
    sl = slots.reshape(-1, g.K).cpu().numpy()
    for name in kernels:
        ref = _oracle_rows(name, cs, rs, pos, sup, sl)
        err = (np.abs(multi[name].cpu().numpy() - ref) / _family_scale(ref, 1e-3)).max(0)
        assert err.max() < 1e-12, (name, err)


def test_physical_world_units_of_a_rotated_bundle():
    cs, rs = BUNDLE
    pose = dict(center=(0.15, -0.1), angle=0.5)
    sc = Scene([Body(bodyId=0, reps=[DiskArrayRep(cs, rs)], **pose)], DEV)
    rng = np.random.default_rng(8)
    pos = _queries(800, seed=8)
    sup = H * rng.uniform(0.6, 1.0, len(pos))
    ps = _ps(pos, sup)
    adj = fixed_adjacency(sc, ps, None, H)
    fw = FusedWall(sc, adj, (FusedGroup("w4", None, "fp64"),))
    got = fw.channels[0, 0].cpu().numpy()
    ba = adj.bodies[0]
    lpos, topo = ba.lpos.cpu().numpy(), ba.reps[0]
    ref = _oracle_rows("w4", cs, rs, lpos, sup, topo.e.reshape(-1, topo.K).cpu().numpy())
    ref[:, [1, 2, 9, 10, 11]] *= np.concatenate([np.tile(sup[:, None], 2), np.tile(sup[:, None] ** 2, 3)], 1)
    ref[:, [3, 4]] /= sup[:, None]
    c, s = math.cos(0.5), math.sin(0.5)
    Rm = np.array([[c, -s], [s, c]])
    for a, b in ((1, 2), (3, 4)):
        ref[:, [a, b]] = ref[:, [a, b]] @ Rm.T
    G = ref[:, 5:9].reshape(-1, 2, 2)
    ref[:, 5:9] = (Rm @ G @ Rm.T).reshape(-1, 4)
    M2 = np.stack([ref[:, 9], ref[:, 10], ref[:, 10], ref[:, 11]], 1).reshape(-1, 2, 2)
    M2 = Rm @ M2 @ Rm.T
    ref[:, 9], ref[:, 10], ref[:, 11] = M2[:, 0, 0], M2[:, 0, 1], M2[:, 1, 1]
    assert (np.abs(got - ref) / _family_scale(ref, 1e-6)).max() < 1e-12


def _layout_outputs(scene, pos):
    ab = AnalyticBoundary(scene)
    ag = ab.aggregate(_ps(pos), H, "w2", laplacian=True)
    out = {k: v.sum(0) for k, v in ag.out.items()}
    more = ag.evaluate((WallOutput("m1", 0, "m1"), WallOutput("m2", 0, "m2")))
    out.update({k: v.sum(0) for k, v in more.items()})
    return out


def test_bundle_equals_the_sum_of_polygon_disks():
    cs, rs = BUNDLE
    pose = dict(center=(0.15, -0.1), angle=0.5)
    bundle = Scene([Body(bodyId=0, reps=[DiskArrayRep(cs, rs)], **pose)], DEV)
    polys = Scene([Body(bodyId=i, reps=[SurfaceRep.regularPolygon(c, r, max(24, int(math.ceil(2 * math.pi * r / (H / 16)))))],
                        **pose) for i, (c, r) in enumerate(zip(cs, rs))], DEV)
    pos = _queries()
    got, ref = _layout_outputs(bundle, pos), _layout_outputs(polys, pos)
    for k in ("lam", "G", "Cov", "m1", "m2", "cover", "lap", "tens"):
        e = float((got[k] - ref[k]).abs().max()) / max(float(ref[k].abs().max()), 1e-12)
        assert e < (1e-4 if k == "cover" else 5e-6), (k, e)
    assert float(ref["lam"].max()) > 0.5


def test_solid_implicit_disk_is_the_disk_element_and_geometry_queries():
    c, R = (0.3, 0.4), 0.2
    rng = np.random.default_rng(9)
    pos = np.asarray(c) + rng.uniform(-0.6, 0.6, (700, 2))
    outs = (WallOutput("lam", 0, "lam"), WallOutput("g", 0, "g0"), WallOutput("cov", 0, "cov"), WallOutput("m2", 0, "m2"))
    res = []
    for rep in (ImplicitRep(DiskBody(c, R)), DiskArrayRep([c], [R])):
        sc = Scene([Body(bodyId=0, reps=[rep], center=(0.05, -0.02), angle=0.3)], DEV)
        res.append(FusedWall(sc, fixed_adjacency(sc, _ps(pos), None, H), (FusedGroup("w4"),)).evaluate(outs))
    for k in res[0]:
        assert torch.equal(res[0][k], res[1][k]), k
    cs, rs = BUNDLE
    sc = Scene([Body(bodyId=0, center=(0.2, 0.1), angle=0.9, reps=[DiskArrayRep(cs, rs)])], DEV)
    b = sc.bodies[0]
    wpos = b.pose.toWorld(torch.as_tensor(_queries(400, seed=5), dtype=F64, device=DEV))
    d, n, hit = sc.signed_distance(wpos)
    lp = b.pose.toLocal(wpos).cpu().numpy()
    dist = np.linalg.norm(lp[:, None, :] - np.asarray(cs)[None], axis=2) - np.asarray(rs)[None]
    assert np.abs(d.cpu().numpy() - dist.min(1)).max() < 1e-12 and bool(hit.all())
    kk = dist.argmin(1)
    nl = lp - np.asarray(cs)[kk]
    nl = nl / np.linalg.norm(nl, axis=1, keepdims=True)
    nw = b.pose.vecToWorld(torch.as_tensor(nl, dtype=F64, device=DEV))
    assert float((nw - n).abs().max()) < 1e-12
    assert bool((sc.inside(wpos) == (d < 0)).all())


def test_adjacency_and_moments_replay_in_a_captured_graph():
    """the two entry points allocate nothing and do not synchronise: captured once into a HIP graph, every replay
    (with eager work and synchronisation between the replays) equals the eager result bitwise."""
    cs, rs = BUNDLE
    g = DiskArrayRep(cs, rs).to(DEV).geom(H, DEV)
    lp = torch.as_tensor(_queries(3000, seed=4), dtype=F64, device=DEV)
    ls = torch.full((len(lp),), H, dtype=F64, device=DEV)
    slots = torch.empty(len(lp) * g.K, dtype=torch.int32, device=DEV)
    counts = torch.empty(len(lp), dtype=torch.int32, device=DEV)
    out = torch.empty((len(lp), 12), dtype=F64, device=DEV)
    g.adjacency(lp, ls, slots=slots, counts=counts)
    eager = g.moments("w4", slots, counts, lp, ls, out=out).clone()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        with torch.cuda.graph(graph, stream=stream):
            g.adjacency(lp, ls, slots=slots, counts=counts)
            g.moments("w4", slots, counts, lp, ls, out=out)
    torch.cuda.synchronize()
    for rep in range(3):
        out.fill_(float(rep))
        slots.fill_(7)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep


def _numeric_cone_areas(centres, radii, x, axis, half, H, nth=200000):
    """area of (disks cap disk(x, H) cap wedge(x, axis, half)) and of (disks cap disk(x, H)) by integrating the ray
    lengths over the angle (midpoint rule on the wedge's own angle interval and on the full circle; independent of the
    closed form)."""
    def sweep(th, dth):
        d = np.stack([np.cos(th), np.sin(th)], 1)
        tot = 0.0
        for c, r in zip(centres, radii):
            v = np.asarray(c) - x
            b = d @ v
            disc = b * b - (v @ v - r * r)
            hit = disc > 0
            sq = np.sqrt(np.where(hit, disc, 0.0))
            lo, hi = np.clip(b - sq, 0, H), np.clip(b + sq, 0, H)
            tot += np.where(hit, 0.5 * (hi ** 2 - lo ** 2), 0.0).sum() * dth
        return tot
    full = sweep((np.arange(nth) + 0.5) * (2 * np.pi / nth) - np.pi, 2 * np.pi / nth)
    if half >= np.pi:
        return full, full
    a0 = np.arctan2(axis[1], axis[0]) - half
    return sweep(a0 + (np.arange(nth) + 0.5) * (2 * half / nth), 2 * half / nth), full


@pytest.mark.parametrize("half", [math.pi / 6, math.pi / 2, 2.5, math.pi])
def test_cone_area_closed_form_equals_the_numerical_angular_integral(half):
    cs, rs = [(0.0, 0.0), (0.45, 0.1), (-0.3, 0.5)], [0.15, 0.25, 0.4]
    sc = Scene([Body(bodyId=0, center=(0.1, 0.2), angle=0.7, reps=[DiskArrayRep(cs, rs)])], DEV)
    rng = np.random.default_rng(3)
    lpos_h = rng.uniform(-0.6, 0.9, (120, 2))
    b = sc.bodies[0]
    pos = b.pose.toWorld(torch.as_tensor(lpos_h, dtype=F64, device=DEV))
    ax = torch.as_tensor(rng.normal(size=(120, 2)), dtype=F64, device=DEV)
    fw = FusedWall(sc, fixed_adjacency(sc, _ps(pos), None, H), (FusedGroup("w2"),))
    got = fw.cone_area(ax, half).cpu().numpy()
    lp = b.pose.toLocal(pos).cpu().numpy()
    la = (ax @ b.pose.R.to(ax.device)).cpu().numpy()                      # world axis -> body frame (R^T a)
    worst = 0.0
    for i in range(0, 120, 4):
        w, f = _numeric_cone_areas(cs, rs, lp[i], la[i], half, H)
        worst = max(worst, abs(got[0, i] - w), abs(got[1, i] - f))
    assert worst < 1e-7 * H * H * 100, worst
    assert float(got[1].max()) > 0.1 * H * H


def test_dir_extreme_equals_a_sampled_maximum():
    cs, rs = [(0.0, 0.0), (0.45, 0.1), (-0.3, 0.5)], [0.15, 0.25, 0.4]
    sc = Scene([Body(bodyId=0, center=(0.1, 0.2), angle=0.7, reps=[DiskArrayRep(cs, rs)])], DEV)
    rng = np.random.default_rng(12)
    b = sc.bodies[0]
    pos = b.pose.toWorld(torch.as_tensor(rng.uniform(-0.6, 0.9, (80, 2)), dtype=F64, device=DEV))
    vec = torch.as_tensor(rng.normal(size=(1, 80, 2)), dtype=F64, device=DEV)
    fw = FusedWall(sc, fixed_adjacency(sc, _ps(pos), None, H), (FusedGroup("w2"),))
    got = fw.dir_extreme(vec)[0].cpu().numpy()
    lp = b.pose.toLocal(pos).cpu().numpy()
    lv = (vec[0] @ b.pose.R.to(DEV)).cpu().numpy()
    # sampled: points of every disk on a fine polar grid, kept within H of the query
    ref = np.zeros(80)
    for c, r in zip(cs, rs):
        rr, tt = np.meshgrid(np.linspace(0, r, 400), np.linspace(0, 2 * np.pi, 1600, endpoint=False))
        pts = np.asarray(c) + np.stack([rr.ravel() * np.cos(tt.ravel()), rr.ravel() * np.sin(tt.ravel())], 1)
        for i in range(80):
            v = pts - lp[i]
            dist = np.linalg.norm(v, axis=1)
            ok = (dist < H) & (dist > 0)
            if ok.any():
                cosv = np.abs(v[ok] @ lv[i]) / (dist[ok] * np.linalg.norm(lv[i]))
                ref[i] = max(ref[i], cosv.max())
    assert np.all(got >= ref - 1e-12) and np.abs(got - ref).max() < 5e-3, np.abs(got - ref).max()
    assert (got > 0).sum() > 20


def test_bad_disk_array_is_rejected():
    import ctypes
    with pytest.raises(ValueError):
        L.check(L.load().wsb_disks2d_create(np.zeros(2).ctypes.data, np.array([-1.0]).ctypes.data, 1, 0.1,
                                            ctypes.byref(ctypes.c_void_p()), ctypes.byref(ctypes.c_int())))
