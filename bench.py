"""Benchmark of the wall-integral hot path on one MI355X.

Workload (the 2-D at-scale configuration): one star polygon of 4,096 edges, 1,048,576 particles (30 % within one
support of the wall), h = 0.004, kernels cubic, w2, w4, w6.  One step = the adjacency rebuild (bit-exact cell-list
slots + winding indicator) followed by all twelve moments for each of the four kernels (4 x 12 x N particle-moment
evaluations), inputs resident in HBM.  ``value`` is the mixed form's throughput; the all-FP64 plain path (plan
interpreter, one thread per pair) and the FP64 closed form are timed beside it, the mixed form's error is measured
against the FP64 closed form on the device, and the CPU oracle is timed on a bounded subsample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12          # B/s
PEAK_FP64 = 78.6e12        # flop/s (MI355X FP64 vector; the matrix FP64 peak is the same figure)
METRIC = "particle-moment evals/s @≤1e-6 rel err vs FP64; served p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def active_block_counts(torch, verts, edges, slots, counts, lpos, lsup, K, kname):
    """number of active (pair, block) chords of every block radius, from the pair geometry (for the op model)."""
    from warpsphboundaries_amd.plan import block_profiles
    rows = torch.repeat_interleave(torch.arange(len(counts), device=counts.device), counts.long())
    off = torch.cumsum(counts.long(), 0) - counts.long()
    k = torch.arange(len(rows), device=rows.device) - off[rows]
    e = slots.view(-1, K)[rows, k].long()
    h = lsup[rows]
    p = (verts[edges[e, 0]] - lpos[rows]) / h[:, None]
    q = (verts[edges[e, 1]] - lpos[rows]) / h[:, None]
    d = q - p
    ell = torch.sqrt((d * d).sum(1))
    z = (p[:, 0] * q[:, 1] - p[:, 1] * q[:, 0]) / ell
    s0 = (d * p).sum(1) / ell
    s1 = (d * q).sum(1) / ell
    out = []
    for b in block_profiles(kname):
        R = float(b.R)
        L = torch.sqrt(torch.clamp((R - z.abs()) * (R + z.abs()), min=0))
        act = (z.abs() < R) & (torch.maximum(s0, -L) < torch.minimum(s1, L))
        out.append(int(act.sum()))
    return len(rows), out


def active_radius_counts(torch, verts, edges, slots, counts, lpos, lsup, K, knames):
    """{block radius: active (pair, radius) chords} over the union of the kernels' block radii (the fused pass
    computes each radius's chord primitives once for every kernel that has a block there)."""
    from warpsphboundaries_amd.plan import block_profiles
    radii = sorted({float(b.R) for k in knames for b in block_profiles(k)})
    res = {}
    for k in knames:
        npairs, act = active_block_counts(torch, verts, edges, slots, counts, lpos, lsup, K, k)
        for b, a in zip(block_profiles(k), act):
            res[float(b.R)] = a
    assert set(res) == set(radii)
    return npairs, res


def cpu_baseline(work, n_sample=65536, seed=7, runs=5):
    """the reference's CPU edge backend restated on torch CPU tensors (``oracle/torch2d.py``: FP64, every host thread
    this process may use): slot fill of the static cell list, winding indicator, twelve channels per pair and row
    sums.  The star workload on a stratified subsample that keeps its near-wall fraction (four kernels), and the
    box + 64-gon scene in full (w4, two representations); one warm-up, then the median of ``runs`` runs of each."""
    import torch
    from oracle import torch2d as OT
    from warpsphboundaries_amd import workloads as WL
    from warpsphboundaries_amd.scene import SurfaceRep
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)

    def prepared(loops, h):
        out = []
        for loop, solid in loops:
            rep = SurfaceRep.polygon(loop, solid)
            out.append(OT.Scene2D(rep.vertices.numpy(), rep.edges.numpy(), rep.background, h))
        return out

    def timed(scenes, pos, sup, kernels):
        def run():
            for sc in scenes:
                OT.twelve(sc, pos, sup, kernels)
        run()
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            run()
            times.append(time.perf_counter() - t0)
        return float(np.median(times))

    sub = WL.subsample(work, n_sample, seed)
    pos, sup = torch.tensor(sub.positions), torch.tensor(sub.supports)
    t_star = timed(prepared(sub.loops, sub.h), pos, sup, sub.kernels)
    star = 12 * len(sub.kernels) * len(pos) / t_star
    c1 = WL.config1()
    t_box = timed(prepared(c1.loops, c1.h), torch.tensor(c1.positions), torch.tensor(c1.supports), c1.kernels)
    box = 12 * len(c1.kernels) * len(c1.positions) / t_box
    return {"value": star, "unit": "particle-moment evals/s", "cores": cores, "kind": "port",
            "sample": f"{len(pos)}-particle stratified subsample of the star workload ({int(sub.near.sum())} near-wall), "
                      f"slot fill + winding indicator + twelve channels + row sums for cubic/w2/w4/w6, torch-CPU FP64 "
                      f"restatement of the reference's CPU edge backend on {cores} threads, median of {runs} runs of "
                      f"{t_star:.2f} s",
            "box_64gon_full": {"value": box, "particles": len(c1.positions), "seconds": t_box,
                               "note": "16,384 particles, w4, both representations, same restatement and threads"}}


_SERVED = {}


def served_scenes(n_scenes=256, seed=5):
    """the served workload's scenes with their resident geometry built once (2-D SurfaceReps cache their device
    geometry; 3-D Mesh3Ds), shared by every server of this process."""
    key = (n_scenes, seed)
    if key not in _SERVED:
        import torch
        from warpsphboundaries_amd import workloads as WL
        from warpsphboundaries_amd.mesh3d import Mesh3D
        from warpsphboundaries_amd.scene import SurfaceRep
        scenes = WL.config5(n_scenes=n_scenes, seed=seed)
        built = {}
        for sc in scenes:
            if sc.dim == 2:
                built[sc.name] = [SurfaceRep.polygon(v, solid) for v, solid in sc.loops]
            else:
                built[sc.name] = Mesh3D(sc.mesh[0], sc.mesh[1], sc.h, device=torch.device("cuda:0"))
        _SERVED[key] = (scenes, built)
    return _SERVED[key]


def serve_measure(form, qps, n_scenes=256, steps=1, graphs=False, streams=4, seed=5, batched=None, device_requests=False,
                  max_batch_rows=1 << 20):
    """the served mixed-size workload: every scene resident, one request per scene and timestep arriving as a Poisson
    process at `qps`; continuous batching over `streams` streams.  Returns latency percentiles and throughput."""
    import torch
    from warpsphboundaries_amd import workloads as WL
    from warpsphboundaries_amd.serve import BoundaryServer
    scenes, built = served_scenes(n_scenes, seed)
    srv = BoundaryServer(streams=streams, form=form, graphs=graphs, hbm_budget=64 << 30, batched=batched,
                         max_batch_rows=max_batch_rows)
    for sc in scenes:
        if sc.dim == 2:
            srv.add_scene(sc.name, built[sc.name], sc.h)
        else:
            srv.add_scene(sc.name, None, sc.h, mesh=built[sc.name])
    work = []
    for t in range(steps):
        for i, sc in enumerate(scenes):
            pos = WL.jitter(sc.positions, sc.h, 1000 * t + i)
            if device_requests:             # the simulations' particle arrays live on this GPU
                pos = torch.tensor(pos, dtype=torch.float64, device=srv.device)
            work.append((sc.name, pos, sc.h, sc.kernel, 12 if sc.dim == 2 else 22))
    order = WL.rng(seed + 1).permutation(len(work))
    work = [work[i] for i in order]
    arrivals = WL.poisson_arrivals(len(work), qps, seed + 2)

    def arrival_pass():
        t0 = time.perf_counter()
        reqs = []
        i = 0
        while i < len(work) or srv.pending or any(b is not None for b in srv.busy):
            now = time.perf_counter() - t0
            while i < len(work) and arrivals[i] <= now:
                name, pos, h, k, _ = work[i]
                r = srv.submit(name, pos, h, k)
                r.arrival = t0 + arrivals[i]
                reqs.append(r)
                i += 1
            srv.poll()
        return reqs, time.perf_counter() - t0

    # warm pass, not timed: the same requests all at once (per-stream buffers of every row bucket, first kernel
    # loads), then once with the same arrival schedule (the batches a server in steady state has captured), then the
    # timed pass
    for name, pos, h, k, _ in work:
        srv.submit(name, pos, h, k)
    srv.run()
    arrival_pass()
    torch.cuda.synchronize()
    srv.completed.clear()
    srv.launched_batches = 0
    srv.relaunched_batches = 0
    srv.captured_batches = 0
    srv.replayed_batches = 0
    srv.host_seconds = {k: 0.0 for k in srv.host_seconds}
    reqs, wall = arrival_pass()
    rows = sum(r.n for r in reqs)
    width = {name: wd for name, _, _, _, wd in work}
    evals = sum(r.n * width[r.scene] for r in reqs)
    rows3 = sum(r.n for r in reqs if width[r.scene] == 22)
    st = srv.latency_stats(reqs)
    st.update({"form": form, "qps_offered": qps, "requests": len(reqs), "rows": rows, "rows_3d": rows3, "wall_s": wall,
               "captured_batches": srv.captured_batches, "replayed_batches": srv.replayed_batches,
               "evals_per_s": evals / wall, "batches": srv.launched_batches, "relaunched_3d_batches": srv.relaunched_batches,
               "streams": streams, "graphs": graphs, "cross_scene_batches": srv.batched,
               "device_resident_requests": bool(device_requests), "max_batch_rows": max_batch_rows,
               "host_seconds": dict(srv.host_seconds)})
    return st


def search3d_compare(torch, dev, reps=3):
    """the two 3-D face searches (cell lists, face tree) on the 64 meshes of the served workload: face lists and
    winding numbers compared, per-scene adjacency launches timed with HIP events and summed."""
    from warpsphboundaries_amd import workloads as WL
    from warpsphboundaries_amd.mesh3d import Mesh3D
    scenes = [s for s in WL.config5(n_scenes=256, seed=5) if s.dim == 3]
    tot = {"grid": 0.0, "bvh": 0.0}
    pairs, rows = 0, 0
    for s in scenes:
        V, F = s.mesh
        m = Mesh3D(V, F, s.h, device=dev)
        P = torch.tensor(s.positions, dtype=torch.float64, device=dev)
        H = torch.full((len(P),), s.h, dtype=torch.float64, device=dev)
        o = m.sort(P)
        a = m.adjacency(P, H, order=o, search="grid")
        b = m.adjacency(P, H, order=o, search="bvh")
        if not (a.count == b.count and torch.equal(a.faces[:a.count], b.faces[:b.count]) and torch.equal(a.chi, b.chi)):
            raise RuntimeError(f"face searches differ on {s.name}")
        pairs += a.count
        rows += len(P)
        for search in ("grid", "bvh"):
            m.adjacency(P, H, order=o, search=search, capacity=a.faces.numel())
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                m.adjacency(P, H, order=o, search=search, capacity=a.faces.numel())
            e1.record()
            torch.cuda.synchronize()
            tot[search] += e0.elapsed_time(e1) / reps
        m.close()
    return {"scenes": len(scenes), "rows": rows, "pairs": pairs, "adjacency_ms_summed": tot,
            "equal": True}


def mesh3d_measure(torch, dev, steps, warmup):
    """the 3-D mesh path on the 81,920-face sphere cavity workload: adjacency rebuild + the 22 stored moments (w4)."""
    from warpsphboundaries_amd import workloads as WL
    from warpsphboundaries_amd.mesh3d import Mesh3D
    log("generating the 3-D workload ...")
    w = WL.config3()
    N = len(w.positions)
    t0 = time.perf_counter()
    mesh = Mesh3D(w.vertices, w.faces, w.h, device=dev)
    t_build = time.perf_counter() - t0
    pos = torch.tensor(w.positions, dtype=torch.float64, device=dev)
    sup = torch.tensor(w.supports, dtype=torch.float64, device=dev)
    adj = mesh.adjacency(pos, sup)
    cap = adj.faces.numel()
    adj_s = mesh.adjacency(pos, sup, capacity=cap, layout="sorted")
    out = torch.empty((N, 22), dtype=torch.float64, device=dev)
    kern = w.kernels[0]

    def step(form, layout="sorted"):
        # cell sort, face adjacency + winding numbers (entries at the cell-order positions), 22 moments at the rows
        a = mesh.adjacency(pos, sup, capacity=cap, layout=layout)
        mesh.moments(a, pos, sup, kern, form, out=out)

    res = {}
    for form in ("fp64", "mixed"):
        for _ in range(warmup):
            step(form)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(form)
        torch.cuda.synchronize()
        res[form] = (time.perf_counter() - t0) / steps
        res[form + "_out"] = out.clone()
        log(f"3-D {form}: {res[form] * 1e3:.3f} ms/step")
    # the particle (rows) layout of the same step: the same values bitwise
    t0 = time.perf_counter()
    for _ in range(steps):
        step("mixed", "rows")
    torch.cuda.synchronize()
    res["rows_mixed"] = (time.perf_counter() - t0) / steps
    assert torch.equal(out, res["mixed_out"])
    stream = torch.cuda.current_stream()
    per = {}
    for name, fn in (("adjacency", lambda: mesh.adjacency(pos, sup, capacity=cap, layout="sorted")),
                     ("adjacency_rows_layout", lambda: mesh.adjacency(pos, sup, capacity=cap)),
                     ("adjacency_face_tree", lambda: mesh.adjacency(pos, sup, capacity=cap, search="bvh")),
                     ("moments_fp64", lambda: mesh.moments(adj_s, pos, sup, kern, "fp64", out=out)),
                     ("moments_mixed", lambda: mesh.moments(adj_s, pos, sup, kern, "mixed", out=out)),
                     ("moments_mixed_rows_layout", lambda: mesh.moments(adj, pos, sup, kern, "mixed", out=out))):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        per[name] = e0.elapsed_time(e1) / 1e3 / steps
    ref = res["fp64_out"]
    scale = ref.abs().amax(0).clamp(min=1e-300)
    err = float(((res["mixed_out"] - ref).abs().amax(0) / scale).max())
    chi = adj.chi
    evals = 22 * N
    # algorithmic bytes of the moment launch: positions, support, offsets, winding number in; 22 float64 out;
    # the face list and the face records of every pair
    bytes_mom = N * (24 + 8 + 8 + 8 + 22 * 8) + adj.count * (4 + 36 * 8)
    return {
        "workload": f"{w.name} (box [0,1]^3 minus a level-6 icosphere of radius 0.25), {N:,} particles, w4, h=0.01; "
                    "step = cell sort + face adjacency + winding rebuild (sorted-position layout) + 22 stored moments",
        "N": N, "faces": int(len(w.faces)), "pairs": int(adj.count), "pairs_per_particle": adj.count / N,
        "winding_values": sorted({float(v) for v in torch.unique(chi).cpu().tolist()}),
        "static_build_s": t_build, "grid": list(mesh.grid_shape), "K": mesh.K,
        "ms_per_step": {"fp64": res["fp64"] * 1e3, "mixed": res["mixed"] * 1e3,
                        "mixed_rows_layout": res["rows_mixed"] * 1e3},
        "evals_per_s": {"fp64": evals / res["fp64"], "mixed": evals / res["mixed"]},
        "launch_ms": {k: v * 1e3 for k, v in per.items()},
        "moments_hbm_GBps_algorithmic": bytes_mom / per["moments_fp64"] / 1e9,
        "moments_algorithmic_bytes": bytes_mom,
        "max_signal_relative_error_mixed_vs_fp64": err,
    }


def disks_measure(torch, dev, steps, warmup):
    """the disk element on a fibre bundle (DiskArrayRep): 4,096 solid fibres, 1,048,576 particles outside them,
    h = 0.006; step = disk slot list + twelve moments for each of cubic, w2, w4, w6 (FP64) in one pass; the same step
    with one launch per kernel is timed beside it."""
    from warpsphboundaries_amd import workloads as WL
    from warpsphboundaries_amd.scene import DiskArrayRep
    log("generating the fibre bundle workload ...")
    w = WL.fibre_bundle()
    N = len(w.positions)
    g = DiskArrayRep(w.centres, w.radii).to(dev).geom(w.h, dev)
    lpos = torch.tensor(w.positions, dtype=torch.float64, device=dev)
    lsup = torch.tensor(w.supports, dtype=torch.float64, device=dev)
    slots = torch.empty(N * g.K, dtype=torch.int32, device=dev)
    counts = torch.empty(N, dtype=torch.int32, device=dev)
    outs = {k: torch.empty((N, 12), dtype=torch.float64, device=dev) for k in w.kernels}

    def step():
        g.adjacency(lpos, lsup, slots=slots, counts=counts)
        g.moments_multi(w.kernels, slots, counts, lpos, lsup, outs=outs)

    def step_single():
        g.adjacency(lpos, lsup, slots=slots, counts=counts)
        for k in w.kernels:
            g.moments(k, slots, counts, lpos, lsup, out=outs[k])

    t = {}
    for name, fn in (("single", step_single), ("multi", step), ("single", step_single)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        t[name] = (time.perf_counter() - t0) / steps      # the single-launch step: its second timing
    t_step = t["multi"]
    stream = torch.cuda.current_stream()
    per = {}
    for name, fn in [("adjacency", lambda: g.adjacency(lpos, lsup, slots=slots, counts=counts)),
                     ("moments_multi", lambda: g.moments_multi(w.kernels, slots, counts, lpos, lsup, outs=outs))] + \
            [("moments_" + k, (lambda k=k: g.moments(k, slots, counts, lpos, lsup, out=outs[k]))) for k in w.kernels]:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        per[name] = e0.elapsed_time(e1) / 1e3 / steps
    pairs = int(counts.sum())
    evals = 12 * N * len(w.kernels)
    # algorithmic bytes of one moments launch: position, support, count and slots in, twelve float64 out, one 32-byte
    # disk record per pair
    bytes_mom = N * (16 + 8 + 4 + 4 * g.K + 96) + pairs * 32
    return {
        "workload": f"{w.name}: {len(w.radii)} solid fibres (radii 0.2-0.3 of the 1/64 lattice, jittered), {N:,} "
                    f"particles outside them, h = {w.h}; step = disk slot list + twelve moments for cubic, w2, w4, w6 "
                    f"in one pass",
        "N": N, "disks": int(len(w.radii)), "pairs": pairs, "K": g.K,
        "ms_per_step": t_step * 1e3, "evals_per_s": evals / t_step,
        "ms_per_step_one_launch_per_kernel": t["single"] * 1e3,
        "launch_ms": {k: v * 1e3 for k, v in per.items()},
        "pairs_per_s_w4": pairs / per["moments_w4"],
        "moments_w4_hbm_GBps_algorithmic": bytes_mom / per["moments_w4"] / 1e9,
        "bound": "fp64 VALU issue (16-node Gauss panels of the arc-angle form per pair)",
    }


def _traffic(kernel):
    """HBM bytes per launch of the mixed-form fused twelve-moment pass on the star workload, from the committed
    FETCH_SIZE / WRITE_SIZE counter runs scaled on a copy kernel of known size (None when the summary is absent)."""
    fn = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic_rows12_multi_star4096_1M.json")
    try:
        with open(fn) as f:
            return json.load(f)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def _valu(kernel):
    """VALU busy and issue fractions of ``kernel`` from the committed SQ counter runs (None when absent)."""
    fn = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_valu_hot_kernels.json")
    try:
        with open(fn) as f:
            d = json.load(f)["kernels"][kernel]
        return {"valu_busy": d["valu_busy"], "valu_issue": d["valu_issue"], "fp64_share_of_valu": d["fp64_share_of_valu"],
                "wait_share": d["wait_share"], "source": "profiles/pmc_valu_hot_kernels.json (rocprofv3 --pmc, "
                                                         "SQ_ACTIVE_INST_VALU x 4 / (1024 SIMDs x GRBM_GUI_ACTIVE / 8))"}
    except (OSError, KeyError, ValueError):
        return None


CH2 = ["m00", "m10", "m01", "Gx", "Gy", "g1xx", "g1xy", "g1yx", "g1yy", "m20", "m11", "m02"]


def precision_mp(torch, dev):
    """Per-moment error of every form on the 4,096-particle wedge subsample whose values were computed with mpmath
    at 50 digits (tests/golden/stress2d_mp.npz): signal-relative (max |v - r| / max |r| over one wedge's particles,
    then the maximum over wedges and kernels) and absolute in units of h."""
    from warpsphboundaries_amd.scene import SurfaceRep
    fn = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "stress2d_mp.npz")
    d = np.load(fn)
    off, poly, pts, h = d["verts_offsets"], d["poly"], d["positions"], float(d["h"])
    kernels = [str(k) for k in d["kernels"]]
    forms = ("plain", "fp64", "mixed")
    got = {(f, k): np.zeros((len(pts), 12)) for f in forms for k in kernels}
    for i in range(len(off) - 1):
        sel = np.nonzero(poly == i)[0]
        if len(sel) == 0:
            continue
        geo = SurfaceRep.polygon(d["verts"][off[i]:off[i + 1]], "inside").geom(h, dev)
        lp = torch.tensor(pts[sel], dtype=torch.float64, device=dev)
        ls = torch.full((len(sel),), h, dtype=torch.float64, device=dev)
        slots, counts, ind = geo.adjacency(lp, ls)
        for f in forms:
            for k in kernels:
                got[(f, k)][sel] = geo.moments(k, slots, counts, ind, lp, ls, form=f).cpu().numpy()
    out = {"particles": int(len(pts)), "wedges": int(len(off) - 1), "channels": CH2, "reference": "mpmath, 50 digits"}
    for f in forms:
        sig, ab, per_k = np.zeros(12), np.zeros(12), {}
        for ki, k in enumerate(kernels):
            ref, v = d["values"][ki], got[(f, k)]
            worst = np.zeros(12)
            for c in np.unique(poly):
                m = poly == c
                e = np.abs(v[m] - ref[m]).max(0)
                s = np.abs(ref[m]).max(0)
                worst = np.maximum(worst, np.where(s > 0, e / np.where(s > 0, s, 1.0), e))
            sig = np.maximum(sig, worst)
            ab = np.maximum(ab, np.abs(v - ref).max(0))
            per_k[k] = float(worst.max())
        out[f] = {"signal_relative": [float(x) for x in sig], "absolute_units_of_h": [float(x) for x in ab],
                  "worst_signal_relative_per_kernel": per_k}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--qps", type=float, default=2000.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--only-3d", action="store_true", help="run the 3-D section alone and print it")
    ap.add_argument("--search-3d", action="store_true",
                    help="compare the two 3-D face searches on the served workload's meshes, print it and exit")
    ap.add_argument("--no-disks", action="store_true")
    ap.add_argument("--only-disks", action="store_true", help="run the fibre-bundle disk section alone and print it")
    ap.add_argument("--probe", action="store_true",
                    help="counter probe: one mixed step of the star workload and a 512 MiB device copy, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the wall-integral path shards by particle ranges with replicated geometry "
              "(no collective during evaluation), and this build runs on a single MI355X; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from warpsphboundaries_amd import workloads as WL
    from warpsphboundaries_amd import opmodel
    from warpsphboundaries_amd.scene import SurfaceRep
    from warpsphboundaries_amd import _lib as L

    dev = torch.device("cuda:0")
    if args.search_3d:
        print(json.dumps({"search_3d": search3d_compare(torch, dev)}), flush=True)
        return
    if args.only_3d:
        print(json.dumps(mesh3d_measure(torch, dev, args.steps, args.warmup)), flush=True)
        return
    if args.only_disks:
        print(json.dumps(disks_measure(torch, dev, args.steps, args.warmup)), flush=True)
        return
    log("generating the star workload ...")
    work = WL.config2(n=args.n)
    loop, solid = work.loops[0]
    rep = SurfaceRep.polygon(loop, solid)
    lpos = torch.tensor(work.positions, dtype=torch.float64, device=dev)
    lsup = torch.tensor(work.supports, dtype=torch.float64, device=dev)
    N = len(work.positions)
    geo = rep.geom(work.h, dev)
    K = geo.K
    log(f"N={N} edges={len(loop)} K={K} cells={geo.info['nx']}x{geo.info['ny']}")
    slots = torch.empty(N * K, dtype=torch.int32, device=dev)
    counts = torch.empty(N, dtype=torch.int32, device=dev)
    ind = torch.empty(N, dtype=torch.float64, device=dev)
    order = torch.empty(N, dtype=torch.int32, device=dev)
    sort_scratch = torch.empty(int(L.load().wsb_sort2d_scratch_bytes(geo.handle, N)), dtype=torch.uint8, device=dev)
    corder = torch.empty(N, dtype=torch.int32, device=dev)
    nwork = torch.empty(1, dtype=torch.int32, device=dev)
    corder_scratch = torch.empty(int(L.load().wsb_count_order2d_scratch_bytes()), dtype=torch.uint8, device=dev)
    outs = {k: torch.empty((N, 12), dtype=torch.float64, device=dev) for k in work.kernels}
    # the sorted-position layout of the step: position i holds row order[i]'s slots, count and (x, y, h, indicator)
    slots_s = torch.empty(N * K, dtype=torch.int32, device=dev)
    counts_s = torch.empty(N, dtype=torch.int32, device=dev)
    prow = torch.empty((N, 4), dtype=torch.float64, device=dev)

    def evaluate(form, scratch=None):
        if form == "plain":                 # the plan interpreter: one launch per kernel
            for k in work.kernels:
                geo.moments(k, slots, counts, ind, lpos, lsup, form=form, out=outs[k], scratch=scratch)
        else:                               # the closed form: the four kernels in one pass sharing the chords,
            # one row per thread, rows served in cell order from the sorted-position layout
            geo.moments_multi_sorted(work.kernels, order, slots_s, counts_s, prow, form=form, outs=outs)

    def evaluate_rows(form):
        """the same pass over the row layout (each row's record read at the row)"""
        geo.moments_multi(work.kernels, slots, counts, ind, lpos, lsup, form=form, outs=outs, order=order)

    def evaluate_by_count(form):
        """the same pass in the count order (rows with slots grouped by count, indicator terms stored apart)"""
        geo.moments_multi(work.kernels, slots, counts, ind, lpos, lsup, form=form, outs=outs, order=corder, nwork=nwork)

    def step(form, scratch=None):
        geo.sort(lpos, order=order, scratch=sort_scratch)
        if form == "plain":
            geo.adjacency(lpos, lsup, pad=False, slots=slots, counts=counts, ind=ind, order=order)
        else:
            geo.adjacency_sorted(lpos, lsup, order, slots=slots_s, counts=counts_s, prow=prow)
        evaluate(form, scratch)

    def step_rows(form):
        geo.sort(lpos, order=order, scratch=sort_scratch)
        geo.adjacency(lpos, lsup, pad=False, slots=slots, counts=counts, ind=ind, order=order)
        evaluate_rows(form)

    def timed(form, steps, warmup, scratch=None, fn=None):
        fn = fn or (lambda: step(form, scratch))
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    if args.probe:
        step("mixed")
        step("mixed")
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)                      # a plain copy of known size: 512 MiB read, 512 MiB written
        torch.cuda.synchronize()
        print(json.dumps({"probe": "star", "N": N,
                          "fused_pass_algorithmic_bytes": N * opmodel.bytes_per_particle_2d(len(work.kernels)),
                          "copy_bytes": src.numel() * 8}), flush=True)
        return

    # the first section a process times on the GPU runs slower (measured: 0.44 against 0.40 ms for the same step
    # timed later), so the row layout (adjacency and records at the rows; the same values as the headline's layout,
    # bitwise, checked below) is timed once before the headline, discarded, and again after it
    for f in ("mixed", "fp64"):
        timed(f, args.steps, args.warmup, fn=lambda f=f: step_rows(f))
    res_rows = {k: outs[k].clone() for k in work.kernels}
    # ---- headline: mixed form, cell sort + adjacency rebuild + evaluate
    t_mixed = timed("mixed", args.steps, args.warmup)
    log(f"mixed: {t_mixed * 1e3:.3f} ms/step")
    res_mixed = {k: outs[k].clone() for k in work.kernels}
    t_fp64 = timed("fp64", args.steps, args.warmup)
    log(f"fp64: {t_fp64 * 1e3:.3f} ms/step")
    res_fp64 = {k: outs[k].clone() for k in work.kernels}
    assert all(torch.equal(res_rows[k], res_fp64[k]) for k in work.kernels)
    t_rows = {f: timed(f, args.steps, args.warmup, fn=lambda f=f: step_rows(f)) for f in ("mixed", "fp64")}
    log("row layout ms/step:", {f: v * 1e3 for f, v in t_rows.items()})
    step_rows("mixed")
    assert all(torch.equal(outs[k], res_mixed[k]) for k in work.kernels)
    nbytes = int(L.load().wsb_moments2d_scratch_bytes(N, K))
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    t_plain = timed("plain", args.plain_steps, 1, scratch)
    log(f"plain: {t_plain * 1e3:.3f} ms/step")
    res_plain = {k: outs[k].clone() for k in work.kernels}
    del scratch

    # ---- error of the mixed form (and the plain path) against the FP64 closed form, signal-relative per channel
    err = {}
    for k in work.kernels:
        ref = res_fp64[k]
        scale = ref.abs().amax(0).clamp(min=1e-300)
        err[k] = {"mixed": float(((res_mixed[k] - ref).abs().amax(0) / scale).max()),
                  "plain": float(((res_plain[k] - ref).abs().amax(0) / scale).max())}
    log("signal-relative error vs fp64:", err)

    # ---- evaluate-only timing (events on the launch stream): the fused pass, and each kernel's own launch beside it
    geo.sort(lpos, order=order, scratch=sort_scratch)
    geo.adjacency(lpos, lsup, pad=False, slots=slots, counts=counts, ind=ind, order=order)
    geo.count_order(counts, base=order, order=corder, nwork=nwork, scratch=corder_scratch)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    reps = max(5, args.steps)
    geo.adjacency_sorted(lpos, lsup, order, slots=slots_s, counts=counts_s, prow=prow)
    torch.cuda.synchronize()
    t_multi, t_multi_count, t_multi_rows = {}, {}, {}
    for form in ("mixed", "fp64"):
        for fn, dst in ((evaluate, t_multi), (evaluate_by_count, t_multi_count), (evaluate_rows, t_multi_rows)):
            for _ in range(2):
                fn(form)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn(form)
            e1.record(stream)
            torch.cuda.synchronize()
            dst[form] = e0.elapsed_time(e1) / 1e3 / reps
    per_kernel = {}
    for form in ("mixed", "fp64"):
        for k in work.kernels:
            for _ in range(2):
                geo.moments(k, slots, counts, ind, lpos, lsup, form=form, out=outs[k])
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                geo.moments(k, slots, counts, ind, lpos, lsup, form=form, out=outs[k])
            e1.record(stream)
            torch.cuda.synchronize()
            per_kernel[(form, k)] = e0.elapsed_time(e1) / 1e3 / reps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        geo.sort(lpos, order=order, scratch=sort_scratch)
    e1.record(stream)
    torch.cuda.synchronize()
    t_sort = e0.elapsed_time(e1) / 1e3 / args.steps
    e0.record(stream)
    for _ in range(args.steps):
        geo.adjacency_sorted(lpos, lsup, order, slots=slots_s, counts=counts_s, prow=prow)
    e1.record(stream)
    torch.cuda.synchronize()
    t_adj = e0.elapsed_time(e1) / 1e3 / args.steps
    e0.record(stream)
    for _ in range(args.steps):
        geo.adjacency(lpos, lsup, pad=False, slots=slots, counts=counts, ind=ind, order=order)
    e1.record(stream)
    torch.cuda.synchronize()
    t_adj_rows = e0.elapsed_time(e1) / 1e3 / args.steps
    e0.record(stream)
    for _ in range(args.steps):
        geo.adjacency(lpos, lsup, pad=False, slots=slots, counts=counts, ind=ind)
    e1.record(stream)
    torch.cuda.synchronize()
    t_adj_unsorted = e0.elapsed_time(e1) / 1e3 / args.steps
    geo.adjacency(lpos, lsup, pad=False, slots=slots, counts=counts, ind=ind, order=order)
    e0.record(stream)
    for _ in range(args.steps):
        geo.count_order(counts, base=order, order=corder, nwork=nwork, scratch=corder_scratch)
    e1.record(stream)
    torch.cuda.synchronize()
    t_corder = e0.elapsed_time(e1) / 1e3 / args.steps
    # signed distance of every particle (70 % of them farther than one support from the wall: the edge tree)
    sd = geo.signed_distance(lpos, work.h)
    e0.record(stream)
    for _ in range(args.steps):
        sd = geo.signed_distance(lpos, work.h)
    e1.record(stream)
    torch.cuda.synchronize()
    t_sd = e0.elapsed_time(e1) / 1e3 / args.steps
    log("signed distance", t_sd)
    log("fused-pass seconds:", t_multi, "in count order:", t_multi_count, "single-kernel launch seconds:", {f"{a}/{b}": v for (a, b), v in per_kernel.items()},
        "adjacency", t_adj, "sort", t_sort,
        "adjacency in input order", t_adj_unsorted, "count order", t_corder)

    # ---- roofline of the dominant kernel: the mixed-form fused pass over the four kernels
    npairs, active = active_radius_counts(torch, rep.vertices.to(dev), rep.edges.to(dev).long(), slots, counts, lpos,
                                          lsup, K, work.kernels)
    flops = opmodel.multi_flops(work.kernels, npairs, active, N)
    bytes_alg = N * opmodel.bytes_per_particle_2d(len(work.kernels))
    t_dom = t_multi["mixed"]
    t_roof = max(bytes_alg / PEAK_HBM, flops / PEAK_FP64)
    compute_bound = flops / PEAK_FP64 >= bytes_alg / PEAK_HBM
    kname = "rows12_multi_kernel<15, 2, false>"
    roof = {"bound": "mfma" if compute_bound else "hbm",
            "pipe": "FP64 VALU (no MFMA: transcendental gather-reduce; the MI355X FP64 vector peak equals its FP64 matrix peak)",
            "kernel": kname + " (cubic, w2, w4, w6 in one pass, mixed form)",
            "achieved": (flops / t_dom / 1e12) if compute_bound else (bytes_alg / t_dom / 1e9),
            "peak": PEAK_FP64 / 1e12 if compute_bound else PEAK_HBM / 1e9,
            "unit": "TFLOP/s" if compute_bound else "GB/s",
            "frac": t_roof / t_dom, "traffic": _traffic(kname), "valu": _valu(kname),
            "algorithmic_flops": flops, "algorithmic_bytes": bytes_alg, "pairs": npairs,
            "active_chords_by_radius": {str(r): a for r, a in sorted(active.items())},
            "launch_seconds": t_dom, "hbm_GBps_algorithmic": bytes_alg / t_dom / 1e9}

    evals = 12 * len(work.kernels) * N
    value = evals / t_mixed
    line = {
        "metric": METRIC, "value": value, "unit": "particle-moment evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_mixed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 (mixed form: FP64 arithmetic, FP32 hardware angle)", "data": "synthetic",
        "config": {"workload": "star polygon 4096 edges, 1,048,576 particles (30% within one support), h=0.004, "
                               "kernels cubic/w2/w4/w6; step = cell sort + adjacency rebuild (sorted-position layout) + "
                               "twelve moments x 4 kernels",
                   "N": N, "edges": len(loop), "K": K, "pairs": npairs},
        "forms": {"mixed": {"evals_per_s": evals / t_mixed, "ms_per_step": t_mixed * 1e3},
                  "fp64": {"evals_per_s": evals / t_fp64, "ms_per_step": t_fp64 * 1e3},
                  "plain_fp64": {"evals_per_s": evals / t_plain, "ms_per_step": t_plain * 1e3}},
        "speedup_mixed_vs_plain": t_plain / t_mixed,
        "evaluate_only_evals_per_s": {"mixed": evals / t_multi["mixed"], "fp64": evals / t_multi["fp64"]},
        "evaluate_only_count_order_evals_per_s": {"mixed": evals / t_multi_count["mixed"],
                                                  "fp64": evals / t_multi_count["fp64"],
                                                  "note": "rows with slots grouped by count (wsb_count_order2d, "
                                                          "count_order_ms apart), indicator terms stored by a separate "
                                                          "contiguous pass; pays off when one adjacency is evaluated "
                                                          "more than once"},
        "evaluate_only_single_kernel_launches_evals_per_s": {
            "mixed": evals / sum(per_kernel[("mixed", k)] for k in work.kernels),
            "fp64": evals / sum(per_kernel[("fp64", k)] for k in work.kernels)},
        "row_layout": {"ms_per_step": {f: v * 1e3 for f, v in t_rows.items()},
                       "evaluate_only_evals_per_s": {f: evals / t_multi_rows[f] for f in ("mixed", "fp64")},
                       "adjacency_ms": t_adj_rows * 1e3,
                       "note": "adjacency and row records at the rows (wsb_adjacency2d, wsb_moments2d_multi): the "
                               "same values bitwise; the headline step keeps them at the sorted positions "
                               "(wsb_adjacency2d_sorted, wsb_moments2d_multi_sorted)"},
        "adjacency_ms": t_adj * 1e3, "cell_sort_ms": t_sort * 1e3, "adjacency_input_order_ms": t_adj_unsorted * 1e3,
        "count_order_ms": t_corder * 1e3,
        "signed_distance_ms": t_sd * 1e3,
        "max_signal_relative_error_vs_fp64": err,
        "served_latency": None,
        "roofline": roof,
    }
    line["precision_mp_wedges"] = precision_mp(torch, dev)
    log("per-moment error against mpmath:", {f: max(line["precision_mp_wedges"][f]["signal_relative"])
                                             for f in ("plain", "fp64", "mixed")})
    if not args.no_3d:
        line["mesh3d"] = mesh3d_measure(torch, dev, args.steps, args.warmup)
        log("3-D:", line["mesh3d"])
    if not args.no_disks:
        line["disks"] = disks_measure(torch, dev, args.steps, args.warmup)
        log("disks:", line["disks"])
    if not args.no_serve:
        log("served workload ...")
        # headline: the simulations' particle arrays are resident on this GPU (device-tensor requests and results)
        sat = 1e9                          # every request arrives at t = 0: saturation throughput
        # the first server of a process pays one-time costs its own warm pass does not absorb (its timed pass ran
        # at ~2/3 of the rate of every later server); a long-running server has paid them, so one whole served
        # measurement runs first, reported apart and kept out of the ratios
        cold = serve_measure("mixed", sat, device_requests=True)
        log("served mixed, saturation, first server of the process:", cold)
        mixed_sat = serve_measure("mixed", sat, device_requests=True)
        log("served mixed, saturation:", mixed_sat)
        plain_sat = serve_measure("plain", sat, device_requests=True)
        log("served plain FP64, saturation:", plain_sat)
        fp64_sat = serve_measure("fp64", sat, device_requests=True)
        log("served FP64 closed form, saturation:", fp64_sat)
        mixed_q = serve_measure("mixed", args.qps, device_requests=True)
        log("served mixed at", args.qps, "QPS:", mixed_q)
        plain_q = serve_measure("plain", args.qps, device_requests=True)
        log("served plain FP64 at", args.qps, "QPS:", plain_q)
        # host-array requests: every request and result crosses PCIe (pinned staging)
        mixed_host = serve_measure("mixed", args.qps)
        log("served mixed, host requests:", mixed_host)
        mixed_host_sat = serve_measure("mixed", sat)
        log("served mixed, host requests, saturation:", mixed_host_sat)
        graph_q = serve_measure("mixed", args.qps, graphs=True, batched=False, steps=2)
        log("served mixed, per-scene captured graphs, host requests:", graph_q)
        # captured cross-scene batches: the same workloads as the eager headline, one graph launch per batch
        mixed_sat_g = serve_measure("mixed", sat, device_requests=True, graphs=True)
        log("served mixed, saturation, captured cross-scene batches:", mixed_sat_g)
        mixed_q_g = serve_measure("mixed", args.qps, device_requests=True, graphs=True)
        log("served mixed at", args.qps, "QPS, captured cross-scene batches:", mixed_q_g)
        line["served_latency"] = {
            "p50_ms": mixed_q.get("p50_ms"), "p99_ms": mixed_q.get("p99_ms"), "qps": args.qps,
            "saturation_evals_per_s": {"mixed": mixed_sat["evals_per_s"], "plain_fp64": plain_sat["evals_per_s"],
                                       "fp64": fp64_sat["evals_per_s"]},
            "mixed_over_plain_fp64_at_saturation": mixed_sat["evals_per_s"] / plain_sat["evals_per_s"],
            "mixed": mixed_q, "plain_fp64": plain_q, "mixed_saturation": mixed_sat, "plain_fp64_saturation": plain_sat,
            "fp64_saturation": fp64_sat, "mixed_host_requests_pcie": mixed_host,
            "mixed_host_requests_pcie_saturation": mixed_host_sat, "mixed_per_scene_graphs_host_requests": graph_q,
            "mixed_saturation_first_server_of_process": cold,
            "mixed_saturation_captured": mixed_sat_g, "mixed_captured": mixed_q_g,
            "host_launch_seconds": {"eager_saturation": mixed_sat["host_seconds"]["launch"],
                                    "captured_saturation": mixed_sat_g["host_seconds"]["launch"],
                                    "eager_at_qps": mixed_q["host_seconds"]["launch"],
                                    "captured_at_qps": mixed_q_g["host_seconds"]["launch"]},
            "plain_denominator_note": "the served plain FP64 form is the row-parallel plan interpreter (one thread per "
                                      "row summing its pairs in slot order; batch2d.hip), and its 3-D scenes run the "
                                      "FP64 closed form; the per-pair plain path of the headline (plain_pairs_kernel) "
                                      "is timed in forms.plain_fp64",
            "workload": "256 scenes: 192 2-D star polygons (16-8192 edges) and 64 3-D closed meshes (box around an "
                        "icosphere cavity of 5,120 or 20,480 faces, or a box subdivided into 1,452-76,800 faces), "
                        "1e3-2e5 particles each, ~40 neighbours, w2/w4/w6, one jittered timestep per scene, Poisson "
                        "arrivals; 2-D requests in cross-scene batched launches, 3-D requests batched per scene, 4 "
                        "streams; 12 moments per 2-D row, 22 per 3-D row; a plain-form server runs its 3-D scenes in "
                        "the FP64 closed form (the 3-D path has no separate plan interpreter)"}
    if not args.no_cpu:
        log("timing the CPU oracle ...")
        line["cpu_baseline"] = cpu_baseline(work)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
