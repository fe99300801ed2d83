"""The continuous-batching server: results equal alone and in the mix (bitwise), captured-graph replays equal
eager launches on every replay, HBM admission, a poisoned request isolated from its batch."""
import numpy as np
import pytest
import torch

from warpsphboundaries_amd import workloads as WL
from warpsphboundaries_amd.serve import BoundaryServer

pytestmark = pytest.mark.gpu


def _scenes(k=6, seed=21):
    return WL.config5(n_scenes=k, seed=seed, n_min=500, n_max=20000, e_min=16, e_max=512)


def _serve(scenes, requests, **kw):
    srv = BoundaryServer(**kw)
    for s in scenes:
        srv.add_scene(s.name, s.loops, s.h)
    reqs = [srv.submit(name, pos, h, kern) for name, pos, h, kern in requests]
    srv.run()
    return srv, reqs


def _requests(scenes, steps=3):
    out = []
    for t in range(steps):
        for i, s in enumerate(scenes):
            out.append((s.name, WL.jitter(s.positions, s.h, 100 * t + i), s.h, s.kernel))
    return out


def test_results_equal_alone_and_in_the_mix():
    scenes = _scenes()
    reqs = _requests(scenes)
    _, mixed = _serve(scenes, reqs, streams=4)
    assert all(r.error is None and r.result is not None for r in mixed)
    for (name, pos, h, kern), m in zip(reqs, mixed):
        _, (alone,) = _serve([s for s in scenes if s.name == name], [(name, pos, h, kern)], streams=1)
        assert np.array_equal(alone.result, m.result), name


def test_graph_replays_equal_eager_launches():
    scenes = _scenes(3, seed=22)
    reqs = _requests(scenes, steps=4)
    _, eager = _serve(scenes, reqs, streams=2)
    srv, graphed = _serve(scenes, reqs, streams=2, graphs=True)
    for e, g in zip(eager, graphed):
        assert np.array_equal(e.result, g.result)
    assert any(len(s.graphs) for pool in srv.slots for s in pool.values())


def test_hbm_admission_queues_and_rejects():
    scenes = _scenes(2, seed=23)
    srv = BoundaryServer(hbm_budget=64 << 20, streams=4)
    for s in scenes:
        srv.add_scene(s.name, s.loops, s.h)
    big = np.random.default_rng(0).uniform(0, 1, (400000, 2))
    r = srv.submit(scenes[0].name, big, scenes[0].h, "w2")
    assert isinstance(r.error, MemoryError)
    reqs = [srv.submit(s.name, s.positions, s.h, s.kernel) for s in scenes for _ in range(3)]
    srv.run()
    assert all(q.result is not None for q in reqs)
    with pytest.raises(ValueError):
        srv.submit(scenes[0].name, scenes[0].positions, 2 * scenes[0].h, "w2")


def test_poisoned_request_is_isolated():
    scenes = _scenes(1, seed=24)
    s = scenes[0]
    good = s.positions
    bad = good.copy()
    bad[::7] = np.nan
    srv, (r_bad, r_good) = _serve(scenes, [(s.name, bad, s.h, s.kernel), (s.name, good, s.h, s.kernel)], streams=1)
    _, (alone,) = _serve(scenes, [(s.name, good, s.h, s.kernel)], streams=1)
    assert np.array_equal(r_good.result, alone.result)
    assert srv.launched_batches == 1          # both were in one batch


def test_cross_scene_batches_equal_per_scene_launches():
    scenes = _scenes(5, seed=25)
    box = WL.config1(n=3000, seed=26)                       # a two-representation scene (box + obstacle)
    scenes.append(WL.ServedScene("box+gon", box.loops, 0.02, "w4", box.positions))
    reqs = _requests(scenes, steps=2)
    srv_b, batched = _serve(scenes, reqs, streams=2, batched=True)
    srv_p, per_scene = _serve(scenes, reqs, streams=2, batched=False)
    for b, p in zip(batched, per_scene):
        assert b.error is None and p.error is None
        assert np.array_equal(b.result, p.result), b.scene
    assert srv_b.launched_batches < srv_p.launched_batches


def test_device_resident_requests_equal_host_requests():
    scenes = _scenes(4, seed=27)
    reqs = _requests(scenes, steps=2)
    _, host = _serve(scenes, reqs, streams=3)
    dev_reqs = [(n, torch.tensor(p, device="cuda:0"), h, k) for n, p, h, k in reqs]
    _, dev = _serve(scenes, dev_reqs, streams=3)
    for a, b in zip(host, dev):
        assert isinstance(b.result, torch.Tensor) and b.result.is_cuda
        assert np.array_equal(a.result, b.result.cpu().numpy())


def test_plain_form_batches_equal_per_scene_plain():
    scenes = _scenes(4, seed=28)
    reqs = _requests(scenes, steps=1)
    _, b = _serve(scenes, reqs, streams=2, form="plain", batched=True)
    _, p = _serve(scenes, reqs, streams=2, form="plain", batched=False)
    for x, y in zip(b, p):
        assert np.array_equal(x.result, y.result)
