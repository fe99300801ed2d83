"""wsb_moments3d_static (no allocation: the heavy-row list in the library's own device array) against wsb_moments3d
with a caller-supplied scratch buffer: bitwise equal on two calls in a row, the second with fewer particles (the
shared array keeps the first call's contents; its counter must start again from zero), both forms; sizes beyond the
array's capacity are rejected."""
import numpy as np
import pytest
import torch

from warpsphboundaries_amd import _lib as L
from warpsphboundaries_amd import mesh3d as MM
from warpsphboundaries_amd.kernels import KERNEL_ID

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _scene():
    # a fine sphere inside a box with a support of ~10 face sizes: many particles have more than 64 faces
    V, F = MM.box_surface((0, 0, 0), (1, 1, 1))
    Vs, Fs = MM.icosphere(4, 0.25, (0.5, 0.5, 0.5))
    V, F = np.vstack([V, Vs]), np.vstack([F, (Fs + 8)[:, ::-1]])
    rng = np.random.default_rng(5)
    d = rng.normal(size=(1400, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    P = np.vstack([0.5 + d * (0.25 + rng.uniform(0.0, 0.1, 1400))[:, None], rng.uniform(0, 1, (600, 3))])
    return V, F, P[rng.permutation(len(P))], 0.12


def _static(m, adj, Q, H, kernel, form):
    n = Q.shape[0]
    out = torch.full((n, 22), float("nan"), dtype=torch.float64, device=DEV)
    L.check(L.load().wsb_moments3d_static(m._h, KERNEL_ID[kernel], L.form_id(form), L.ptr(adj.offsets),
                                          L.ptr(adj.faces), adj.faces.numel(), L.ptr(adj.chi), L.ptr(Q), L.ptr(H), n,
                                          int(L.OUT_PHYSICAL), L.ptr(out), L.stream_ptr(Q.device)))
    return out


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_static_moments_equal_the_scratch_entry_point(form):
    V, F, P, h = _scene()
    m = MM.Mesh3D(V, F, h, device=DEV)
    for rows in (len(P), len(P) // 3):                       # two calls in a row, the second smaller
        Q = torch.tensor(P[:rows], device=DEV)
        H = torch.full((rows,), h, dtype=torch.float64, device=DEV)
        adj = m.adjacency(Q, H)
        counts = np.diff(adj.offsets.cpu().numpy())
        assert (counts > 64).sum() > 20
        ref = m.moments(adj, Q, H, "w4", form)
        got = _static(m, adj, Q, H, "w4", form)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (rows, form)


def test_static_moments_reject_too_many_rows():
    V, F, P, h = _scene()
    m = MM.Mesh3D(V, F, h, device=DEV)
    big = int(L.load().wsb_moments3d_static_max_rows()) + 1
    assert big - 1 == 1 << 21
    rc = L.load().wsb_moments3d_static(m._h, KERNEL_ID["w4"], 1, None, None, 0, None, None, None, big, 0, None, None)
    assert rc == 1 and "at most" in L.last_error()
