"""Edge evaluator API (reference ``edge/warpbc.py:919-991`` and the batched polygon signatures of
``edge/np2d.py:501-565`` / ``edge/warp2d.py:424-475``), computed by the HIP library.

    edge_channels(pair_q, pair_e, positions, supports, vertices, edges, kernel, ...)   -> c[P, 9]  (units of h)
    second_channels(...)                                                               -> c[P, 3]  (m20, m11, m02)
    indicator_vector(kernel)                                                           -> u[9]
    channels12(...)                                                                    -> c[P, 12] (device order)
    value / gradient / moment / moment_gradient(verts[N,K,2], x[N,2], kernel, alpha, h) -> per polygon, physical units

The closed body total is  sum_edges c + indicator * u  (the reference's edge-local decomposition).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import precision
from .device import edge_channels12, plan_channels
from .kernels import disk_moment, plane_channels
from .scene import kernelName

F64 = torch.float64
_FORM_OF_STABLE = {None: "fp64", False: "fp64"}


def _dev(device):
    d = torch.device(device)
    if d.type != "cuda":
        raise RuntimeError("warpsphboundaries_amd computes on the MI355X (HIP) device: no CPU fallback")
    return d


def _t(x, dtype, device):
    return torch.as_tensor(x, dtype=dtype).to(device).contiguous()


def channels12(pair_q, pair_e, positions, supports, vertices, edges, kernel, device="cuda:0", form="fp64"):
    d = _dev(device)
    return edge_channels12(_t(pair_q, torch.int32, d), _t(pair_e, torch.int32, d), _t(positions, F64, d), _t(supports, F64, d),
                           _t(vertices, F64, d), _t(edges, torch.int32, d), kernelName(kernel), form)


def edge_channels(pair_q, pair_e, positions, supports, vertices, edges, kernel, device="cuda:0", plan=None, as_torch=True,
                  stable=None, channels=None, form=None):
    """edge-local channels c[P, 9] (lam, m10, m01, Gx, Gy, g1xx, g1xy, g1yx, g1yy) in units of h of every (query, edge)
    pair; edges counter-clockwise around the solid.  ``channels`` keeps a subset (the others are zero).  ``form``:
    'fp64' (closed form), 'mixed', or 'plain' (the plan interpreter); None: the process precision
    (``precision.DEFAULT_FORM``, from warpSPHCore_PRECISION)."""
    if isinstance(stable, tuple):
        raise NotImplementedError("the Chebyshev quadrature route is not built: the closed forms are FP64-accurate")
    c = channels12(pair_q, pair_e, positions, supports, vertices, edges, kernel, device, precision.default_form(form))[:, :9]
    if channels is not None:
        keep = torch.zeros(9, dtype=F64, device=c.device)
        keep[[int(k) for k in channels]] = 1.0
        c = c * keep
    return c if as_torch else c.cpu().numpy()


def second_channels(pair_q, pair_e, positions, supports, vertices, edges, kernel, device="cuda:0", stable=None, form=None):
    """edge-local m20, m11, m02 in units of h; add indicator * plane moment."""
    if isinstance(stable, tuple):
        raise NotImplementedError("the Chebyshev quadrature route is not built")
    return channels12(pair_q, pair_e, positions, supports, vertices, edges, kernel, device, precision.default_form(form))[:, 9:]


def indicator_vector(kernel):
    """u[9]: channel value of a body that contains the whole support: (1,0,0, 0,0, 1,0, 0,1)."""
    return np.array(plane_channels(kernelName(kernel))[:9])


# ================================================================================================ polygon element API
def _cross2(ax, ay, bx, by):
    # torch elementwise products are rounded separately: the same predicate as the device kernels
    return ax * by - ay * bx


def polygon_indicator(rel):
    """winding indicator of the origin for polygons ``rel`` [N, K, 2] (ccw, units of h): crossing number with
    half-open rules, on the boundary the angle sum of the non-collinear edges / 2 pi (reference
    scene/scene.py:333-371)."""
    a = rel
    c = torch.roll(rel, -1, dims=1)
    cross = _cross2(a[..., 0], a[..., 1], c[..., 0], c[..., 1])
    up = (a[..., 1] <= 0) & (0 < c[..., 1]) & (cross > 0)
    down = (c[..., 1] <= 0) & (0 < a[..., 1]) & (cross < 0)
    w = up.to(F64).sum(1) - down.to(F64).sum(1)
    on = (cross == 0) & ((a[..., 0] * c[..., 0] + a[..., 1] * c[..., 1]) <= 0)
    boundary = on.any(1)
    if bool(boundary.any()):
        ang = torch.where(cross == 0, torch.zeros_like(cross), torch.atan2(cross, a[..., 0] * c[..., 0] + a[..., 1] * c[..., 1]))
        w = torch.where(boundary, ang.sum(1) / (2 * np.pi), w)
    return w


def _prep(verts, x, h, device):
    d = _dev(device)
    V = torch.as_tensor(np.asarray(verts, dtype=np.float64)).to(d)
    X = torch.as_tensor(np.asarray(x, dtype=np.float64)).to(d)
    if V.ndim == 2:
        V, X = V[None], X[None]
    N, K = V.shape[0], V.shape[1]
    H = torch.as_tensor(np.broadcast_to(np.asarray(h, dtype=np.float64), (N,)).copy()).to(d)
    rel = (V - X[:, None, :]) / H[:, None, None]
    a2 = _cross2(rel[..., 0], rel[..., 1], torch.roll(rel, -1, 1)[..., 0], torch.roll(rel, -1, 1)[..., 1]).sum(1)
    flip = a2 < 0
    V = torch.where(flip[:, None, None], V.flip(1), V)
    rel = torch.where(flip[:, None, None], rel.flip(1), rel)
    verts_flat = V.reshape(N * K, 2).contiguous()
    k = torch.arange(K, device=d)
    edges = torch.stack([k, (k + 1) % K], 1)[None] + (torch.arange(N, device=d) * K)[:, None, None]
    pair_q = torch.arange(N, device=d, dtype=torch.int32).repeat_interleave(K)
    pair_e = torch.arange(N * K, device=d, dtype=torch.int32)
    return dict(N=N, K=K, V=verts_flat, E=edges.reshape(-1, 2).to(torch.int32).contiguous(), X=X.contiguous(), H=H, rel=rel,
                pq=pair_q, pe=pair_e)


def _sum(c, N, K):
    return c.view(N, K, -1).sum(1)


def value(verts, x, kernel, h=1, device="cuda:0"):
    return moment(verts, x, kernel, (0, 0), h, device)


def gradient(verts, x, kernel, h=1, device="cuda:0"):
    return moment_gradient(verts, x, kernel, (0, 0), h, device)


def moment(verts, x, kernel, alpha, h=1, device="cuda:0"):
    """m_alpha = int_T y^alpha W dA per polygon (y = x' - x), physical units (h^|alpha|)."""
    k = kernelName(kernel)
    a = (int(alpha[0]), int(alpha[1]))
    g = _prep(verts, x, h, device)
    twelve = {(0, 0): 0, (1, 0): 1, (0, 1): 2, (2, 0): 9, (1, 1): 10, (0, 2): 11}
    if a in twelve:
        c = edge_channels12(g["pq"], g["pe"], g["X"], g["H"], g["V"], g["E"], k)[:, twelve[a]]
    else:
        c = plan_channels(g["pq"], g["pe"], g["X"], g["H"], g["V"], g["E"], k, a)[:, 0]
    ind = polygon_indicator(g["rel"])
    tot = _sum(c[:, None], g["N"], g["K"])[:, 0] + ind * float(disk_moment(k, a))
    return tot * g["H"] ** (a[0] + a[1])


def moment_gradient(verts, x, kernel, alpha, h=1, device="cuda:0"):
    """grad_x m_alpha = -sum_e n_e int_chord y^alpha W ds per polygon -> [N, 2] (h^(|alpha|-1))."""
    k = kernelName(kernel)
    a = (int(alpha[0]), int(alpha[1]))
    g = _prep(verts, x, h, device)
    if a == (0, 0):
        c = edge_channels12(g["pq"], g["pe"], g["X"], g["H"], g["V"], g["E"], k)[:, 3:5]
    else:
        c = plan_channels(g["pq"], g["pe"], g["X"], g["H"], g["V"], g["E"], k, a, grad=True)[:, 0:2]
    return _sum(c, g["N"], g["K"]) * (g["H"] ** (a[0] + a[1] - 1))[:, None]
