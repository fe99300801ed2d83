"""Solid-sphere closed form lambda_sph(R, d) = int_ball W(|x' - x|) dV (3-D, unit support), exact and on the device.

For a query at signed distance d from the surface of a ball of radius R (d > 0 outside, d < 0 inside; reference
``curvbound/symbolic.py:166-178`` for the outside, ``mesh3d/radial.py:102-140`` / ``fluxbridge/radial.py`` for both
sides) the sphere of radius r about the query meets the ball in a cap of area  pi r (r - d)(2R + d - r)/(R + d)
for |d| < r < 2R + d, and lies wholly inside it for r < -d.  With the 3-D kernel blocks P_j (pi W as truncated
monomials of radius R_j) this gives, per block,

    lambda_j = 4 M3(a1) + [ -dM4 + 2 (R + d) dM3 - d (2R + d) dM2 ] / (R + d),
    M_k(r) = sum_n c_n r^(n+k) / (n+k),   dM_k = M_k(b) - M_k(a),
    a1 = min(R_j, max(0, -d)),  a = |d|,  b = min(R_j, 2R + d)   (the bracket only when a < b),

a rational function of (R, d) on each branch, evaluated here exactly (``Fraction``) and on the device in FP64.  At
the centre of the ball (R + d = 0) the bracket vanishes with the cap range.  For R -> infinity the bracket tends
to the half-space value without cancellation (d (2R + d)/(R + d) -> 2d), which is why no branch needs a series.
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from typing import Tuple

import numpy as np

from .kernels import NAMES, kernel as get_kernel

F = Fraction


@lru_cache(maxsize=None)
def blocks3d(name: str) -> Tuple[Tuple[Fraction, Tuple[Fraction, ...]], ...]:
    """3-D blocks: pi W = sum_n c_n r^n 1[r <= R_j] with the 3-D normalisation (c3/c2 times the 2-D blocks)."""
    k = get_kernel(name)
    s = k.c3_pi / k.c2_pi
    return tuple((b.R, tuple(c * s for c in b.coeffs)) for b in k.blocks)


def _M(coeffs, k, r):
    return sum((c * r ** (n + k) / (n + k) for n, c in enumerate(coeffs) if c), F(0))


def lambda_exact(name: str, R, d) -> Fraction:
    """exact lambda_sph(R, d) for rational R > 0, -R <= d (units of h)."""
    R, d = F(R), F(d)
    if R <= 0:
        raise ValueError("radius must be positive")
    if d < -R:
        raise ValueError("d < -R: the query would be outside on the far side")
    D = R + d
    tot = F(0)
    for Rj, c in blocks3d(name):
        a1 = min(Rj, max(F(0), -d))
        tot += 4 * _M(c, 3, a1)
        a, b = abs(d), min(Rj, 2 * R + d)
        if a < b and D > 0:
            dM2 = _M(c, 2, b) - _M(c, 2, a)
            dM3 = _M(c, 3, b) - _M(c, 3, a)
            dM4 = _M(c, 4, b) - _M(c, 4, a)
            tot += (-dM4 + 2 * D * dM3 - d * (2 * R + d) * dM2) / D
    return tot


def half_space_exact(name: str, d) -> Fraction:
    """lambda of the half space at signed distance d (the R -> infinity limit)."""
    d = F(d)
    tot = F(0)
    for Rj, c in blocks3d(name):
        a1 = min(Rj, max(F(0), -d))
        tot += 4 * _M(c, 3, a1)
        a, b = abs(d), Rj
        if a < b:
            # limit of the bracket: -0 + 2 dM3 - 2 d dM2
            tot += 2 * (_M(c, 3, b) - _M(c, 3, a)) - 2 * d * (_M(c, 2, b) - _M(c, 2, a))
    return tot


def device_tables():
    """per kernel: [(R_j, [c_n/(n+2)], [c_n/(n+3)], [c_n/(n+4)])] as floats, for the generated header."""
    out = {}
    for name in NAMES:
        rows = []
        for Rj, c in blocks3d(name):
            rows.append((float(Rj), [float(x / (n + 2)) for n, x in enumerate(c)], [float(x / (n + 3)) for n, x in enumerate(c)],
                         [float(x / (n + 4)) for n, x in enumerate(c)]))
        out[name] = rows
    return out


def closed_form(R, d, kernel: str):
    """lambda_sph on the device (FP64): R, d float64 tensors of equal shape, units of h."""
    import torch
    from . import _lib as L
    from .kernels import KERNEL_ID
    L.require_device(R, d)
    Rt = R.to(torch.float64).contiguous()
    dt = d.to(torch.float64).contiguous()
    out = torch.empty_like(Rt)
    n = Rt.numel()
    if n:
        L.check(L.load().wsb_sphere_closed_form(KERNEL_ID[kernel], L.ptr(Rt), L.ptr(dt), n, L.ptr(out), L.stream_ptr(Rt.device)))
    return out


def sample_pairs(n=65536, seed=3, h=1.0):
    """the sphere check of the 3-D configuration: R/h log-uniform in [0.05, 1e3], d/h uniform in [-1, 1]
    (clipped to d >= -R, the deepest point of the ball)."""
    g = np.random.Generator(np.random.PCG64(seed))
    R = 10.0 ** g.uniform(np.log10(0.05), 3.0, n)
    d = g.uniform(-1.0, 1.0, n)
    d = np.maximum(d, -R)
    return R * h, d * h
