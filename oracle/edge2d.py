"""Numpy restatement of the 2-D edge reductions per (query, edge) pair, in FP64 or the extended precisions of
``oracle.arith``.  TEST INFRASTRUCTURE (see oracle/__init__).

This follows the stage-1 structure of the reference -- monomial-by-monomial block recursion, not the flattened
plan the product uses -- so that the product's plan compiler and device kernels are checked against an
independent formulation:

* geometry of an edge relative to the query in units of h     : ``edge/geometry.py:126-145`` (z, s0, s1, n, t)
  with the rounded orientation product of ``edge/predicates.py:21-32`` (no fused multiply-add);
* chord clipping                                               : ``edge/geometry.py:166-177``
* primitives I_m, S_{j,m}, dangle                              : ``edge/primitives.py:19-94``
* block value / gradient / moment recursion / moment gradient  : ``edge/core.py:68-166``
* kernel level sums, 1/pi                                      : ``edge/core.py:196-212``

Everything is returned edge-local (no indicator): the total of a closed body is the sum over its edges plus
``indicator * plane_vector`` (``scene/scene.py:410-412``).
"""
from __future__ import annotations

from fractions import Fraction
from math import comb

import numpy as np

from . import arith as A
from .kernels import blocks as kernel_blocks, plane_vector_exact


class EdgeGeo:
    """per-pair edge geometry, arrays of shape (P,)."""

    def __init__(self, p, q):
        p = A.asarray(p)
        q = A.asarray(q)
        d = q - p
        ell = A.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
        ok = ell > 0
        safe = np.where(ok, ell, 1.0)
        self.t0, self.t1 = d[:, 0] / safe, d[:, 1] / safe
        self.n0, self.n1 = self.t1, -self.t0
        cross = p[:, 0] * q[:, 1] - p[:, 1] * q[:, 0]          # numpy rounds each product: no contraction
        self.z = cross / safe
        self.s0 = (d[:, 0] * p[:, 0] + d[:, 1] * p[:, 1]) / safe
        self.s1 = (d[:, 0] * q[:, 0] + d[:, 1] * q[:, 1]) / safe
        self.ok = ok
        self._chords = {}

    def chord(self, R):
        key = float(R)
        v = self._chords.get(key)
        if v is None:
            Rf = float(R)
            az = np.abs(self.z)
            L = A.sqrt(np.maximum((Rf - az) * (Rf + az), 0.0))
            lo = np.maximum(self.s0, -L)
            hi = np.minimum(self.s1, L)
            m = (az < Rf) & (lo < hi) & self.ok
            v = (np.where(m, lo, 0.0), np.where(m, hi, 0.0), m)
            self._chords[key] = v
        return v


def I_all(M, s, z):
    """[I_0 .. I_M](s): odd antiderivatives of r^m (edge/primitives.py:19-32)."""
    zz = z * z
    r2 = s * s + zz
    r = A.sqrt(r2)
    out = [s]
    if M >= 1:
        safe = np.where(z == 0, 1.0, np.abs(z))
        tail = np.where(z == 0, 0.0, zz * A.arcsinh(s / safe))
        out.append((s * r + tail) / 2)
    for m in range(2, M + 1):
        rpow = r2 ** (m // 2) if m % 2 == 0 else r * r2 ** ((m - 1) // 2)
        out.append((s * rpow + m * zz * out[m - 2]) / (m + 1))
    return out


def S(j, m, s, z):
    """int s^j r^m ds for m >= 0 via s^2 = r^2 - z^2 (edge/primitives.py:46-54)."""
    zz = z * z
    if j % 2 == 0:
        h = j // 2
        Is = I_all(m + 2 * h, s, z)
        return sum(comb(h, i) * (-zz) ** (h - i) * Is[m + 2 * i] for i in range(h + 1))
    h = (j - 1) // 2
    r2 = s * s + zz
    tot = 0.0
    for i in range(h + 1):
        mm = m + 2 * i
        rp = r2 ** ((mm + 2) // 2) if mm % 2 == 0 else A.sqrt(r2) * r2 ** ((mm + 1) // 2)
        tot = tot + comb(h, i) * (-zz) ** (h - i) * rp / (mm + 2)
    return tot


def dangle(z, lo, hi):
    """atan2(z (hi - lo), z^2 + lo hi), 0 at z == 0 (edge/primitives.py:86-94)."""
    safe = np.where(z == 0, 1.0, z)
    return np.where(z == 0, 0.0, A.arctan2(safe * (hi - lo), z * z + lo * hi))


def ypoly(g: EdgeGeo, beta):
    """coefficients of y^beta on the edge line, y = z n + s t (edge/core.py:42-53)."""
    a, b = beta
    c1 = [g.z * g.n0, g.t0]
    c2 = [g.z * g.n1, g.t1]
    out = [np.ones_like(g.z)]

    def mul(u, v):
        r = [0.0] * (len(u) + len(v) - 1)
        for i, x in enumerate(u):
            for j, y in enumerate(v):
                r[i + j] = r[i + j] + x * y
        return r
    for _ in range(a):
        out = mul(out, c1)
    for _ in range(b):
        out = mul(out, c2)
    return out


class _Ctx:
    def __init__(self, g: EdgeGeo):
        self.g = g
        self.cache = {}

    def sdiff(self, j, m, R):
        key = ("S", j, m, float(R))
        v = self.cache.get(key)
        if v is None:
            lo, hi, msk = self.g.chord(R)
            z = self.g.z
            v = np.where(msk, S(j, m, hi, z) - S(j, m, lo, z), 0.0)
            self.cache[key] = v
        return v

    # ---- blocks, edge-local parts (edge/core.py:68-166)
    def block_value(self, n, R):
        key = ("val", n, float(R))
        v = self.cache.get(key)
        if v is None:
            lo, hi, msk = self.g.chord(R)
            z = self.g.z
            Rf = float(R)
            In = self.sdiff(0, n, R)
            v = np.where(msk, (z * In - Rf ** (n + 2) * dangle(z, lo, hi)) / (n + 2), 0.0)
            self.cache[key] = v
        return v

    def block_grad(self, n, R):
        d = self.sdiff(0, n, R)
        return -self.g.n0 * d, -self.g.n1 * d

    def block_moment(self, alpha, n, R):
        key = ("mom", alpha, n, float(R))
        if key in self.cache:
            return self.cache[key]
        a, b = alpha
        if a + b == 0:
            val = self.block_value(n, R)
            self.cache[key] = val
            return val
        Rf = float(R)
        i = 0 if a > 0 else 1
        beta = (a - 1, b) if i == 0 else (a, b - 1)
        n2 = n + 2
        g = self.g
        ni = g.n0 if i == 0 else g.n1
        acc = 0.0
        for j, cj in enumerate(ypoly(g, beta)):
            acc = acc + cj * (self.sdiff(j, n2, R) - Rf ** n2 * self.sdiff(j, 0, R))
        t1 = ni * acc / n2
        bi = beta[i]
        t2 = 0.0
        if bi >= 1:
            bm = (beta[0] - (1 if i == 0 else 0), beta[1] - (1 if i == 1 else 0))
            t2 = bi * (self.block_moment(bm, n2, R) - Rf ** n2 * self.block_moment(bm, 0, R)) / n2
        val = t1 - t2
        self.cache[key] = val
        return val

    def block_moment_grad(self, alpha, n, R):
        acc = 0.0
        for j, cj in enumerate(ypoly(self.g, alpha)):
            acc = acc + cj * self.sdiff(j, n, R)
        return -self.g.n0 * acc, -self.g.n1 * acc


def _sum_blocks(kname, fn):
    tot = None
    for R, coeffs in kernel_blocks(kname):
        for n, c in enumerate(coeffs):
            if c == 0:
                continue
            v = fn(n, R)
            w = A.num(c) / A.pi()
            v = tuple(w * x for x in v) if isinstance(v, tuple) else w * v
            if tot is None:
                tot = v
            elif isinstance(v, tuple):
                tot = tuple(a + b for a, b in zip(tot, v))
            else:
                tot = tot + v
    return tot


def edge_channels12(p, q, kname):
    """[P, 12] edge-local channels in device order (lam, m10, m01, Gx, Gy, g1xx, g1xy, g1yx, g1yy, m20, m11, m02),
    units of h, for edges p -> q given relative to the query in units of h (solid on the left)."""
    g = EdgeGeo(p, q)
    c = _Ctx(g)
    lam = _sum_blocks(kname, lambda n, R: c.block_value(n, R))
    m10 = _sum_blocks(kname, lambda n, R: c.block_moment((1, 0), n, R))
    m01 = _sum_blocks(kname, lambda n, R: c.block_moment((0, 1), n, R))
    G = _sum_blocks(kname, lambda n, R: c.block_grad(n, R))
    gx = _sum_blocks(kname, lambda n, R: c.block_moment_grad((1, 0), n, R))
    gy = _sum_blocks(kname, lambda n, R: c.block_moment_grad((0, 1), n, R))
    m20 = _sum_blocks(kname, lambda n, R: c.block_moment((2, 0), n, R))
    m11 = _sum_blocks(kname, lambda n, R: c.block_moment((1, 1), n, R))
    m02 = _sum_blocks(kname, lambda n, R: c.block_moment((0, 2), n, R))
    # g1[d, j] = d_j m_(e_d) + delta_dj m00   (validation/adversarial.py:144-147)
    out = np.stack([lam, m10, m01, G[0], G[1], gx[0] + lam, gx[1], gy[0], gy[1] + lam, m20, m11, m02], axis=1)
    return np.where(g.ok[:, None], out, 0.0)


def moment_edges(p, q, kname, alpha):
    """edge-local m_alpha of arbitrary order (units of h) per pair."""
    c = _Ctx(EdgeGeo(p, q))
    a = (int(alpha[0]), int(alpha[1]))
    return _sum_blocks(kname, lambda n, R: c.block_moment(a, n, R))


def moment_grad_edges(p, q, kname, alpha):
    c = _Ctx(EdgeGeo(p, q))
    a = (int(alpha[0]), int(alpha[1]))
    gx, gy = _sum_blocks(kname, lambda n, R: c.block_moment_grad(a, n, R))
    return np.stack([gx, gy], 1)


def plane_vector(kname):
    """the twelve-channel whole-support vector in the working precision (exact rationals rounded once)."""
    return A.asarray([A.num(v) for v in plane_vector_exact(kname)]) if A.mode() != "mp" else \
        np.array([A.num(v) for v in plane_vector_exact(kname)], dtype=object)
