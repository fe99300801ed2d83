"""Exact plan compiler: kernel blocks -> edge terms of the twelve surface channels.

Every channel of a polygon is a sum over its edges of line integrals of polynomials in r = |y| times powers of
the edge-line coordinate, plus the indicator of the point times a whole-support constant.  The reduction is the
compact-potential recursion of the reference (``edge/np2d.py:46-101``; derivation in the reference's
``docs/derivations/moments-recursion.md``): for a block profile P(r) = sum_n c_n r^n, 1[r <= R],

    Phi[P](r) = -int_r^R t P(t) dt        (a polynomial that vanishes at r = R)

and  y^alpha P = div(e_i y^beta Phi) - beta_i y^(beta - e_i) Phi  with alpha = beta + e_i, i = 0 if alpha_0 > 0
else 1; the recursion ends at alpha = 0 with the radial potential y Phi / r^2 (the "value" term whose point
singularity at y = 0 is the indicator).  Everything is exact (``Fraction``); floats are produced only when a plan
is flattened for a device.

Two outputs:

* ``compile_moment(kernel, alpha)`` -> (edges, values) exactly as the recursion produces them;
* ``surface_plan(kernel)`` -> the flattened twelve-channel term table of the edge-local (truncated-potential)
  channels that the reference evaluates per (query, edge) pair (``edge/warpbc.py:578-638`` for channels 0-8,
  ``:970-991`` for 9-11).  The plain FP64 device kernel interprets this table term by term;
* ``block_profiles(kernel)`` -> per block the four profile polynomials (P, P/(n+2), Phi, Phi/(n+2)) and the two
  indicator weights that the optimized kernels use with shared primitive tables.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from fractions import Fraction
from functools import lru_cache
from typing import Dict, List, Tuple

import numpy as np

from .kernels import kernel as get_kernel

F = Fraction

# device channel order (units of h):
#   0 lam  1 m10  2 m01  3 Gx  4 Gy  5 g1xx  6 g1xy  7 g1yx  8 g1yy  9 m20  10 m11  11 m02
NCH = 12
CHANNEL_NAMES = ("lam", "m10", "m01", "Gx", "Gy", "g1xx", "g1xy", "g1yx", "g1yy", "m20", "m11", "m02")
# report / API order m00 m10 m01 m20 m11 m02 Gx Gy g1xx g1xy g1yx g1yy -> device channel
REPORT_ORDER = (0, 1, 2, 9, 10, 11, 3, 4, 5, 6, 7, 8)
REPORT_NAMES = ("m00", "m10", "m01", "m20", "m11", "m02", "Gx", "Gy", "g1xx", "g1xy", "g1yx", "g1yy")


def phi(P: Dict[int, Fraction], R) -> Dict[int, Fraction]:
    """Phi[P](r) = -int_r^R t P(t) dt as {power: coefficient} (exact)."""
    R = F(R)
    out: Dict[int, Fraction] = {}
    const = F(0)
    for n, c in P.items():
        out[n + 2] = out.get(n + 2, F(0)) + c / (n + 2)
        const -= c * R ** (n + 2) / (n + 2)
    out[0] = out.get(0, F(0)) + const
    return {n: c for n, c in sorted(out.items()) if c != 0}


def mass_R(P: Dict[int, Fraction], R) -> Fraction:
    """m_R[P] = int_0^R t P(t) dt = sum c_n R^(n+2)/(n+2): the weight of the indicator (x 2 pi) and of the angle term."""
    R = F(R)
    return sum((c * R ** (n + 2) / (n + 2) for n, c in P.items()), F(0))


def _recurse(P, R, alpha, mult, edges, values):
    a, b = alpha
    if a + b == 0:
        d = values.setdefault(R, {})
        for n, c in P.items():
            d[n] = d.get(n, F(0)) + mult * c
        return
    i = 0 if a > 0 else 1
    beta = (a - 1, b) if i == 0 else (a, b - 1)
    Ph = phi(P, R)
    d = edges.setdefault((i, beta, R), {})
    for n, c in Ph.items():
        d[n] = d.get(n, F(0)) + mult * c
    bi = beta[i]
    if bi >= 1:
        low = (beta[0] - (1 if i == 0 else 0), beta[1] - (1 if i == 1 else 0))
        _recurse(Ph, R, low, -mult * bi, edges, values)


@lru_cache(maxsize=None)
def compile_moment(kname: str, alpha: Tuple[int, int]):
    """exact plan of m_alpha:  edges {(i, beta, R): {n: c}}  (n_i int_chord y^beta sum c_n r^n ds)
    and values {R: {n: c}}  (z int_chord sum c_n/(n+2) r^n ds - m_R dangle, plus 2 m_R ind); pi-free."""
    k = get_kernel(kname)
    edges: Dict = {}
    values: Dict = {}
    for blk in k.blocks:
        _recurse(blk.profile(), blk.R, (int(alpha[0]), int(alpha[1])), F(1), edges, values)
    edges = {key: {n: c for n, c in sorted(d.items()) if c != 0} for key, d in edges.items()}
    values = {R: {n: c for n, c in sorted(d.items()) if c != 0} for R, d in values.items()}
    return ({k_: v for k_, v in edges.items() if v}, {R: v for R, v in values.items() if v})


# ----------------------------------------------------------------------------------------------------------- flat plan
@dataclass
class SurfacePlan:
    """flattened edge-local twelve-channel plan of one kernel (float64 constants, pi-free)."""
    kernel: str
    radii: List[float] = field(default_factory=list)
    cn: List[int] = field(default_factory=list)          # power of r of every coefficient
    cc: List[float] = field(default_factory=list)        # coefficient
    # edge terms: (ch, i, a, b, R index, c0, c1, scale)   -> scale * n_i int_chord y^(a,b) sum cc r^cn ds
    E: List[Tuple[int, int, int, int, int, int, int]] = field(default_factory=list)
    # value terms: (ch, R index, c0, c1, mR)               -> z int_chord sum cc/(cn+2) r^cn ds - mR dangle
    V: List[Tuple[int, int, int, int, float]] = field(default_factory=list)

    def _r(self, R):
        f = float(R)
        if f not in self.radii:
            self.radii.append(f)
        return self.radii.index(f)

    def _coef(self, P, scale):
        c0 = len(self.cc)
        for n, c in sorted(P.items()):
            self.cn.append(int(n))
            self.cc.append(float(c * scale))
        return c0, len(self.cc)

    def edge(self, ch, i, beta, R, P, scale=F(1)):
        c0, c1 = self._coef(P, scale)
        self.E.append((ch, i, beta[0], beta[1], self._r(R), c0, c1))

    def value(self, ch, R, P, scale=F(1)):
        c0, c1 = self._coef(P, scale)
        self.V.append((ch, self._r(R), c0, c1, float(mass_R(P, R) * scale)))

    def moment(self, ch, alpha, scale=F(1)):
        edges, values = compile_moment(self.kernel, tuple(alpha))
        for (i, beta, R), P in edges.items():
            self.edge(ch, i, beta, R, P, scale)
        for R, P in values.items():
            self.value(ch, R, P, scale)

    # ---- arrays for the device
    def arrays(self):
        Ea = np.array(self.E or [(0,) * 7], dtype=np.int32).reshape(-1, 7)
        Va = np.array([v[:4] for v in self.V] or [(0,) * 4], dtype=np.int32).reshape(-1, 4)
        mR = np.array([v[4] for v in self.V] or [0.0], dtype=np.float64)
        return dict(radii=np.array(self.radii, dtype=np.float64), cn=np.array(self.cn or [0], dtype=np.int32),
                    cc=np.array(self.cc or [0.0], dtype=np.float64), E=np.ascontiguousarray(Ea),
                    V=np.ascontiguousarray(Va), mR=mR, nE=len(self.E), nV=len(self.V))


def _profile(kname):
    return {blk.R: blk.profile() for blk in get_kernel(kname).blocks}


@lru_cache(maxsize=None)
def surface_plan(kname: str, channels: Tuple[int, ...] = tuple(range(NCH))) -> SurfacePlan:
    """the edge-local terms of the twelve channels (truncated potentials only: a single open edge has no
    closed polygon for an inner potential's constant to drop out of; reference ``edge/warpbc.py:119-136``,
    ``:608-633``, ``:970-991``)."""
    pl = SurfacePlan(kname)
    want = set(channels)
    alphas = {0: (0, 0), 1: (1, 0), 2: (0, 1), 9: (2, 0), 10: (1, 1), 11: (0, 2)}
    for ch in (0, 1, 2):
        if ch in want:
            pl.moment(ch, alphas[ch])
    prof = _profile(kname)
    # g_alpha[j] = int y^alpha d_j W = -sum_e n_j int_chord y^alpha P ds + alpha_j m_(alpha - e_j)
    for k, al in enumerate(((0, 0), (1, 0), (0, 1))):
        for j in (0, 1):
            ch = 3 + 2 * k + j
            if ch not in want:
                continue
            for R, P in prof.items():
                pl.edge(ch, j, al, R, P, F(-1))
            if al[j]:
                pl.moment(ch, (al[0] - (j == 0), al[1] - (j == 1)), F(al[j]))
    for ch in (9, 10, 11):
        if ch in want:
            pl.moment(ch, alphas[ch])
    return pl


@lru_cache(maxsize=None)
def general_plan(kname: str, alpha: Tuple[int, int], grad: bool = False) -> SurfacePlan:
    """one-channel plan (channel 0) of m_alpha, or with ``grad`` the two channels (0, 1) of
    grad_x m_alpha = -sum_e n_e int_chord y^alpha W ds  (reference ``edge/core.py:150-166``)."""
    pl = SurfacePlan(kname)
    if grad:
        for j in (0, 1):
            for R, P in _profile(kname).items():
                pl.edge(j, j, tuple(alpha), R, P, F(-1))
    else:
        pl.moment(0, tuple(alpha))
    return pl


def indicator_weights(kname: str, alpha=(0, 0)) -> Fraction:
    """the exact whole-support constant that multiplies the indicator for m_alpha: sum over value terms of 2 m_R."""
    _, values = compile_moment(kname, tuple(alpha))
    return sum((2 * mass_R(P, R) for R, P in values.items()), F(0))


# ----------------------------------------------------------------------------------------------------------- profiles
@dataclass(frozen=True)
class BlockProfile:
    """one kernel block in the form the optimized kernels consume (pi-free, exact)."""
    R: Fraction
    P: Tuple[Fraction, ...]          # P(r) = sum P[n] r^n
    Phi: Tuple[Fraction, ...]        # Phi[P](r)
    mR_P: Fraction                   # m_R[P]  (angle weight of lam)
    mR_Phi: Fraction                 # m_R[Phi] (angle weight of the radial term of m20 / m02)


def _dense(d: Dict[int, Fraction]):
    if not d:
        return (F(0),)
    out = [F(0)] * (max(d) + 1)
    for n, c in d.items():
        out[n] = c
    return tuple(out)


@lru_cache(maxsize=None)
def block_profiles(kname: str) -> Tuple[BlockProfile, ...]:
    out = []
    for blk in get_kernel(kname).blocks:
        P = blk.profile()
        Ph = phi(P, blk.R)
        out.append(BlockProfile(blk.R, _dense(P), _dense(Ph), mass_R(P, blk.R), mass_R(Ph, blk.R)))
    return tuple(out)


def max_power(kname: str) -> int:
    """largest power of r any profile of the kernel carries (the I_m table length - 1)."""
    return max(len(b.Phi) - 1 for b in block_profiles(kname))
