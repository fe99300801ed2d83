"""Exact tables of the 3-D face-local closed form (closed oriented triangle surfaces).

For a solid V bounded by outward-oriented triangles f (unit normal n_f, plane offset z_f = n_f . (p_f - x)) and a
kernel block P(r) = sum_n c_n r^n 1[r <= R] (units of h, 3-D normalisation, times pi), every stored moment of
``baseline(3)[:22]`` (reference ``fluxbridge/contract.py:41-46``) splits into compact face terms, which vanish
for faces outside the support ball, plus the winding number chi = Omega / 4 pi times the whole-ball value:

    M0      = sum_f z_f S_f[psi_P - m_R r^-3]                    + 4 pi chi m_R
    M1_i    = sum_f n_i S_f[Phi]
    M2_ik   = sum_f n_i (z n_k S_f[Phi] + sum_e mu_e,k L_e[Phi2]) - delta_ik V_Phi          (i <= k)
    V_Phi   = sum_f z_f S_f[psi_Phi - m_RPhi r^-3]                + 4 pi chi m_RPhi
    G_j     = -sum_f n_j S_f[P]
    g1_dj   = -sum_f n_j (z n_d S_f[P] + sum_e mu_e,d L_e[Phi]) + delta_dj M0

with the compact potentials Phi = -int_r^R t P, Phi2 = -int_r^R t Phi, the volume potentials psi_g = sum g_m
r^m/(m+3), m_g = sum g_m R^(m+3)/(m+3), S_f[g] = int_f g(r) 1[r <= R] dA and L_e[g] = int_e g(r) 1[r <= R] ds.
This is the recursion of ``mesh3d/reference.py:219-262`` / ``fluxbridge/plan.py:38-113`` summed over the
profile and regrouped by radial function; the whole-ball vector is [1, 0, 0, 0, s, 0, 0, s, 0, s, G = 0, g1 = I].
"""
from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction
from functools import lru_cache
from typing import Dict, Tuple

from .kernels import NAMES, kernel as get_kernel

F = Fraction

# baseline(3)[:22]: moments in graded order, then kernel gradients of (0,0,0),(1,0,0),(0,1,0),(0,0,1) per axis
MOMENT_ALPHAS = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1),
                 (0, 0, 2))
GRADIENT_ALPHAS = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))
REQUESTS22 = tuple((a, ()) for a in MOMENT_ALPHAS) + tuple((a, (j,)) for a in GRADIENT_ALPHAS for j in range(3))
CHANNEL_NAMES3 = ("m000", "m100", "m010", "m001", "m200", "m110", "m101", "m020", "m011", "m002",
                  "Gx", "Gy", "Gz", "g1xx", "g1xy", "g1xz", "g1yx", "g1yy", "g1yz", "g1zx", "g1zy", "g1zz")
NCH3 = 22
# physical scaling exponent of h per channel: |alpha| - |derivative| (fluxbridge/evaluate.py:76-78)
H_POWER3 = tuple(sum(a) - len(d) for a, d in REQUESTS22)


def _poly_eval(coeffs: Dict[int, Fraction], x: Fraction) -> Fraction:
    return sum((c * x ** m for m, c in coeffs.items()), F(0))


def _prim(coeffs: Dict[int, Fraction], shift: int) -> Dict[int, Fraction]:
    """{m: g_m} -> {m + shift: g_m / (m + shift)}: the monomial map of int t^(shift-1) g(t) dt."""
    return {m + shift: c / (m + shift) for m, c in coeffs.items() if c}


def _potential(g: Dict[int, Fraction], R: Fraction) -> Dict[int, Fraction]:
    """-int_r^R t g(t) dt as {power: coeff}."""
    p = _prim(g, 2)
    out = dict(p)
    out[0] = out.get(0, F(0)) - _poly_eval(p, R)
    return {m: c for m, c in out.items() if c}


@dataclass(frozen=True)
class Block3:
    """one kernel block with the radial functions the face closed form needs (3-D normalised, times pi)."""
    R: Fraction
    P: Dict[int, Fraction]          # the block polynomial
    Phi: Dict[int, Fraction]        # -int_r^R t P
    Phi2: Dict[int, Fraction]       # -int_r^R t Phi
    psiP: Dict[int, Fraction]       # sum P_m r^m/(m+3)
    psiF: Dict[int, Fraction]       # sum Phi_m r^m/(m+3)
    mR: Fraction                    # sum P_m R^(m+3)/(m+3)
    mRF: Fraction                   # sum Phi_m R^(m+3)/(m+3)


@lru_cache(maxsize=None)
def blocks3(name: str) -> Tuple[Block3, ...]:
    k = get_kernel(name)
    ratio = k.c3_pi / k.c2_pi
    out = []
    for b in k.blocks:
        P = {n: c * ratio for n, c in enumerate(b.coeffs) if c}
        Phi = _potential(P, b.R)
        Phi2 = _potential(Phi, b.R)
        psiP = _prim(P, 3)
        psiF = _prim(Phi, 3)
        psiP = {m - 3: c for m, c in psiP.items()}
        psiF = {m - 3: c for m, c in psiF.items()}
        mR = _poly_eval(psiP, b.R) * b.R ** 3
        mRF = _poly_eval(psiF, b.R) * b.R ** 3
        out.append(Block3(b.R, P, Phi, Phi2, psiP, psiF, mR, mRF))
    return tuple(out)


def ball_second_moment(name: str) -> Fraction:
    """int_ball y_x^2 W dV at unit support = -4 sum_j m_RPhi,j (exact)."""
    return -4 * sum(b.mRF for b in blocks3(name))


def ball_mass(name: str) -> Fraction:
    return 4 * sum(b.mR for b in blocks3(name))


def whole_vector3(name: str):
    """22-channel value of a solid that covers the whole support (request order), as floats."""
    s = float(ball_second_moment(name))
    m = float(ball_mass(name))
    v = [m, 0.0, 0.0, 0.0, s, 0.0, 0.0, s, 0.0, s, 0.0, 0.0, 0.0]
    v += [m if d == j else 0.0 for d in range(3) for j in range(3)]
    return v


def max_power3(name: str) -> int:
    return max(max(b.Phi2) for b in blocks3(name))


def dense(coeffs: Dict[int, Fraction], lo: int, hi: int):
    return [coeffs.get(m, F(0)) for m in range(lo, hi + 1)]


__all__ = ["Block3", "blocks3", "whole_vector3", "ball_mass", "ball_second_moment", "REQUESTS22", "CHANNEL_NAMES3",
           "NCH3", "H_POWER3", "MOMENT_ALPHAS", "GRADIENT_ALPHAS", "NAMES", "max_power3", "dense"]
