"""Exact SPH kernel table as truncated monomials (the algebra every plan is compiled from).

A kernel is a compactly supported radial function of q = r/h on [0, 1].  In two dimensions

    W(r, h) = (1/pi) * (1/h^2) * sum_j sum_n c_{j,n} q^n 1[q <= R_j]

with exact rational coefficients (``Fraction``).  The shapes and the 2-D normalisation constants follow
``curvbound/kernels.py:1-50`` of the reference (Wendland C2/C4/C6 and the cubic spline):

    cubic : (1-q)^3 - 4 (1/2-q)_+^3        C2 * pi = 80/7     C3 * pi = 16
    w2    : (1-q)^4 (1 + 4q)                C2 * pi = 7        C3 * pi = 21/2
    w4    : (1-q)^6 (1 + 6q + 35/3 q^2)     C2 * pi = 9        C3 * pi = 495/32
    w6    : (1-q)^8 (1 + 8q + 25q^2 + 32q^3) C2 * pi = 78/7    C3 * pi = 1365/64

The block decomposition ("truncated monomials", reference ``edge/kernels.py:60-121``): the outermost piece
becomes the block with R = 1, and every inner knot k_j contributes the difference of the piece below and the
piece above it, truncated at R = k_j.  Everything here is exact; floats appear only when a plan is flattened.
"""
from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction
from math import comb
from typing import Dict, Tuple

F = Fraction

__all__ = ["Block", "Kernel", "KERNELS", "kernel", "disk_moment", "plane_mass", "plane_second_moment",
           "plane_channels", "NAMES", "AUX_NAMES", "P5_NAMES", "ALL_NAMES", "CLOSED_NAMES", "KERNEL_ID"]


def _poly_mul(a, b):
    out = [F(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def _poly_pow(a, k):
    out = [F(1)]
    for _ in range(k):
        out = _poly_mul(out, a)
    return out


def _poly_add(a, b, sb=1):
    n = max(len(a), len(b))
    return [(a[i] if i < len(a) else F(0)) + sb * (b[i] if i < len(b) else F(0)) for i in range(n)]


def _trim(a):
    a = list(a)
    while len(a) > 1 and a[-1] == 0:
        a.pop()
    return a


@dataclass(frozen=True)
class Block:
    """pi * W restricted to r <= R as a polynomial  sum_n coeffs[n] r^n  (units of h)."""
    R: Fraction
    coeffs: Tuple[Fraction, ...]

    def profile(self) -> Dict[int, Fraction]:
        return {n: c for n, c in enumerate(self.coeffs) if c != 0}


@dataclass(frozen=True)
class Kernel:
    name: str
    c2_pi: Fraction                      # 2-D normalisation times pi
    c3_pi: Fraction                      # 3-D normalisation times pi
    pieces: Tuple[Tuple[Fraction, Fraction, Tuple[Fraction, ...]], ...]   # (lo, hi, coeffs of pi*W) ascending
    blocks: Tuple[Block, ...]            # descending R

    @property
    def radii(self):
        return tuple(b.R for b in self.blocks)

    def w_pi(self, q: Fraction) -> Fraction:
        """pi * W(q) (2-D, h = 1), exact at rational q."""
        q = F(q)
        if q >= 1:
            return F(0)
        tot = F(0)
        for b in self.blocks:
            if q <= b.R:
                tot += sum(c * q ** n for n, c in enumerate(b.coeffs))
        return tot


def _shape_pieces(kind, n=0, poly=()):
    """pieces of the unnormalised shape on [0, 1] as ascending coefficient lists in q."""
    one_minus_q = [F(1), F(-1)]
    if kind == "cubic":
        outer = _poly_pow(one_minus_q, 3)
        inner = _poly_add(outer, [4 * c for c in _poly_pow([F(1, 2), F(-1)], 3)], -1)
        return [(F(0), F(1, 2), inner), (F(1, 2), F(1), outer)]
    p = [F(0)] * (max(k for k, _ in poly) + 1)
    for k, c in poly:
        p[k] += F(c)
    return [(F(0), F(1), _poly_mul(_poly_pow(one_minus_q, n), p))]


def _build(name, kind, c2_pi, c3_pi, n=0, poly=()):
    pieces = [(lo, hi, tuple(_trim([c2_pi * c for c in cs]))) for lo, hi, cs in _shape_pieces(kind, n, poly)]
    blocks = []
    for j in range(len(pieces) - 1, -1, -1):
        lo, hi, cs = pieces[j]
        if j == len(pieces) - 1:
            coeffs = list(cs)
        else:
            coeffs = _poly_add(list(cs), list(pieces[j + 1][2]), -1)
        blocks.append(Block(F(hi), tuple(_trim(coeffs))))
    return Kernel(name, F(c2_pi), F(c3_pi), tuple(pieces), tuple(blocks))


KERNELS: Dict[str, Kernel] = {
    "cubic": _build("cubic", "cubic", F(80, 7), F(16)),
    "w2": _build("w2", "wendland", F(7), F(21, 2), 4, ((0, 1), (1, 4))),
    "w4": _build("w4", "wendland", F(9), F(495, 32), 6, ((0, 1), (1, 6), (2, F(35, 3)))),
    "w6": _build("w6", "wendland", F(78, 7), F(1365, 64), 8, ((0, 1), (1, 8), (2, 25), (3, 32))),
}
NAMES = ("cubic", "w2", "w4", "w6")


# ---- auxiliary kernels of the wall operations (reference edge/kernels.py:168-243, scene/viscosity.py:40-60,
# scene/tensile.py:31-40): truncated powers of u = 1 - q with knot 1, normalised so that int_{R^2} W dA = 1
def _from_u_terms(name, terms):
    """terms [(coefficient, power)]: shape(q) = sum c (1 - q)^p on [0, 1]; C = 1 / (2 int_0^1 q shape dq)."""
    shape = [F(0)]
    for c, p in terms:
        shape = _poly_add(shape, [F(c) * x for x in _poly_pow([F(1), F(-1)], p)])
    qs = _poly_mul([F(0), F(1)], shape)
    c2_pi = 1 / (2 * sum(x / (n + 1) for n, x in enumerate(qs)))
    coeffs = tuple(_trim([c2_pi * x for x in shape]))
    return Kernel(name, c2_pi, F(0), ((F(0), F(1), coeffs),), (Block(F(1), coeffs),))


def _power_terms(k, base, p0):
    """(shape)^k as u-power terms for a Wendland shape u^p0 * sum_m base[m] u^m (reference POWER_FAMILIES)."""
    pol = [F(1)]
    for _ in range(k):
        pol = _poly_mul(pol, [F(b) for b in base])
    return [(c, p0 * k + m) for m, c in enumerate(pol) if c != 0]


_W2_U = ([F(5), F(-4)], 4)                              # (1-q)^4 (1 + 4q)        = u^4 (5 - 4u)
_W4_U = ([F(56, 3), F(-88, 3), F(35, 3)], 6)           # (1-q)^6 (1+6q+35/3 q^2) = u^6 (56/3 - 88/3 u + 35/3 u^2)
KERNELS["cone"] = _from_u_terms("cone", [(1, 1)])                       # the cover vector's kernel
KERNELS["lw2"] = _from_u_terms("lw2", [(1, 3)])                         # L = W'/r of w2 (wall Laplacian)
KERNELS["lw4"] = _from_u_terms("lw4", [(6, 5), (-5, 6)])               # L = W'/r of w4
KERNELS["w2p5"] = _from_u_terms("w2p5", _power_terms(5, *_W2_U))       # W^5 of w2 (tensile), degree 25
KERNELS["w4p5"] = _from_u_terms("w4p5", _power_terms(5, *_W4_U))       # W^5 of w4, degree 40
AUX_NAMES = ("cone", "lw2", "lw4")          # evaluated by the closed form, like NAMES
P5_NAMES = ("w2p5", "w4p5")                 # gradient channels only, by chord quadrature
ALL_NAMES = NAMES + AUX_NAMES + P5_NAMES
CLOSED_NAMES = NAMES + AUX_NAMES
KERNEL_ID = {n: i for i, n in enumerate(ALL_NAMES)}


def kernel(name: str) -> Kernel:
    try:
        return KERNELS[name]
    except KeyError:
        raise KeyError(f"unknown kernel {name!r}; expected one of {sorted(KERNELS)}") from None


def _radial_moment(k: Kernel, power: int) -> Fraction:
    """int_0^1 r^power (pi W) dr, exact."""
    tot = F(0)
    for lo, hi, cs in k.pieces:
        for n, c in enumerate(cs):
            tot += c * (hi ** (n + power + 1) - lo ** (n + power + 1)) / (n + power + 1)
    return tot


def disk_moment(name: str, alpha) -> Fraction:
    """exact whole-support moment  int_{R^2} y^alpha W dA  (h = 1)."""
    a, b = int(alpha[0]), int(alpha[1])
    if a % 2 or b % 2:
        return F(0)

    def dfact(m):
        out = 1
        while m > 1:
            out *= m
            m -= 2
        return out
    ang = F(2 * dfact(a - 1) * dfact(b - 1), dfact(a + b))      # (oint w^alpha dtheta) / pi
    return ang * _radial_moment(kernel(name), a + b + 1)


def plane_mass(name: str) -> Fraction:
    return disk_moment(name, (0, 0))


def plane_second_moment(name: str) -> Fraction:
    """int_{R^2} y_x^2 W dA = int y_y^2 W dA (h = 1)."""
    return disk_moment(name, (2, 0))


def plane_channels(name: str):
    """twelve-channel value of a solid that covers the whole support (device channel order
    lam, m10, m01, Gx, Gy, g1xx, g1xy, g1yx, g1yy, m20, m11, m02), as floats."""
    m, s = float(plane_mass(name)), float(plane_second_moment(name))
    return [m, 0.0, 0.0, 0.0, 0.0, m, 0.0, 0.0, m, s, 0.0, s]
