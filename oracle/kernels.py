"""Oracle's own kernel table (independent of the product's ``warpsphboundaries_amd.kernels``).  TEST INFRASTRUCTURE.

Derived with sympy from the shapes and 2-D normalisations documented in the reference
(``curvbound/kernels.py:1-50``) and the truncated-monomial block rule of ``edge/kernels.py:1-12,96-108``: the
outermost piece is the block with R = 1, each inner knot k contributes (piece below - piece above), truncated at k.
Returns pi-free blocks [(R, [c_0, c_1, ...])] with exact ``Fraction`` coefficients.
"""
from fractions import Fraction
from functools import lru_cache

import sympy as sp

_q = sp.Symbol("q")
_SPEC = {
    "cubic": (sp.Rational(80, 7), [((0, sp.Rational(1, 2)), (1 - _q) ** 3 - 4 * (sp.Rational(1, 2) - _q) ** 3),
                                    ((sp.Rational(1, 2), 1), (1 - _q) ** 3)]),
    "w2": (sp.Integer(7), [((0, 1), (1 - _q) ** 4 * (1 + 4 * _q))]),
    "w4": (sp.Integer(9), [((0, 1), (1 - _q) ** 6 * (1 + 6 * _q + sp.Rational(35, 3) * _q ** 2))]),
    "w6": (sp.Rational(78, 7), [((0, 1), (1 - _q) ** 8 * (1 + 8 * _q + 25 * _q ** 2 + 32 * _q ** 3))]),
}


def _normalised(shape):
    """(c2 pi, [((0, 1), shape)]) with 2 int_0^1 q (c2 pi) shape dq = 1 (mass 1 in the plane)."""
    return (1 / (2 * sp.integrate(_q * sp.expand(shape), (_q, 0, 1))), [((0, 1), shape)])


# the wall-operation kernels (reference edge/kernels.py:168-170, scene/viscosity.py:1-40, scene/tensile.py:1-30):
# the cone (1 - q) of the cover vector, L = W'/r of w2 / w4 divided by its constant (the wall Laplacian), and W^5
_W2 = (1 - _q) ** 4 * (1 + 4 * _q)
_W4 = (1 - _q) ** 6 * (1 + 6 * _q + sp.Rational(35, 3) * _q ** 2)
_SPEC["cone"] = _normalised(1 - _q)
_SPEC["lw2"] = _normalised(sp.expand(sp.diff(_W2, _q) / _q) / -20)
_SPEC["lw4"] = _normalised(sp.expand(sp.diff(_W4, _q) / _q) / sp.Rational(-56, 3))
_SPEC["w2p5"] = _normalised(_W2 ** 5)
_SPEC["w4p5"] = _normalised(_W4 ** 5)


def _coeffs(expr):
    p = sp.Poly(sp.expand(expr), _q)
    deg = p.degree()
    out = [Fraction(0)] * (deg + 1)
    for (k,), c in p.terms():
        out[k] = Fraction(int(sp.fraction(c)[0]), int(sp.fraction(c)[1]))
    return out


@lru_cache(maxsize=None)
def blocks(name):
    c2pi, pieces = _SPEC[name]
    out = []
    for j in range(len(pieces) - 1, -1, -1):
        (lo, hi), e = pieces[j]
        if j == len(pieces) - 1:
            poly = c2pi * e
        else:
            poly = c2pi * (e - pieces[j + 1][1])
        out.append((Fraction(int(sp.fraction(sp.nsimplify(hi))[0]), int(sp.fraction(sp.nsimplify(hi))[1])), _coeffs(poly)))
    return tuple(out)


@lru_cache(maxsize=None)
def w_pi(name, q):
    """pi W(q) exactly (sympy Rational q)."""
    c2pi, pieces = _SPEC[name]
    q = sp.nsimplify(q)
    for (lo, hi), e in pieces:
        if lo <= q <= hi and q < 1:
            return c2pi * e.subs(_q, q)
    return sp.Integer(0)


@lru_cache(maxsize=None)
def c3_pi(name):
    """pi C3 from the 3-D normalisation 4 pi C3 int_0^1 q^2 w(q) dq = 1 (curvbound/kernels.py:10-14 lists the values)."""
    _, pieces = _SPEC[name]
    m = sum(sp.integrate(e * _q ** 2, (_q, lo, hi)) for (lo, hi), e in pieces)
    v = sp.nsimplify(1 / (4 * m))
    return Fraction(int(sp.fraction(v)[0]), int(sp.fraction(v)[1]))


def c3_over_c2(name):
    c2 = _SPEC[name][0]
    return c3_pi(name) / Fraction(int(sp.fraction(c2)[0]), int(sp.fraction(c2)[1]))


def radial_integral(name, power):
    """exact int_0^1 r^power pi W dr."""
    c2pi, pieces = _SPEC[name]
    return sum(sp.integrate(c2pi * e * _q ** power, (_q, lo, hi)) for (lo, hi), e in pieces)


@lru_cache(maxsize=None)
def plane_vector_exact(name):
    """twelve-channel whole-support vector (device order) as exact Fractions: mass and int y_x^2 W."""
    mass = 2 * radial_integral(name, 1)                 # 2 pi int r W dr / pi * pi ... (pi W normalised: 2 int r (pi W) dr)
    second = radial_integral(name, 3)                   # int y_x^2 W = (1/pi) pi int r^3 (pi W) dr
    m = Fraction(int(sp.fraction(mass)[0]), int(sp.fraction(mass)[1]))
    s = Fraction(int(sp.fraction(second)[0]), int(sp.fraction(second)[1]))
    z = Fraction(0)
    return (m, z, z, z, z, m, z, z, m, s, z, s)


def plane_vector(name):
    """twelve-channel whole-support vector (device order) in float64."""
    return [float(v) for v in plane_vector_exact(name)]
