"""Arithmetic of the oracle -- TEST INFRASTRUCTURE (see oracle/__init__).

The oracle's restatements (``edge2d``, ``mesh3d``) evaluate their formulas through this module, so the same
operation sequence runs in

* ``"f64"``  IEEE double (numpy float64), the default;
* ``"ld"``   x86 extended precision (numpy longdouble, 64-bit significand, unit roundoff 5.4e-20);
* ``"mp"``   mpmath at ``dps`` decimal digits over numpy object arrays (pure-Python loops: small inputs only; every
  division the formulas guard with ``numpy.where`` must be safe on both branches, which holds for ``edge2d``).

Inputs are converted exactly (a float64 is exact in all three); exact rationals (kernel coefficients) are rounded
once to the working precision, and pi is the working precision's own.  The reference evaluates its formulas in
mpmath (``mesh3d/reference.py``, ``edge/core.py`` with ``edge/mpq.py``); the extended modes are the high-precision
references of the precision stress workload.
"""
from __future__ import annotations

import contextlib
from fractions import Fraction

import mpmath
import numpy as np

_mode = "f64"
_dps = 50


def mode():
    return _mode


@contextlib.contextmanager
def precision(name: str, dps: int = 50):
    """``with precision("ld"): ...`` evaluates the oracle in extended precision inside the block."""
    global _mode, _dps
    if name not in ("f64", "ld", "mp"):
        raise ValueError(f"unknown oracle precision {name!r}")
    old = (_mode, _dps, mpmath.mp.dps)
    _mode, _dps = name, int(dps)
    if name == "mp":
        mpmath.mp.dps = int(dps)
    try:
        yield
    finally:
        _mode, _dps, mpmath.mp.dps = old[0], old[1], old[2]


_mp_sqrt = np.frompyfunc(lambda x: mpmath.sqrt(x), 1, 1)
_mp_asinh = np.frompyfunc(lambda x: mpmath.asinh(x), 1, 1)
_mp_atan = np.frompyfunc(lambda x: mpmath.atan(x), 1, 1)
_mp_atan2 = np.frompyfunc(lambda y, x: mpmath.atan2(y, x), 2, 1)
_mp_log = np.frompyfunc(lambda x: mpmath.log(x), 1, 1)
def _to_mpf(x):
    if isinstance(x, Fraction):
        return mpmath.mpf(x.numerator) / x.denominator
    return mpmath.mpf(x)


_mp_conv = np.frompyfunc(_to_mpf, 1, 1)


def asarray(x):
    """the working type's array of ``x`` (exact conversion of float64 / integer inputs; in "mp" mode also of an
    object array of Fractions)."""
    if _mode == "f64":
        return np.asarray(x, dtype=np.float64)
    if _mode == "ld":
        return np.asarray(x, dtype=np.longdouble)
    a = np.asarray(x)
    if a.dtype == object:
        return _mp_conv(a) if a.ndim else _to_mpf(a.item())
    a = a.astype(np.float64) if a.dtype.kind in "iub" else a
    if a.dtype == np.longdouble:
        raise TypeError("mp mode takes float64 inputs")
    return _mp_conv(a.astype(object)) if a.ndim else mpmath.mpf(float(a))


def num(c):
    """an exact rational (Fraction, int, or a decimal string) rounded once to the working precision."""
    c = Fraction(c)
    if _mode == "f64":
        return float(c)
    if _mode == "ld":
        return np.longdouble(c.numerator) / np.longdouble(c.denominator)
    return mpmath.mpf(c.numerator) / c.denominator


def pi():
    if _mode == "f64":
        return np.pi
    if _mode == "ld":
        return 4 * np.arctan(np.longdouble(1))
    return +mpmath.pi


def sqrt(x):
    return _mp_sqrt(x) if _mode == "mp" else np.sqrt(x)


def arcsinh(x):
    return _mp_asinh(x) if _mode == "mp" else np.arcsinh(x)


def arctan(x):
    return _mp_atan(x) if _mode == "mp" else np.arctan(x)


def arctan2(y, x):
    return _mp_atan2(y, x) if _mode == "mp" else np.arctan2(y, x)


def log(x):
    return _mp_log(x) if _mode == "mp" else np.log(x)


def to_float(x):
    """float64 values of a working-precision array (rounded once)."""
    a = np.asarray(x)
    if a.dtype == object:
        return np.vectorize(float, otypes=[np.float64])(a) if a.size else a.astype(np.float64)
    return a.astype(np.float64)


def zeros(shape):
    """a zero array of the working type."""
    if _mode == "f64":
        return np.zeros(shape)
    if _mode == "ld":
        return np.zeros(shape, dtype=np.longdouble)
    out = np.empty(shape, dtype=object)
    out.fill(mpmath.mpf(0))
    return out
