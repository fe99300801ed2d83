"""The disk element on the host (reference edge/disk.py, tests/edge/test_disk.py, tests/scene/test_disk_array.py):

(a) the oracle's circle identity (``oracle.disk2d.circle_channels``, a restatement of edge/disk.py plus the second
    moments) against the polar integrals evaluated by mpmath at 30 digits: every channel, every kernel family;
(b) the reference's limits: the support inside a large disk gives lam = 1, g1 = I, m2 = the whole-support second
    moment and no gradient; nothing beyond reach; smooth across the circle and converged (24 vs 48 nodes);
(c) the library's own arithmetic (``wsb_disk_channels_host``, the code the device runs, compiled for the host)
    against the oracle at 1e-12 on a grid of radii and distances that crowds the circle, the centre and the knots;
(d) the slot list: the oracle's cell list and slot fill against brute force, K bounded by the 3 x 3 blocks.
No device calls."""
import math
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest
import sympy as sp

from oracle import disk2d as OD
from oracle import kernels as OK
from warpsphboundaries_amd.device import disk_channels_host

KEYS = ("lam", "m1x", "g0x", "g1xx", "g1yy", "m2xx", "m2yy")


def polar_mp(name, D, R, dps=30):
    """the seven channels by their definition in polar coordinates about the query: the arc of the circle of radius
    r inside the disk has half angle alpha(r), so each channel is one radial integral (mpmath quadrature)."""
    mp.mp.dps = dps
    c2pi, pcs = OK._SPEC[name]
    q = OK._q
    polys = []
    for (lo, hi), e in pcs:
        P = sp.Poly(sp.expand(c2pi * e), q)
        dP = P.diff(q)
        to_mp = lambda poly: [(k[0], mp.mpf(sp.Rational(c).p) / sp.Rational(c).q) for k, c in poly.terms()]
        polys.append((mp.mpf(sp.Rational(lo).p) / sp.Rational(lo).q, mp.mpf(sp.Rational(hi).p) / sp.Rational(hi).q,
                      to_mp(P), to_mp(dP)))

    def prof(r, der):
        for lo, hi, P, dP in polys:
            if lo <= r < hi:
                return mp.fsum(c * r ** k for k, c in (dP if der else P)) / mp.pi
        return mp.mpf(0)
    D, R = mp.mpf(D), mp.mpf(R)

    def alpha(r):
        if r <= abs(D - R):
            return mp.pi if D < R else mp.mpf(0)
        if r >= D + R:
            return mp.mpf(0)
        return mp.acos((r * r + D * D - R * R) / (2 * r * D))
    bps = sorted(set([mp.mpf(0), mp.mpf(1)] + [v for v in (abs(D - R), D + R) + tuple(p[1] for p in polys) if 0 < v < 1]))

    def integral(f):
        return mp.fsum(mp.quad(f, [lo, hi]) for lo, hi in zip(bps[:-1], bps[1:]))
    sc = lambda r: mp.sin(alpha(r)) * mp.cos(alpha(r))
    return dict(lam=integral(lambda r: prof(r, 0) * 2 * alpha(r) * r),
                m1x=integral(lambda r: prof(r, 0) * 2 * mp.sin(alpha(r)) * r * r),
                g0x=integral(lambda r: -prof(r, 1) * 2 * mp.sin(alpha(r)) * r),
                g1xx=integral(lambda r: -prof(r, 1) * r * r * (alpha(r) + sc(r))),
                g1yy=integral(lambda r: -prof(r, 1) * r * r * (alpha(r) - sc(r))),
                m2xx=integral(lambda r: prof(r, 0) * r ** 3 * (alpha(r) + sc(r))),
                m2yy=integral(lambda r: prof(r, 0) * r ** 3 * (alpha(r) - sc(r))))


MP_CASES = [("w2", 0.31, 0.3), ("w2", 0.4, 0.9), ("cubic", 0.45000001, 0.05), ("cubic", 0.75, 0.5),
            ("w4", 1.3, 0.5), ("w4", 0.0, 0.5), ("w6", 1.000001, 1.0), ("w6", 2.7, 3.0),
            ("cone", 0.2, 0.35), ("lw2", 0.9, 0.2), ("lw4", 0.6, 1.2), ("w2p5", 0.35, 0.3)]


@pytest.mark.parametrize("name,D,R", MP_CASES)
def test_oracle_circle_identity_equals_the_polar_integrals(name, D, R):
    ref = polar_mp(name, D, R)
    got = OD.circle_channels(name, D, R)
    for k in KEYS:
        assert abs(float(got[k]) - float(ref[k])) <= 1e-13 * max(1.0, abs(float(ref[k]))), (k, float(got[k]), float(ref[k]))


@pytest.mark.parametrize("name", ["w2", "w4", "cubic", "w6"])
def test_oracle_limits(name):
    big = OD.circle_channels(name, np.array([0.0, 0.3, 4.9]), 6.0)      # the support inside a large disk
    second = float(OK.plane_vector_exact(name)[9])
    assert np.allclose(big["lam"], 1.0, atol=1e-14) and np.abs(big["g0x"]).max() < 1e-13
    assert np.allclose(big["g1xx"], 1.0, atol=1e-13) and np.allclose(big["g1yy"], 1.0, atol=1e-13)
    assert np.abs(big["m1x"]).max() < 1e-13
    assert np.allclose(big["m2xx"], second, atol=1e-14) and np.allclose(big["m2yy"], second, atol=1e-14)
    far = OD.circle_channels(name, 3.0, 0.5)
    assert max(abs(float(v)) for v in far.values()) < 1e-14


@pytest.mark.parametrize("R", [0.2, 0.5, 2.0])
def test_oracle_smooth_across_the_surface_and_converged(R):
    h = 0.004
    D = R + h * (np.arange(-15, 15) + 0.5)
    for k in KEYS:
        v = OD.circle_channels("w2", D, R)[k]
        assert np.abs(np.diff(v, 4)).max() < 5e-6 * max(np.abs(v).max(), 0.05) + 1e-7, k
    a, b = OD.circle_channels("w2", D, R, nodes=24), OD.circle_channels("w2", D, R, nodes=48)
    assert max(float(np.abs(a[k] - b[k]).max()) for k in a) < 1e-13


def _grid(R):
    base = [0.0, 2e-12, 0.3 * R, 0.9 * R, R * (1 - 1e-9), R, R * (1 + 1e-12), R + 1e-6, R + 0.37, R + 0.9, R + 0.999999,
            R - 1e-7, R + 1e-3, 1.0 - R, 1.0 - R + 1e-6, 1.0 - R - 1e-6,
            abs(R - 0.5) + 1e-10, 0.5 + R - 1e-10, abs(1 - R) * (1 + 1e-9), max(R - 0.6, 0.05), R - 1.0 + 1e-9]
    return np.array(sorted({d for d in base if d >= 0}))


@pytest.mark.parametrize("name", ["cubic", "w2", "w4", "w6", "cone", "lw2", "lw4", "w2p5", "w4p5"])
def test_library_arithmetic_equals_the_oracle(name):
    worst = 0.0
    for R in (0.01, 0.1, 0.3, 0.5, 0.55, 0.667, 1.0, 2.0, 6.0):
        D = _grid(R)
        got = disk_channels_host(name, D, R)
        ref = OD.circle_channels(name, D, R)
        for j, k in enumerate(KEYS):
            err = np.abs(got[:, j] - ref[k]) / np.maximum(1.0, np.abs(ref[k]))
            worst = max(worst, float(err.max()))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("name", ["w2", "w6"])
def test_library_next_to_the_centre(name):
    """a query 1e-13 h from the centre: the reference's circle identity evaluates it at D = 1e-12 (its floor for the
    angle panels, edge/disk.py:116), 4e-12 off in g0; the library's radial form is checked against mpmath instead."""
    for R in (0.1, 0.5):
        got = disk_channels_host(name, [1e-13], R)[0]
        ref = polar_mp(name, 1e-13, R)
        for j, k in enumerate(KEYS):
            assert abs(got[j] - float(ref[k])) <= 1e-14 * max(1.0, abs(float(ref[k]))), (R, k)


def test_slots_match_brute_force_and_K_is_bounded():
    rng = np.random.default_rng(11)
    H = 0.3
    centres = np.array([(0.0, 0.0), (0.9, 0.2), (-0.7, 0.6), (0.3, -0.8), (0.35, -0.75)])
    radii = np.array([0.1, 0.6, 2.0, 6.0, 0.01]) * H
    pos = rng.uniform(-1.6, 1.6, (400, 2))
    sup = np.full(400, H) * rng.uniform(0.5, 1.0, 400)
    cl = OD.cells(centres, radii, H)
    slots = OD.disk_slots(cl, centres, radii, pos, sup)
    d = np.linalg.norm(pos[:, None, :] - centres[None], axis=2)
    for i in range(len(pos)):
        want = set(np.nonzero(d[i] < sup[i] + radii)[0].tolist())
        got = [int(m) for m in slots[i] if m >= 0]
        assert set(got) == want and len(got) == len(want), i
        assert list(slots[i][:len(got)]) == got and (slots[i][len(got):] == -1).all()
    assert 1 <= cl[-1] <= len(radii)
    valid = np.arange(400) % 3 != 0
    s2 = OD.disk_slots(cl, centres, radii, pos, sup, valid)
    assert (s2[~valid] == -1).all() and (s2[valid] == slots[valid]).all()
