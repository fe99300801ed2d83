"""Conditioned small / sliver triangle route on the CPU: the oracle restatement (oracle/triangle2d.py) pinned by the
reference's 40-digit golden elements and by the edge-local closed form evaluated in 50-digit mpmath; the library's
host entry point (wsb_triangle2d_host, the device arithmetic) against the oracle; the reference's own properties of
the route (tests/edge/test_triangle_adaptive.py, tests/scene/test_complete_surface_moments.py of the reference).

Error measure (the quadrature's own): per channel |v - r| / s_k over max_k |r_k| / s_k + J, with s_k the channel's
scale from the triangle's extent about the particle (units of h) and J = twice the area / h^2 (a floor for
triangles whose support overlap is a 1e-12 corner, where every value is ~1e-70).  Gate against the fixtures and
mpmath: |v - r| <= estimate + 1e-10 (signal) per channel -- every value within 1e-10 or within the route's own
empirical estimate (a particle 5e-4 h from the long edge of a 1e-3 sliver: 2e-9 .. 6e-9, estimate 3e-6 .. 1e-5;
everything else is below 1e-10; the reference's tests use 2e-10 to 2e-11 on small triangles); host vs oracle 1e-13
(the same arithmetic summed in another order).  The fixtures' g1 diagonal is mapped as moment gradient + value in exact arithmetic: for a
1e-6 h element the two terms cancel to 1e-13 of their size.
"""
from fractions import Fraction as F

import mpmath
import numpy as np
import pytest

from oracle import arith as A
from oracle import edge2d
from oracle import kernels as OK
from oracle import triangle2d as OT
from tests.golden_data import ELEMENTS, KERNELS, rel_units_h
from warpsphboundaries_amd import triangle as T

E = np.array([[0, 1], [1, 2], [2, 0]])


def _scales(V, meta, x, h):
    a = (V[meta[0][0]] - x) / h
    b = (V[meta[0][1]] - V[meta[0][0]]) / h
    c = (V[meta[0][2]] - V[meta[0][0]]) / h
    sx = max(abs(a[0]), abs(a[0] + b[0]), abs(a[0] + c[0]), 1e-150)
    sy = max(abs(a[1]), abs(a[1] + b[1]), abs(a[1] + c[1]), 1e-150)
    return np.array([1, sx, sy, sx, sy, sx * sx, sx * sy, sy * sx, sy * sy, sx * sx, sx * sy, sy * sy])


def _err(v, r, s, jac):
    return float((np.abs(v - r) / s).max() / ((np.abs(r) / s).max() + jac))


def _within(v, r, est, s, jac):
    """every channel within 1e-10 of the signal or within the estimate."""
    norm = (np.abs(r) / s).max() + jac
    return bool((np.abs(v - r) <= est + 1e-10 * norm * s).all())


def _twelve_exact(rec, k):
    e = rec["kernels"][k]
    h = F(rec["h"])
    val = F(e["value"])
    g = [F(x) for x in e["grad"]]
    m = {kk: F(v) for kk, v in e["moments"].items()}
    mg = {kk: [F(a), F(b)] for kk, (a, b) in e["moment_grad"].items()}
    out = [val, m["1,0"] / h, m["0,1"] / h, g[0] * h, g[1] * h, mg["1,0"][0] + val, mg["1,0"][1], mg["0,1"][0],
           mg["0,1"][1] + val, m["2,0"] / h ** 2, m["1,1"] / h ** 2, m["0,2"] / h ** 2]
    return np.array([float(x) for x in out])


CONDITIONED = [r for r in ELEMENTS if len(r["polygon"]) == 3
               and OT.row(rel_units_h(r["polygon"], r["x"], r["h"]),
                          OT.prepare(rel_units_h(r["polygon"], r["x"], r["h"]), E), np.zeros(2), 1.0, "w2",
                          max_panels=1)[2] != 0]


def test_fixture_set_covers_small_and_sliver_elements():
    ids = {r["id"] for r in CONDITIONED}
    assert len(CONDITIONED) == 11, sorted(ids)
    assert {"small_element-1e6-x_inside", "small_element-1e6-x_vertex", "tiny_chord-both_ends",
            "x_at_vertex-acute"} <= ids


@pytest.mark.parametrize("k", KERNELS)
def test_oracle_matches_golden_small_and_sliver_elements(k):
    worst = 0.0
    for rec in CONDITIONED:
        rel = rel_units_h(rec["polygon"], rec["x"], rec["h"])
        meta = OT.prepare(rel, E)
        v, est, st = OT.row(rel, meta, np.zeros(2), 1.0, k)
        assert st == 1, rec["id"]
        ref, s = _twelve_exact(rec, k), _scales(rel, meta, np.zeros(2), 1.0)
        e = _err(v, ref, s, meta[1])
        worst = max(worst, e)
        assert _within(v, ref, est, s, meta[1]), (rec["id"], k, e)
    print(k, "worst", worst)


def _mp_triangle(V, x, h, k):
    """the edge-local closed form of the triangle (counter-clockwise, solid inside) in 50-digit mpmath from the
    float64 inputs taken exactly, plus the indicator from exact orientation tests (points off the boundary)."""
    with A.precision("mp", 50):
        rel = np.array([[(mpmath.mpf(float(p[0])) - mpmath.mpf(float(x[0]))) / mpmath.mpf(float(h)),
                         (mpmath.mpf(float(p[1])) - mpmath.mpf(float(x[1]))) / mpmath.mpf(float(h))] for p in V],
                       dtype=object)
        c = edge2d.edge_channels12(rel, np.roll(rel, -1, axis=0), k).sum(0)
        Vf = [(F(float(p[0])), F(float(p[1]))) for p in V]
        xf = (F(float(x[0])), F(float(x[1])))
        sides = [(Vf[(i + 1) % 3][0] - Vf[i][0]) * (xf[1] - Vf[i][1]) - (Vf[(i + 1) % 3][1] - Vf[i][1]) * (xf[0] - Vf[i][0])
                 for i in range(3)]
        assert all(s != 0 for s in sides)
        ind = 1 if all(s > 0 for s in sides) else 0
        plane = [mpmath.mpf(q.numerator) / q.denominator for q in OK.plane_vector_exact(k)]
        return np.array([float(c[j] + ind * plane[j]) for j in range(12)])


def _cases():
    rng = np.random.default_rng(41)
    out = []
    for size in (1e-9, 1e-6, 1e-3, 1.5e-2):                 # small: longest edge <= 0.02 h
        V = rng.uniform(0.0, 1.0, (3, 2)) * size + [0.31, -0.12]
        c = V.mean(0)
        for x in (c, c + [0.37, -0.21], c + [0.05, 0.6], V[0] + [size * 1e-3, -size * 2e-3]):
            out.append((V, x))
    for aspect in (1e-3, 1e-6, 1e-9):                       # slivers: twice the area <= 0.01 longest^2
        L = rng.uniform(0.3, 1.5)
        V = np.array([[0.0, 0.0], [L, aspect * L * 0.7], [L * 0.4, aspect * L]]) + [0.2, 0.1]
        for x in (V.mean(0) + [0.0, 0.3], V.mean(0) + [0.1, -0.45], np.array([0.9, 0.1]), V[1] + [0.02, 0.01]):
            out.append((V, x))
    return out


@pytest.mark.parametrize("k", KERNELS)
def test_oracle_matches_mpmath_closed_form(k):
    worst = 0.0
    for V, x in _cases():
        P = V if (V[1, 0] - V[0, 0]) * (V[2, 1] - V[0, 1]) - (V[1, 1] - V[0, 1]) * (V[2, 0] - V[0, 0]) > 0 else V[::-1]
        meta = OT.prepare(P, E)
        v, est, st = OT.row(P, meta, x, 1.0, k)
        assert st == 1
        ref, s = _mp_triangle(P, x, 1.0, k), _scales(P, meta, x, 1.0)
        e = _err(v, ref, s, meta[1])
        worst = max(worst, e)
        assert _within(v, ref, est, s, meta[1]), (k, V.tolist(), x.tolist(), e)
    print(k, "worst", worst)


@pytest.mark.parametrize("k", KERNELS)
def test_host_entry_matches_oracle(k):
    for V, x in _cases()[::3]:
        for solid in ("inside", "outside"):
            from warpsphboundaries_amd.scene import SurfaceRep
            rep = SurfaceRep.polygon(V, solid)
            Vr, Er = rep.vertices.numpy(), rep.edges.numpy()
            P = np.stack([x, x + [0.2, 0.1], x + [3.0, 0.0]])
            H = np.array([1.0, 0.7, 1.0])
            hv, he, hs = T.channels_host(Vr, P, H, rep._triangle, k, rep.background)
            ov, oe, os_ = OT.channels(Vr, Er, P, H, k, rep.background)
            assert np.array_equal(hs, os_)
            for i in range(len(P)):
                meta = OT.prepare(Vr, Er)
                s = _scales(Vr, meta, P[i], H[i])
                assert _err(hv[i], ov[i], s, meta[1] / H[i] ** 2) <= 1e-13, (k, solid, i)
                assert _err(he[i], oe[i], s, meta[1] / H[i] ** 2) <= 1e-11


def test_orientation_outside_rotation_and_support_scaling():
    V = np.array([[0.0, 0.0], [1.2, 1e-7], [0.5, 3e-7]]) + [0.1, 0.2]
    x = np.array([0.16, 0.41])
    from warpsphboundaries_amd.scene import SurfaceRep
    for k in ("cubic", "w4"):
        rin, rout = SurfaceRep.polygon(V, "inside"), SurfaceRep.polygon(V, "outside")
        vi, _, si = T.channels_host(rin.vertices.numpy(), x[None], 1.0, rin._triangle, k, 0)
        vo, _, so = T.channels_host(rout.vertices.numpy(), x[None], 1.0, rout._triangle, k, 1)
        assert si[0] == so[0] == 1 and rin._triangle[2] == 1 and rout._triangle[2] == -1
        # reversed loop: outside = whole support - inside
        np.testing.assert_allclose(vi[0] + vo[0], T.plane_channels(k), rtol=0, atol=1e-15)
        # 90 degree rotation: vectors rotate, tensors rotate on both sides
        rot = lambda y: np.stack((-y[..., 1], y[..., 0]), axis=-1)
        rr = SurfaceRep.polygon(rot(V), "inside")
        vr, _, _ = T.channels_host(rr.vertices.numpy(), rot(x)[None], 1.0, rr._triangle, k, 0)
        perm = [0, 2, 1, 4, 3, 8, 7, 6, 5, 11, 10, 9]
        sign = np.array([1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1])
        np.testing.assert_allclose(vr[0], vi[0][perm] * sign, rtol=1e-10, atol=1e-25)
        # support scaling: all lengths times s leaves the values in units of h unchanged
        for s in (0.125, 8.0):
            rs = SurfaceRep.polygon(V * s, "inside")
            vs, _, _ = T.channels_host(rs.vertices.numpy(), (x * s)[None], s, rs._triangle, k, 0)
            np.testing.assert_allclose(vs[0], vi[0], rtol=1e-10, atol=1e-25)


def test_route_selection_budget_and_degenerate_geometry():
    from warpsphboundaries_amd.scene import SurfaceRep
    big = SurfaceRep.polygon([[0.0, 0.0], [0.3, 0.0], [0.1, 0.25]])
    v, e, s = T.channels_host(big.vertices.numpy(), np.array([[0.1, 0.1]]), 1.0, big._triangle, "w4")
    assert s[0] == 0 and not v.any() and not e.any()              # neither small nor a sliver: the closed form
    v, e, s = T.channels_host(big.vertices.numpy(), np.array([[0.1, 0.1]]), 100.0, big._triangle, "w4")
    assert s[0] == 1                                                # longest edge < 0.02 h: the quadrature
    sliver = SurfaceRep.polygon([[0.0, 0.0], [1.0, 0.0], [0.5, 1e-7]])
    v, e, s = T.channels_host(sliver.vertices.numpy(), np.array([[0.1, 0.2]]), 1.0, sliver._triangle, "w4",
                              tolerance=1e-30, max_panels=1)
    assert s[0] == -1 and np.isnan(v).all() and np.isfinite(e).all()
    valid = np.array([0], dtype=np.uint8)
    v, e, s = T.channels_host(sliver.vertices.numpy(), np.array([[0.1, 0.2]]), 1.0, sliver._triangle, "w4", valid=valid)
    assert s[0] == 0 and not v.any()
    assert SurfaceRep.polygon([[0, 0], [1, 0], [2, 0]])._triangle is None            # collinear
    assert SurfaceRep.polygon([[0, 0], [1, 0], [1, 1], [0, 1]])._triangle is None    # not a triangle
    assert SurfaceRep.polygon([[0, 0], [1, 0], [0, 1]], conditioned=False)._triangle is None
    eps = np.finfo(float).eps
    tri = np.array([[0.0, 0.0], [1.0, 1.0 - eps], [1.0 + eps, 1.0]])
    for order in (tri, tri[::-1]):
        rep = SurfaceRep.polygon(order)
        assert rep._triangle is not None and rep._triangle[2] > 0 and rep._triangle[1] == eps * eps
        out = SurfaceRep.polygon(order, "outside")
        assert out._triangle[2] < 0 and out.background == 1
    with pytest.raises(ValueError):
        T.channels_host(sliver.vertices.numpy(), np.array([[0.1, 0.2]]), 1.0, sliver._triangle, "w4", tolerance=0.0)
