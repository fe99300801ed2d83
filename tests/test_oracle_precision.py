"""The oracle's extended-precision modes (x86 long double, mpmath) pinned by the reference's fixtures: the 40-digit
2-D element totals (tests/golden/edge2d_golden.json) and the 70-digit 3-D tetrahedron and near-face cases
(tests/golden/fluxbridge-v1.json, fluxbridge-nearplane-v1.json).  These modes are the high-precision references of
the precision stress workload.

Tolerances: long double 1e-15 (1 + |ref|) in 2-D (the element coordinates are rounded once to float64 relative to
the query, a 1e-16 perturbation) and 5e-15 in 3-D; mpmath at 50 digits on exact rational inputs 1e-33 (1 + |ref|)
(the fixtures carry 40 digits).  The float64 mode on the same cases is 100 to 5000 times less accurate (4e-13 in
2-D, 4e-12 in 3-D for w6), which the checks below also show, so a mode that silently fell back to float64 fails.
Points on an edge or a vertex are left out of the 2-D checks: there the float64-rounded coordinates put the point
off the edge line in extended precision while the float64 indicator puts it on the edge.  The long double checks
take the indicator of the float64-rounded coordinates; the mpmath checks run on the exact coordinates and take the
fixture's own inside / outside classification (points 1e-20 h from an edge included).
"""
from fractions import Fraction as F

import mpmath
import numpy as np
import pytest

from oracle import adjacency as oadj
from oracle import arith as A
from oracle import edge2d
from oracle import mesh3d as OM
from tests.golden_data import ELEMENTS, KERNELS, ccw, rel_units_h, twelve

OFF_BOUNDARY = [r for r in ELEMENTS if r["where"] in ("inside", "outside")]


def _indicator(rel):
    n = len(rel)
    edges = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
    return oadj.indicator_points(np.asarray(rel, dtype=np.float64), edges, 0, np.zeros((1, 2)), np.ones(1))[0]


def _total(rel, k, ind):
    p, q = rel, np.roll(rel, -1, axis=0)
    return edge2d.edge_channels12(p, q, k).sum(0) + ind * edge2d.plane_vector(k)


@pytest.mark.parametrize("k", KERNELS)
def test_long_double_2d_matches_golden(k):
    worst = {"f64": 0.0, "ld": 0.0}
    for mode in worst:
        with A.precision(mode):
            for rec in OFF_BOUNDARY:
                rel = rel_units_h(rec["polygon"], rec["x"], rec["h"])
                ref = twelve(rec, k)
                got = A.to_float(_total(A.asarray(rel), k, _indicator(rel)))
                worst[mode] = max(worst[mode], (np.abs(got - ref) / (1 + np.abs(ref))).max())
    assert worst["ld"] <= 1e-15, worst
    if k in ("w4", "w6"):                 # the float64 oracle's own error is visible at these kernels
        assert worst["ld"] * 30 < worst["f64"], worst


def _mp(s):
    f = F(s)
    return mpmath.mpf(f.numerator) / f.denominator


def _twelve_mp(rec, k):
    """fixture values mapped to the twelve channels in units of h, at mpmath precision."""
    e = rec["kernels"][k]
    h = _mp(rec["h"])
    val = _mp(e["value"])
    g = [_mp(s) for s in e["grad"]]
    m = {kk: _mp(v) for kk, v in e["moments"].items()}
    mg = {kk: [_mp(a), _mp(b)] for kk, (a, b) in e["moment_grad"].items()}
    return [val, m["1,0"] / h, m["0,1"] / h, g[0] * h, g[1] * h, mg["1,0"][0] + val, mg["1,0"][1], mg["0,1"][0],
            mg["0,1"][1] + val, m["2,0"] / h ** 2, m["1,1"] / h ** 2, m["0,2"] / h ** 2]


@pytest.mark.parametrize("k", KERNELS)
def test_mpmath_2d_matches_golden(k):
    with A.precision("mp", 50):
        worst = mpmath.mpf(0)
        for rec in OFF_BOUNDARY[::2]:
            P = ccw([(F(a), F(b)) for a, b in rec["polygon"]])
            xf, hf = (F(rec["x"][0]), F(rec["x"][1])), F(rec["h"])
            rel = np.array([[(a - xf[0]) / hf, (b - xf[1]) / hf] for a, b in P], dtype=object)
            ind = 1 if rec["where"] == "inside" else 0      # exact inputs: the fixture's own classification
            got = _total(A.asarray(rel), k, ind)
            ref = _twelve_mp(rec, k)
            for g, r in zip(got, ref):
                worst = max(worst, abs(g - r) / (1 + abs(r)))
        assert worst <= mpmath.mpf("1e-33"), (k, worst)


def test_long_double_3d_matches_fixtures():
    import json
    import os
    from warpsphboundaries_amd.plan3d import REQUESTS22
    gold = os.path.join(os.path.dirname(__file__), "golden")
    worst = {"f64": 0.0, "ld": 0.0}
    tets = [c for c in json.load(open(os.path.join(gold, "fluxbridge-v1.json")))["cases"]
            if c["dimension"] == 3 and c["measure"] == "volume"]
    near = json.load(open(os.path.join(gold, "fluxbridge-nearplane-v1.json")))["cases"][::4]    # baseline(3) requests
    for mode in worst:
        with A.precision(mode):
            for c in tets + near:
                V = np.array([[float(x) for x in v] for v in c["vertices"]])
                q = np.array([[float(x) for x in c["query"]]])
                if c in tets:                   # the case's own first 28 requests: baseline(3)
                    req = [(tuple(r["alpha"]), tuple(r["derivative"])) for r in c["requests"]]
                    idx = np.arange(min(len(req), 28))
                    req, ref = [req[i] for i in idx], np.array([float(v) for v in c["values"]])[idx]
                else:
                    req, ref = list(REQUESTS22), np.array([float(v) for v in c["values"]])[:22]
                got = A.to_float(OM.closed_surface_requests(V, OM.tetrahedron_faces(V), q, c["kernel"], float(c["h"]),
                                                            req)[0])
                worst[mode] = max(worst[mode], (np.abs(got - ref) / (1 + np.abs(ref))).max())
    assert worst["ld"] <= 5e-15, worst
    assert worst["ld"] * 30 < worst["f64"], worst


def test_precision_switch_restores_float64():
    with A.precision("ld"):
        assert A.asarray([1.0]).dtype == np.longdouble
        with A.precision("mp", 30):
            assert isinstance(A.pi(), mpmath.mpf) and mpmath.mp.dps == 30
        assert A.mode() == "ld"
    assert A.mode() == "f64" and A.asarray([1.0]).dtype == np.float64
    with pytest.raises(ValueError):
        with A.precision("f128"):
            pass
