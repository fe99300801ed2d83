"""Loader of the reference's 40-digit golden fixtures (tests/golden/edge2d_golden.json, data only)."""
import json
import math
import os
from fractions import Fraction as F

import numpy as np

PATH = os.path.join(os.path.dirname(__file__), "golden", "edge2d_golden.json")
DATA = json.load(open(PATH))
CASES = DATA["cases"]
ELEMENTS = [c for c in CASES if c["kind"] == "element"]
MESHES = [c for c in CASES if c["kind"] == "mesh"]
KERNELS = ["cubic", "w2", "w4", "w6"]


def num(s):
    return float(F(s)) if "/" in s else float(s)


def ccw(poly):
    """exact orientation: counter-clockwise around the solid."""
    n = len(poly)
    a2 = sum(poly[i][0] * poly[(i + 1) % n][1] - poly[(i + 1) % n][0] * poly[i][1] for i in range(n))
    return poly if a2 > 0 else poly[::-1]


def rel_units_h(poly, x, h):
    """(v - x)/h exactly, rounded once to float64; the query sits at the origin."""
    P = ccw([(F(a), F(b)) for a, b in poly])
    xf = (F(x[0]), F(x[1]))
    hf = F(h)
    return np.array([[float((a - xf[0]) / hf), float((b - xf[1]) / hf)] for a, b in P])


def twelve(rec, k):
    """fixture values mapped to the twelve device channels in units of h."""
    e = rec["kernels"][k]
    h = float(F(rec["h"]))
    val = num(e["value"])
    g = [num(s) for s in e["grad"]]
    m = {kk: num(v) for kk, v in e["moments"].items()}
    mg = {kk: [num(a), num(b)] for kk, (a, b) in e["moment_grad"].items()}
    return np.array([val, m["1,0"] / h, m["0,1"] / h, g[0] * h, g[1] * h,
                     mg["1,0"][0] + val, mg["1,0"][1], mg["0,1"][0], mg["0,1"][1] + val,
                     m["2,0"] / h ** 2, m["1,1"] / h ** 2, m["0,2"] / h ** 2])
