"""Write ``tables.inc``: the exact kernel-block profiles rounded once to float64 constants, as compile-time
tables for the optimized HIP kernels (one template specialisation per kernel, fully unrolled loops).

Run by the build (``__graft_entry__.build`` / ``warpsphboundaries_amd.build``); output next to this file.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from warpsphboundaries_amd import plan as PL          # noqa: E402
from warpsphboundaries_amd.kernels import NAMES, plane_channels   # noqa: E402

NMAX = 16          # fixed array extent (>= max power + 1 of every profile)


def _arr(vals):
    vals = list(vals) + [0] * (NMAX - len(vals))
    return "{" + ", ".join(repr(float(v)) for v in vals) + "}"


def render():
    out = ["// generated by gen_tables.py from the exact kernel table -- do not edit",
           "#pragma once", "", "template <int KID> struct KTab;", ""]
    for kid, name in enumerate(NAMES):
        bl = PL.block_profiles(name)
        nb = len(bl)
        degP = [len(b.P) - 1 for b in bl]
        degF = [len(b.Phi) - 1 for b in bl]
        QP = [[c / (n + 2) for n, c in enumerate(b.P)] for b in bl]
        QF = [[c / (n + 2) for n, c in enumerate(b.Phi)] for b in bl]
        out.append(f"// {name}")
        out.append(f"template <> struct KTab<{kid}> {{")
        out.append(f"  static constexpr int NB = {nb};")
        out.append(f"  static constexpr int NMAXP = {max(max(degP), max(degF))};")
        out.append(f"  static constexpr double R[NB] = {{{', '.join(repr(float(b.R)) for b in bl)}}};")
        out.append(f"  static constexpr int DEGP[NB] = {{{', '.join(map(str, degP))}}};")
        out.append(f"  static constexpr int DEGF[NB] = {{{', '.join(map(str, degF))}}};")
        out.append(f"  static constexpr double P[NB][{NMAX}] = {{{', '.join(_arr(b.P) for b in bl)}}};")
        out.append(f"  static constexpr double F[NB][{NMAX}] = {{{', '.join(_arr(b.Phi) for b in bl)}}};")
        out.append(f"  static constexpr double QP[NB][{NMAX}] = {{{', '.join(_arr(q) for q in QP)}}};")
        out.append(f"  static constexpr double QF[NB][{NMAX}] = {{{', '.join(_arr(q) for q in QF)}}};")
        out.append(f"  static constexpr double MRP[NB] = {{{', '.join(repr(float(b.mR_P)) for b in bl)}}};")
        out.append(f"  static constexpr double MRF[NB] = {{{', '.join(repr(float(b.mR_Phi)) for b in bl)}}};")
        out.append(f"  static constexpr double PLANE[12] = {{{', '.join(repr(v) for v in plane_channels(name))}}};")
        out.append("};")
        out.append("")
    # 3-D sphere: per block the moment polynomial coefficients c_n/(n+k), k = 2, 3, 4
    from warpsphboundaries_amd.sphere import device_tables
    out.append("template <int KID> struct STab;")
    for kid, name in enumerate(NAMES):
        rows = device_tables()[name]
        nb = len(rows)
        deg = max(len(r[1]) for r in rows) - 1
        out.append(f"template <> struct STab<{kid}> {{")
        out.append(f"  static constexpr int NB = {nb};")
        out.append(f"  static constexpr int DEG = {deg};")
        out.append(f"  static constexpr double R[NB] = {{{', '.join(repr(r[0]) for r in rows)}}};")
        for k, idx in ((2, 1), (3, 2), (4, 3)):
            out.append(f"  static constexpr double M{k}[NB][{NMAX}] = {{{', '.join(_arr(r[idx]) for r in rows)}}};")
        # the same coefficients split exactly: c = hi + lo (lo carries the rounding of the exact rational)
        from warpsphboundaries_amd.sphere import blocks3d
        from fractions import Fraction
        for k in (2, 3, 4):
            his, los = [], []
            for Rj, c in blocks3d(name):
                exact = [x / (n + k) for n, x in enumerate(c)]
                hi = [float(x) for x in exact]
                lo = [float(x - Fraction(h)) for x, h in zip(exact, hi)]
                his.append(hi)
                los.append(lo)
            out.append(f"  static constexpr double M{k}H[NB][{NMAX}] = {{{', '.join(_arr(x) for x in his)}}};")
            out.append(f"  static constexpr double M{k}L[NB][{NMAX}] = {{{', '.join(_arr(x) for x in los)}}};")
        out.append("};")
        out.append("")
    return "\n".join(out) + "\n"


def main(path=None):
    path = path or os.path.join(HERE, "tables.inc")
    text = render()
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


if __name__ == "__main__":
    print(main())
