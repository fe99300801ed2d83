"""Write ``tables.inc``: the exact kernel-block profiles rounded once to float64 constants, as compile-time
tables for the optimized HIP kernels (one template specialisation per kernel, fully unrolled loops).

Run by the build (``__graft_entry__.build`` / ``warpsphboundaries_amd.build``); output next to this file.
"""
import os
import struct
import sys
from math import gcd

import mpmath

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from warpsphboundaries_amd import plan as PL          # noqa: E402
from fractions import Fraction as F                   # noqa: E402

from warpsphboundaries_amd.kernels import ALL_NAMES, CLOSED_NAMES, NAMES, plane_channels   # noqa: E402

NMAX = 16          # fixed array extent (>= max power + 1 of every profile)


def _literal(x):
    """True when the double nearest x is exact with a zero low word (an FP64 instruction's 32-bit literal)"""
    v = float(x)
    return F(v) == x and struct.pack("<d", v)[:4] == b"\0\0\0\0"


def _arr(vals):
    vals = list(vals) + [0] * (NMAX - len(vals))
    return "{" + ", ".join(repr(float(v)) for v in vals) + "}"


def render():
    out = ["// generated by gen_tables.py from the exact kernel table -- do not edit",
           "#pragma once", "", "template <int KID> struct KTab;", ""]
    for kid, name in enumerate(CLOSED_NAMES):
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
        # the fused pass's integer form: P and Phi times the least common denominator SC of their coefficients, so
        # that every coefficient is an integer that an FP64 instruction takes as its 32-bit literal operand (the
        # high word of the double); the value sums use the shared bases dI_m / (m + 2) and dR_m / m instead of
        # P / (m + 2), and the channel sums are divided by SC with the factor 1/pi at the end of the row
        sc = 1
        for b in bl:
            for c in list(b.P) + list(b.Phi):
                sc = sc * c.denominator // gcd(sc, c.denominator)
        PS = [[c * sc for c in b.P] for b in bl]
        FS = [[c * sc for c in b.Phi] for b in bl]
        assert all(x.denominator == 1 for r in PS + FS for x in r)
        out.append(f"  static constexpr double SC = {float(sc)!r};")
        out.append(f"  static constexpr bool LITERAL = {str(all(_literal(x) for r in PS + FS for x in r)).lower()};")
        out.append(f"  static constexpr double PS[NB][{NMAX}] = {{{', '.join(_arr(r) for r in PS)}}};")
        out.append(f"  static constexpr double FS[NB][{NMAX}] = {{{', '.join(_arr(r) for r in FS)}}};")
        out.append(f"  static constexpr double MRPS[NB] = {{{', '.join(repr(float(b.mR_P * sc)) for b in bl)}}};")
        out.append(f"  static constexpr double MRFS[NB] = {{{', '.join(repr(float(b.mR_Phi * sc)) for b in bl)}}};")
        out.append(f"  static constexpr double IPS = {float(mpmath.mpf(1) / (mpmath.pi * sc))!r};   // 1 / (pi SC)")
        out.append("};")
        out.append("")
    out.append(f"// 1/k, k = 0 .. {NMAX + 1} (entries 0 and 1: 1)")
    out.append(f"constexpr double RK[{NMAX + 2}] = {{{', '.join(repr(1.0 / max(k, 1)) for k in range(NMAX + 2))}}};")
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
    out += render3d()
    out += render_triangle()
    return "\n".join(out) + "\n"


NT = 48           # array extent of the u-basis tables (>= degree + 1 of every kernel, W^5 of w4 included)


def _arrn(vals, n):
    vals = list(vals) + [0] * (n - len(vals))
    return "{" + ", ".join(repr(float(v)) for v in vals) + "}"


def render_triangle():
    """every kernel (the wall-operation kernels included): pi W per support radius in powers of u = R - r (ascending
    radii; the conditioned triangle quadrature and the chord quadrature of the W^5 kernels), the first piece of pi W
    about the origin (its regular -W'(r)/r near r = 0), the whole-support vector, and Gauss-Legendre rules of 8, 16,
    24, 32 nodes on [0, 1]."""
    import numpy as np
    from math import comb
    from warpsphboundaries_amd.kernels import kernel
    out = ["template <int KID> struct TTab;"]
    for kid, name in enumerate(ALL_NAMES):
        k = kernel(name)
        ub = {}
        for b in k.blocks:
            d = ub.setdefault(b.R, {})
            for n, c in enumerate(b.coeffs):
                for j in range(n + 1):
                    d[j] = d.get(j, 0) + c * comb(n, j) * b.R ** (n - j) * (-1) ** j
        radii = sorted(ub)
        deg = max(max(d) for d in ub.values())
        assert deg < NT and len(radii) <= 2
        lo, hi, org = k.pieces[0]
        assert lo == 0 and len(org) <= NT
        out.append(f"template <> struct TTab<{kid}> {{   // {name}")
        out.append(f"  static constexpr int NRAD = {len(radii)};")
        out.append(f"  static constexpr int DEG = {deg};")
        out.append(f"  static constexpr double RAD[NRAD] = {{{', '.join(repr(float(R)) for R in radii)}}};")
        out.append(f"  static constexpr double UB[NRAD][{NT}] = "
                   f"{{{', '.join(_arrn([ub[R].get(j, 0) for j in range(deg + 1)], NT) for R in radii)}}};")
        out.append(f"  static constexpr int NORIG = {len(org)};")
        out.append(f"  static constexpr bool REGULAR = {'true' if len(org) > 1 and org[1] == 0 else 'false'};")
        out.append(f"  static constexpr double ORIG_HI = {float(min(F(1, 8), hi))!r};")
        out.append(f"  static constexpr double ORIG[{NT}] = {_arrn(org, NT)};")
        out.append(f"  static constexpr double PLANE[12] = {{{', '.join(repr(v) for v in plane_channels(name))}}};")
        out.append("};")
        out.append("")
    # the W^5 kernels: normalisation c5 (pi W5 = c5 shape^5 with the unnormalised Wendland shape)
    from warpsphboundaries_amd.kernels import KERNEL_ID, P5_NAMES
    out.append("template <int KID> struct P5Tab;")
    for name in P5_NAMES:
        out.append(f"template <> struct P5Tab<{KERNEL_ID[name]}> {{   // {name}")
        out.append(f"  static constexpr double C = {float(kernel(name).c2_pi)!r};")
        out.append("};")
    out.append("")
    # the kernel as pieces (disk element): pi W = (HI[p] - r)^K[p] sum_n Q[p][n] r^n on [HI[p-1], HI[p]) with the
    # largest K that divides the piece exactly (the Wendland factor (1 - r)^k: Q then has positive coefficients and the
    # evaluation no cancellation), whole-support mass and int y_x^2 W; the W^5 kernels are evaluated from their
    # Wendland shape (DEG = -1)
    from warpsphboundaries_amd.kernels import plane_mass, plane_second_moment
    out.append("template <int KID> struct DTab;")
    for name in ALL_NAMES:
        k = kernel(name)
        p5 = name in P5_NAMES
        pcs = k.pieces
        rows, Ks, deg = [], [], -1
        for lo, hi, c in pcs:
            q, K = list(c), 0
            while not p5 and len(q) > 1:
                # divide by (hi - r): synthetic division at r = hi, exact in rationals
                quo, acc = [F(0)] * (len(q) - 1), F(0)
                for n in range(len(q) - 1, 0, -1):
                    acc = acc * hi + q[n]
                    quo[n - 1] = acc
                rem = acc * hi + q[0]
                if rem != 0:
                    break
                q, K = [-v for v in quo], K + 1
            rows.append(_arrn([] if p5 else q, 12))
            Ks.append(K)
            deg = -1 if p5 else max(deg, len(q) - 1)
        if len(pcs) == 1:
            rows.append(_arrn([], 12))
            Ks.append(0)
        out.append(f"template <> struct DTab<{KERNEL_ID[name]}> {{   // {name}")
        out.append(f"  static constexpr int NP = {len(pcs)};")
        out.append(f"  static constexpr int DEG = {deg};")
        out.append(f"  static constexpr double HI[2] = {{{', '.join(repr(float(hi)) for _, hi, _ in pcs)}{', 1.0' if len(pcs) == 1 else ''}}};")
        out.append(f"  static constexpr int K[2] = {{{Ks[0]}, {Ks[1]}}};")
        out.append(f"  static constexpr double Q[2][12] = {{{', '.join(rows)}}};")
        out.append(f"  static constexpr double MASS = {float(plane_mass(name))!r};")
        out.append(f"  static constexpr double SECOND = {float(plane_second_moment(name))!r};")
        out.append("};")
    out.append("")
    out.append("struct GLTab {")
    for n in (8, 16, 24, 32):
        x, w = np.polynomial.legendre.leggauss(n)
        out.append(f"  static constexpr double X{n}[{n}] = {{{', '.join(repr(float(v)) for v in (x + 1) / 2)}}};")
        out.append(f"  static constexpr double W{n}[{n}] = {{{', '.join(repr(float(v)) for v in w / 2)}}};")
    out.append("};")
    out.append("")
    return out


NM3 = 18           # array extent of the 3-D tables (>= largest power of Phi2 + 3)


def render3d():
    """3-D face closed form: per block the radial functions P, Phi, Phi2, psi_P, psi_Phi (dense by power), their
    K-chain weights g_m/(m+2) (index m), G(R) = sum g_m R^(m+2)/(m+2), and the whole-ball constants."""
    from warpsphboundaries_amd import plan3d as P3
    out = ["template <int KID> struct K3Tab;"]
    for kid, name in enumerate(NAMES):
        bl = P3.blocks3(name)
        nb = len(bl)

        def dense(d):
            assert min(d) >= 0 and max(d) < NM3, d
            return [d.get(m, 0) for m in range(NM3)]

        def q(d):
            return {m: c / (m + 2) for m, c in d.items()}

        def GR(d, R):
            return sum(c * R ** (m + 2) / (m + 2) for m, c in d.items())
        deg = max(max(max(b.Phi2), max(b.psiF) + 2, max(b.Phi) + 2) for b in bl)
        out.append(f"template <> struct K3Tab<{kid}> {{   // {name}")
        out.append(f"  static constexpr int NB = {nb};")
        out.append(f"  static constexpr int DEG = {deg};")
        out.append(f"  static constexpr double R[NB] = {{{', '.join(repr(float(b.R)) for b in bl)}}};")
        for key, attr in (("P", "P"), ("F", "Phi"), ("F2", "Phi2"), ("pP", "psiP"), ("pF", "psiF")):
            rows = [dense(getattr(b, attr)) for b in bl]
            out.append(f"  static constexpr double {key}[NB][{NM3}] = {{{', '.join(_arr3(r) for r in rows)}}};")
            if key != "F2":
                rows = [dense(q(getattr(b, attr))) for b in bl]
                out.append(f"  static constexpr double Q{key}[NB][{NM3}] = {{{', '.join(_arr3(r) for r in rows)}}};")
                out.append(f"  static constexpr double GR{key}[NB] = "
                           f"{{{', '.join(repr(float(GR(getattr(b, attr), b.R))) for b in bl)}}};")
        # Hessian: the radial function W'(r)/r = sum_k k P_k r^(k-2) has the K-chain weights (k P_k)/k = P_k at index
        # k = m + 2 (k >= 1), G(R) = sum_(k>=1) P_k R^k; the in-plane edge integrand is W(r) - W(R) without its
        # constant term: dI_k weights P_k (k >= 1) and -G(R) at k = 0.  HW holds both (index k, -G(R) at 0).
        grl = [sum(c * b.R ** k for k, c in b.P.items() if k >= 1) for b in bl]
        hw = [[(-grl[i] if k == 0 else b.P.get(k, 0)) for k in range(NM3)] for i, b in enumerate(bl)]
        out.append(f"  static constexpr double HW[NB][{NM3}] = {{{', '.join(_arr3(r) for r in hw)}}};")
        out.append(f"  static constexpr double GRL[NB] = {{{', '.join(repr(float(g)) for g in grl)}}};")
        out.append(f"  static constexpr double MR[NB] = {{{', '.join(repr(float(b.mR)) for b in bl)}}};")
        out.append(f"  static constexpr double MRF[NB] = {{{', '.join(repr(float(b.mRF)) for b in bl)}}};")
        out.append(f"  static constexpr double MRSUM = {float(sum(b.mR for b in bl))!r};")
        out.append(f"  static constexpr double MRFSUM = {float(sum(b.mRF for b in bl))!r};")
        out.append("};")
        out.append("")
    return out


def _arr3(vals):
    return "{" + ", ".join(repr(float(v)) for v in vals) + "}"


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
