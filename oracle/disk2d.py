"""Solid disks (fibre bundles): numpy FP64 restatement of the reference's disk element.  TEST INFRASTRUCTURE ONLY.

* ``cells`` / ``disk_slots``: the static cell list over the disk centres and the fixed-capacity slot fill
  (reference ``scene/scene.py:869-888`` ``DiskArrayRep.cells``, ``scene/fixedadj.py:199-222`` ``_disk_slots_kernel``),
  operation for operation (the distance predicate with separately rounded products).
* ``circle_channels``: the five channels lam, m1x, g0x, g1xx, g1yy of one disk by the divergence theorem on its circle
  (reference ``edge/disk.py:36-157``): exact Chebyshev forms of the kernel pieces and of the radial potentials, angle
  panels cut where the circle meets a knot radius and refined geometrically (ratio 4) towards the closest approach,
  24 Gauss-Legendre nodes per panel, the regular unsplit potential M(r)/r^2 over the full circle for lam.
* The reference's disk element has no second moments; ``circle_channels`` adds them by the same identity with the
  regular potential Psi(r) = r^-4 int_0^r t^3 W dt:  div(y_d y_j y Psi) = y_d y_j (4 Psi + r Psi') = y_d y_j W, so
  m2_dj = oint (n . y) y_d y_j Psi dl over the full circle (Psi -> W(0)/4 at r = 0, int_0^1 t^3 W / r^4 outside).
* ``assemble``: the nine body-frame channels of a (query, disk) pair from the five (``edge/warpdisk.py:70-95``), plus
  the second moments, in the twelve-channel device order.

Pinned by (``tests/test_disk2d_cpu.py``): mpmath radial integrals at 30 digits (every channel, every kernel family),
the reference's own checks (the support inside a large disk gives lam = 1, g1 = I, no gradient; nothing beyond reach;
smooth across the circle, 24 vs 48 nodes converged), the whole-support second moment, the polygon limit through
``oracle.edge2d``, and brute force for the slots.
"""
from fractions import Fraction
from functools import lru_cache

import numpy as np
import sympy as sp

from . import kernels as K

NODES = 24
MAXPANELS = 26


def _exact_pieces(name):
    """[(lo, hi, {n: Fraction})] of pi W, ascending pieces."""
    c2pi, pcs = K._SPEC[name]
    out = []
    for (lo, hi), e in pcs:
        co = K._coeffs(c2pi * e)
        out.append((Fraction(lo), Fraction(hi), {n: c for n, c in enumerate(co) if c != 0}))
    return out


def _cheb_exact(P, hi):
    """exact Chebyshev coefficients of r -> sum P[n] r^n on [0, hi] in x = 2 r / hi - 1 (r = hi (x + 1) / 2)."""
    x = sp.Symbol("x")
    poly = sp.Poly(sp.expand(sum(sp.Rational(c.numerator, c.denominator) * (sp.Rational(hi.numerator, hi.denominator)
                                                                            * (x + 1) / 2) ** n for n, c in P.items())),
                   x)
    deg = poly.degree() if P else 0
    mono = [Fraction(0)] * (deg + 1)
    for (k,), c in poly.terms():
        mono[k] = Fraction(int(sp.fraction(c)[0]), int(sp.fraction(c)[1]))
    # x^k in Chebyshev polynomials, exactly
    cheb = [Fraction(0)] * (deg + 1)
    xp = [Fraction(1)]
    for k in range(deg + 1):
        for j, a in enumerate(xp):
            cheb[j] += mono[k] * a
        nxt = [Fraction(0)] * (len(xp) + 1)
        for j, a in enumerate(xp):
            if j == 0:
                nxt[1] += a
            else:
                nxt[j + 1] += a / 2
                nxt[j - 1] += a / 2
        xp = nxt
    return np.array([float(v) for v in cheb])


@lru_cache(maxsize=None)
def _pieces(name):
    """per piece (lo, hi, cheb(pi W), cheb(pi Phi), cheb(regular M / r^2 part), off, cheb(regular Psi part), off4)
    and the totals M_s = int_0^1 t pi W dt, Q_s = int_0^1 t^3 pi W dt (reference edge/disk.py:36-57)."""
    raw = []
    acc = Fraction(0)
    acc4 = Fraction(0)
    for lo, hi, prof in _exact_pieces(name):
        prim = {n + 2: v / (n + 2) for n, v in prof.items()}
        prim4 = {n + 4: v / (n + 4) for n, v in prof.items()}
        off = acc - sum(v * lo ** n for n, v in prim.items())
        acc = off + sum(v * hi ** n for n, v in prim.items())
        off4 = acc4 - sum(v * lo ** n for n, v in prim4.items())
        acc4 = off4 + sum(v * hi ** n for n, v in prim4.items())
        raw.append((lo, hi, prof, prim, off, off4))
    out = []
    for lo, hi, prof, prim, off, off4 in raw:
        compact = dict(prim)
        compact[0] = compact.get(0, Fraction(0)) + off - acc
        regular = {n: v / (n + 2) for n, v in prof.items()}
        regular4 = {n: v / (n + 4) for n, v in prof.items()}
        out.append((float(lo), float(hi), _cheb_exact(prof, hi), _cheb_exact(compact, hi), _cheb_exact(regular, hi),
                    float(off), _cheb_exact(regular4, hi), float(off4)))
    return out, float(acc), float(acc4)


def kernel_profiles(name, r):
    """(W, Phi) at radii r (h = 1): W = (pi W) / pi, Phi = (M(r) - M_s) / pi (reference edge/disk.py:60-74)."""
    pcs, _, _ = _pieces(name)
    r = np.asarray(r, dtype=float)
    W = np.zeros_like(r)
    Phi = np.zeros_like(r)
    for lo, hi, wc, pc, _, _, _, _ in pcs:
        sel = (r >= lo) & (r < hi)
        x = 2.0 * np.clip(r, 0.0, hi) / hi - 1.0
        W = np.where(sel, np.polynomial.chebyshev.chebval(x, wc), W)
        Phi = np.where(sel, np.polynomial.chebyshev.chebval(x, pc), Phi)
    return W / np.pi, Phi / np.pi


def _angle(r, D, R):
    """phi in [0, pi] where the circle meets the radius r (reference edge/disk.py:77-79)."""
    return np.arccos(np.clip((r * r - D * D - R * R) / (2.0 * D * R), -1.0, 1.0))


def _over_rk(pcs, total, r, which):
    """M(r) / r^2 (which = 2) or Psi(r) r^4 / r^4 (which = 4) with the finite origin limit (reference
    edge/disk.py:82-101)."""
    value = np.zeros_like(r)
    outside = r >= pcs[-1][1]
    np.divide(total, r ** which, out=value, where=outside)
    for lo, hi, _, _, qc2, off2, qc4, off4 in pcs:
        qc, off = (qc2, off2) if which == 2 else (qc4, off4)
        sel = (r >= lo) & (r < hi)
        poly = np.polynomial.chebyshev.chebval(2.0 * np.clip(r, 0.0, hi) / hi - 1.0, qc)
        corr = np.zeros_like(r)
        if off != 0:
            np.divide(off, r ** which, out=corr, where=sel)
        value = np.where(sel, corr + poly, value)
    return value


def circle_channels(name, D, R, nodes=NODES):
    """dict(lam, m1x, g0x, g1xx, g1yy, m2xx, m2yy) at broadcast D, R (h = 1, the centre at (D, 0) from the query),
    float64 (reference edge/disk.py:104-157; m2 by the Psi identity of the module docstring)."""
    D, R = np.broadcast_arrays(np.asarray(D, dtype=float), np.asarray(R, dtype=float))
    shape = D.shape
    centre = (D == 0).ravel()
    D = np.maximum(D, 1e-12).ravel()
    R = R.ravel()
    M = len(D)
    pcs, Ms, Qs = _pieces(name)
    knots = [p[1] for p in pcs]
    gx, gw = np.polynomial.legendre.leggauss(nodes)
    edges = [np.full(M, np.pi)]
    w = np.maximum(np.abs(D - R) / np.sqrt(D * R), 1e-12)
    for k in range(MAXPANELS):
        edges.append(np.maximum(np.pi - w * 4.0 ** k, 0.0))
    for kn in knots:
        edges.append(_angle(kn, D, R))
    E = np.sort(np.stack(edges, 1), 1)
    lo, hi = E[:, :-1], E[:, 1:]
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    phi = mid[..., None] + half[..., None] * gx
    wt = half[..., None] * gw
    Dm, Rm = D[:, None, None], R[:, None, None]
    hc = np.cos(phi / 2.0)
    yx, yy = (Dm - Rm) + 2.0 * Rm * hc ** 2, Rm * np.sin(phi)
    r = np.hypot(Dm - Rm, 2.0 * np.sqrt(Dm) * np.sqrt(Rm) * hc)
    nx, ny = np.cos(phi), np.sin(phi)
    W, Phi = kernel_profiles(name, r)
    ndoty = (Rm - Dm) + 2.0 * Dm * hc ** 2
    lam = 2.0 * (wt * ndoty * _over_rk(pcs, Ms, r, 2) / np.pi * Rm).sum((1, 2))
    m1x = 2.0 * (wt * nx * Phi * Rm).sum((1, 2))
    g0x = -2.0 * (wt * nx * W * Rm).sum((1, 2))
    g1xx = -2.0 * (wt * nx * yx * W * Rm).sum((1, 2)) + lam
    g1yy = -2.0 * (wt * ny * yy * W * Rm).sum((1, 2)) + lam
    psi = _over_rk(pcs, Qs, r, 4) / np.pi
    m2xx = 2.0 * (wt * ndoty * yx * yx * psi * Rm).sum((1, 2))
    m2yy = 2.0 * (wt * ndoty * yy * yy * psi * Rm).sum((1, 2))
    if centre.any():
        rc = R[centre]
        wc, _ = kernel_profiles(name, rc)
        within = rc < pcs[-1][1]
        mass = np.full_like(rc, 2.0 * Ms)
        flux = np.zeros_like(rc)
        sec = np.full_like(rc, Qs)
        mass[within] = 2.0 * rc[within] ** 2 * _over_rk(pcs, Ms, rc[within], 2)
        flux[within] = np.pi * rc[within] ** 2 * wc[within]
        sec[within] = rc[within] ** 4 * _over_rk(pcs, Qs, rc[within], 4)
        lam[centre] = mass
        m1x[centre] = 0.0
        g0x[centre] = 0.0
        g1xx[centre] = mass - flux
        g1yy[centre] = g1xx[centre]
        m2xx[centre] = sec        # int y_x^2 W over the disk |y| < R about the query: (1/pi) pi int_0^R r^3 pi W dr
        m2yy[centre] = sec
    out = dict(lam=lam, m1x=m1x, g0x=g0x, g1xx=g1xx, g1yy=g1yy, m2xx=m2xx, m2yy=m2yy)
    return {k: v.reshape(shape) for k, v in out.items()}


def assemble(ch, y):
    """[P, 12] body-frame channels (device order lam, m1, g0, g1 (row-major), m2 (xx, xy, yy)) of (query, disk) pairs
    from ``circle_channels`` at the offsets y [P, 2] = (centre - query) / h (reference edge/warpdisk.py:70-95)."""
    y = np.asarray(y, dtype=float)
    D = np.sqrt(y[:, 0] * y[:, 0] + y[:, 1] * y[:, 1])
    big = D > 1e-12
    cx = np.where(big, y[:, 0] / np.where(big, D, 1.0), 1.0)
    cy = np.where(big, y[:, 1] / np.where(big, D, 1.0), 0.0)
    tx, ty = -cy, cx
    out = np.zeros((len(D), 12))
    out[:, 0] = ch["lam"]
    out[:, 1], out[:, 2] = ch["m1x"] * cx, ch["m1x"] * cy
    out[:, 3], out[:, 4] = ch["g0x"] * cx, ch["g0x"] * cy
    out[:, 5] = ch["g1xx"] * cx * cx + ch["g1yy"] * tx * tx
    out[:, 6] = ch["g1xx"] * cx * cy + ch["g1yy"] * tx * ty
    out[:, 7] = out[:, 6]
    out[:, 8] = ch["g1xx"] * cy * cy + ch["g1yy"] * ty * ty
    out[:, 9] = ch["m2xx"] * cx * cx + ch["m2yy"] * tx * tx
    out[:, 10] = ch["m2xx"] * cx * cy + ch["m2yy"] * tx * ty
    out[:, 11] = ch["m2xx"] * cy * cy + ch["m2yy"] * ty * ty
    return out


# ------------------------------------------------------------------------------------------------------- slot list
def cells(centres, radii, support_max):
    """(lo [2], cell, nx, ny, start [nx ny + 1], items [M], K) of the centres (reference scene/scene.py:869-888)."""
    c = np.asarray(centres, dtype=float).reshape(-1, 2)
    rad = np.asarray(radii, dtype=float).reshape(-1)
    cell = float(support_max) + float(rad.max())
    lo = c.min(0) - cell
    hi = c.max(0) + cell
    dims = np.maximum(np.floor((hi - lo) / cell).astype(np.int64) + 1, 1)
    nx, ny = int(dims[0]), int(dims[1])
    ci = np.floor((c - lo) / cell).astype(np.int64)
    cid = ci[:, 0] * ny + ci[:, 1]
    order = np.argsort(cid, kind="stable")
    counts = np.bincount(cid, minlength=nx * ny).reshape(nx, ny)
    pad = np.pad(counts, 1)
    block = sum(pad[1 + dx:1 + dx + nx, 1 + dy:1 + dy + ny] for dx in (-1, 0, 1) for dy in (-1, 0, 1))
    Kmax = max(1, int(block.max()))
    start = np.concatenate([[0], np.cumsum(counts.reshape(-1))])
    return lo, cell, nx, ny, start.astype(np.int32), order.astype(np.int32), Kmax


def disk_slots(cl, centres, radii, lpos, lsup, valid=None):
    """slots [N, K] (-1 padded) of the rows (reference scene/fixedadj.py:199-222): the 3 x 3 cells around the row's
    cell, x outer and y inner, cell-list order, kept when sqrt(ex^2 + ey^2) < h + R."""
    lo, cell, nx, ny, start, items, Kmax = cl
    c = np.asarray(centres, dtype=float).reshape(-1, 2)
    rad = np.asarray(radii, dtype=float).reshape(-1)
    N = len(lpos)
    out = np.full((N, Kmax), -1, dtype=np.int32)
    for r in range(N):
        if valid is not None and not valid[r]:
            continue
        px, py, h = float(lpos[r, 0]), float(lpos[r, 1]), float(lsup[r])
        cx = int(np.floor((px - lo[0]) / cell))
        cy = int(np.floor((py - lo[1]) / cell))
        n = 0
        for ddx in (-1, 0, 1):
            for ddy in (-1, 0, 1):
                ix, iy = cx + ddx, cy + ddy
                if 0 <= ix < nx and 0 <= iy < ny:
                    cid = ix * ny + iy
                    for i in range(start[cid], start[cid + 1]):
                        m = int(items[i])
                        ex = px - c[m, 0]
                        ey = py - c[m, 1]
                        if np.sqrt(ex * ex + ey * ey) < h + rad[m]:
                            if n < Kmax:
                                out[r, n] = m
                                n += 1
    return out
