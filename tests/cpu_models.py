"""numpy models of the device algorithms, used on the CPU to check the product's plan tables and the algebra of
the closed-form kernel before any GPU run (test code only)."""
from math import comb

import numpy as np

from warpsphboundaries_amd import plan as PL

INV_PI = 1.0 / np.pi


def _geo(p, q):
    d = q - p
    ell = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
    t0, t1 = d[:, 0] / ell, d[:, 1] / ell
    z = (p[:, 0] * q[:, 1] - p[:, 1] * q[:, 0]) / ell
    s0 = (d[:, 0] * p[:, 0] + d[:, 1] * p[:, 1]) / ell
    s1 = (d[:, 0] * q[:, 0] + d[:, 1] * q[:, 1]) / ell
    return t0, t1, t1, -t0, z, s0, s1


def _Im(m, s, z):
    zz = z * z
    r2 = s * s + zz
    r = np.sqrt(r2)
    val = s.copy()
    idx = 0
    if m % 2 == 1:
        safe = np.where(z == 0, 1.0, np.abs(z))
        tail = np.where(z == 0, 0.0, zz * np.arcsinh(s / safe))
        val = (s * r + tail) / 2
        idx = 1
    while idx < m:
        idx += 2
        rp = r2 ** (idx // 2) if idx % 2 == 0 else r * r2 ** ((idx - 1) // 2)
        val = (s * rp + idx * zz * val) / (idx + 1)
    return val


def _S(j, m, s, z):
    zz = z * z
    tot = 0.0
    if j % 2 == 0:
        h = j // 2
        for i in range(h + 1):
            tot = tot + comb(h, i) * (-1) ** (h - i) * zz ** (h - i) * _Im(m + 2 * i, s, z)
    else:
        h = (j - 1) // 2
        r2 = s * s + zz
        r = np.sqrt(r2)
        for i in range(h + 1):
            mm = m + 2 * i
            rp = r2 ** ((mm + 2) // 2) if mm % 2 == 0 else r * r2 ** ((mm + 1) // 2)
            tot = tot + comb(h, i) * (-1) ** (h - i) * zz ** (h - i) * rp / (mm + 2)
    return tot


def _dangle(z, lo, hi):
    safe = np.where(z == 0, 1.0, z)
    return np.where(z == 0, 0.0, np.arctan2(safe * (hi - lo), z * z + lo * hi))


def plan_interpreter(pl: PL.SurfacePlan, p, q):
    """the flattened-plan interpreter (the plain device form) in numpy: [P, 12]."""
    a = pl.arrays()
    t0, t1, n0, n1, z, s0, s1 = _geo(p, q)
    az = np.abs(z)
    out = np.zeros((len(p), 12))
    for (ch, i, aa, bb, Ri, c0, c1) in a["E"][:a["nE"]]:
        R = a["radii"][Ri]
        L = np.sqrt(np.maximum((R - az) * (R + az), 0))
        lo, hi = np.maximum(s0, -L), np.minimum(s1, L)
        m = (az < R) & (lo < hi)
        lo, hi = np.where(m, lo, 0), np.where(m, hi, 0)
        tot = 0.0
        for j in range(aa + bb + 1):
            yc = 0.0
            for ja in range(aa + 1):
                jb = j - ja
                if 0 <= jb <= bb:
                    yc = yc + comb(aa, ja) * t0 ** ja * (z * n0) ** (aa - ja) * comb(bb, jb) * t1 ** jb * (z * n1) ** (bb - jb)
            inner = 0.0
            for k in range(c0, c1):
                inner = inner + a["cc"][k] * (_S(j, a["cn"][k], hi, z) - _S(j, a["cn"][k], lo, z))
            tot = tot + yc * inner
        ni = n1 if i == 1 else n0
        out[:, ch] += np.where(m, ni * tot, 0)
    for t, (ch, Ri, c0, c1) in enumerate(a["V"][:a["nV"]]):
        R = a["radii"][Ri]
        L = np.sqrt(np.maximum((R - az) * (R + az), 0))
        lo, hi = np.maximum(s0, -L), np.minimum(s1, L)
        m = (az < R) & (lo < hi)
        lo, hi = np.where(m, lo, 0), np.where(m, hi, 0)
        poly = 0.0
        for k in range(c0, c1):
            poly = poly + a["cc"][k] / (a["cn"][k] + 2) * (_S(0, a["cn"][k], hi, z) - _S(0, a["cn"][k], lo, z))
        out[:, ch] += np.where(m, z * poly - a["mR"][t] * _dangle(z, lo, hi), 0)
    return out * INV_PI


def closed_form(kname, p, q, h=1.0):
    """the shared-primitive closed form of the optimized device forms (same geometry path, recurrences and algebra),
    for edges p -> q relative to the query (physical units; support h): [P, 12]."""
    ux, uy, vx, vy = p[:, 0], p[:, 1], q[:, 0], q[:, 1]
    dx, dy = vx - ux, vy - uy
    ln = np.sqrt(dx * dx + dy * dy)
    tx, ty, iell = dx / ln, dy / ln, 1.0 / ln
    ih = 1.0 / h
    cr = ux * vy - uy * vx
    mag = np.abs(ux * vy) + np.abs(uy * vx)
    slow = ~((mag > 1e-280) & (np.abs(cr) > 1.7763568394002505e-15 * mag))
    cn = (ux / h) * (vy / h) - (uy / h) * (vx / h)
    cr = np.where(slow, cn * h * h, cr)
    z = cr * ih * iell
    px, py, qx, qy = ux * ih, uy * ih, vx * ih, vy * ih
    s0, s1 = tx * px + ty * py, tx * qx + ty * qy
    n0, n1 = ty, -tx
    az, zz = np.abs(z), z * z
    r0sq, r1sq = px * px + py * py, qx * qx + qy * qy
    c = np.zeros((len(p), 12))
    for b in PL.block_profiles(kname):
        R = float(b.R)
        P = [float(x) for x in b.P]
        Fp = [float(x) for x in b.Phi]
        QP = [float(x) / (n + 2) for n, x in enumerate(b.P)]
        QF = [float(x) / (n + 2) for n, x in enumerate(b.Phi)]
        NM = max(len(P), len(Fp)) - 1
        clo = (s0 < 0) & (r0sq > R * R)
        chi = (s1 > 0) & (r1sq > R * R)
        L = np.sqrt(np.maximum((R - az) * (R + az), 0))
        lo, hi = np.where(clo, -L, s0), np.where(chi, L, s1)
        m = (az < R) & (lo < hi)
        lo, hi = np.where(m, lo, 0.0), np.where(m, hi, 1.0)
        rl = np.where(clo, R, np.sqrt(r0sq))
        rh = np.where(chi, R, np.sqrt(r1sq))
        rl, rh = np.where(m, rl, 1.0), np.where(m, rh, 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            Nh, Dh = np.where(hi >= 0, hi + rh, zz), np.where(hi >= 0, 1.0, rh - hi)
            Nl, Dl = np.where(lo >= 0, lo + rl, zz), np.where(lo >= 0, 1.0, rl - lo)
            dA = np.where(z != 0, np.log((Nh * Dl) / (Dh * Nl)), 0.0)
        dI = [hi - lo, 0.5 * (hi * rh - lo * rl + zz * dA)]
        ph, pl = rh.copy(), rl.copy()
        AP = VP = AF = VF = BP = BF = 0.0
        for mm in range(0, NM + 3):
            if mm >= 2:
                ph, pl = ph * rh, pl * rl
                if mm <= NM:
                    dI.append(((hi * ph - lo * pl) + mm * zz * dI[mm - 2]) * (1.0 / (mm + 1)))
            if mm <= NM:
                g = lambda arr: arr[mm] if mm < len(arr) else 0.0
                AP, VP, AF, VF = AP + g(P) * dI[mm], VP + g(QP) * dI[mm], AF + g(Fp) * dI[mm], VF + g(QF) * dI[mm]
            if mm >= 2:
                dR = ph - pl
                k = mm - 2
                BP = BP + (QP[k] if k < len(QP) else 0.0) * dR
                BF = BF + (QF[k] if k < len(QF) else 0.0) * dR
        th = _dangle(z, lo, hi)
        lam = z * VP - float(b.mR_P) * th
        valF = z * VF - float(b.mR_Phi) * th
        YxP, YyP = z * n0 * AP + tx * BP, z * n1 * AP + ty * BP
        YxF, YyF = z * n0 * AF + tx * BF, z * n1 * AF + ty * BF
        blk = np.stack([lam, n0 * AF, n1 * AF, -n0 * AP, -n1 * AP, lam - n0 * YxP, -n1 * YxP, -n0 * YyP, lam - n1 * YyP,
                        n0 * YxF - valF, n0 * YyF, n1 * YyF - valF], 1)
        c += np.where(m[:, None], blk, 0.0)
    return c * INV_PI
