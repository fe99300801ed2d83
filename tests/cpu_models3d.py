"""numpy model of the device 3-D face closed form (compact face terms + winding x whole-ball vector).

Test helper only: it mirrors the arithmetic structure of ``csrc/mesh3d_device.h`` (per pair, per block, per edge:
one clipped chord, the dI_k recurrence, two chord angles, the w-weighted K_m chains) so that the algebra is checked
against the oracle on the CPU before and independently of the HIP build.
"""
from __future__ import annotations

import numpy as np

from warpsphboundaries_amd import plan3d as P3

TWO_PI = 2.0 * np.pi


def _f(c):
    return float(c)


def _tables(name):
    out = []
    for b in P3.blocks3(name):
        def lst(d):
            return {m: float(c) for m, c in d.items()}
        GR = {}
        for key, g in (("P", b.P), ("F", b.Phi), ("pP", b.psiP), ("pF", b.psiF)):
            GR[key] = float(sum(c * b.R ** (m + 2) / (m + 2) for m, c in g.items()))
        out.append(dict(R=float(b.R), P=lst(b.P), F=lst(b.Phi), F2=lst(b.Phi2), pP=lst(b.psiP), pF=lst(b.psiF),
                        mR=float(b.mR), mRF=float(b.mRF), GR=GR))
    return out


def _dlog(l, u, rl, ru, c2):
    """asinh(u/c) - asinh(l/c) as one log of a ratio (no cancellation for chords on one side of the foot)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        pos = np.log((u + ru) / (l + rl))
        neg = np.log((rl - l) / (ru - u))
        c = np.sqrt(c2)
        mid = np.arcsinh(u / c) - np.arcsinh(l / c)
    return np.where(l >= 0, pos, np.where(u <= 0, neg, mid))


def face_pairs(V, x, h, name):
    """per (query, face) pair: the 23 accumulator contributions (M0, M1[3], VPhi, M2[6], G[3], g1[9]) in units of h,
    before the winding term, 1/pi and physical scaling.  V[P,3,3], x[P,3], h[P]."""
    V = np.asarray(V, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    h = np.broadcast_to(np.asarray(h, dtype=np.float64), (len(V),))
    p = (V - x[:, None, :]) / h[:, None, None]
    Np = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    n = Np / np.linalg.norm(Np, axis=1)[:, None]
    z = np.einsum("ij,ij->i", n, p[:, 0])
    a = np.abs(z)
    sz = np.sign(z)
    E = []
    for e in range(3):
        v, u = p[:, e], p[:, (e + 1) % 3]
        D = u - v
        t = D / np.linalg.norm(D, axis=1)[:, None]
        mu = np.cross(t, n)
        w = np.einsum("ij,ij->i", mu, v)
        lo = np.einsum("ij,ij->i", t, v)
        hi = np.einsum("ij,ij->i", t, u)
        E.append((w, lo, hi, mu))
    W = np.stack([e[0] for e in E])
    theta_all = np.zeros(len(V))
    for w, lo, hi, _ in E:
        b = np.abs(w)
        theta_all += np.where(w != 0, np.sign(w) * np.arctan2(b * (hi - lo), w * w + lo * hi), 0.0)
    Theta = np.where((W > 0).all(axis=0), TWO_PI, np.where((W < 0).any(axis=0), 0.0, theta_all))

    sP = np.zeros(len(V)); sF = np.zeros(len(V)); zs0 = np.zeros(len(V)); zsV = np.zeros(len(V))
    l1 = np.zeros((len(V), 3)); l2 = np.zeros((len(V), 3))
    for T in _tables(name):
        R = T["R"]
        act = a < R
        S = {k: np.zeros(len(V)) for k in ("P", "F", "pP", "pF")}
        zS3 = np.zeros(len(V))
        M = max(max(T["F2"]), max(T["pF"]) + 2, max(T["F"]) + 2) + 1
        for w, lo, hi, mu in E:
            c2 = w * w + a * a
            ch = act & (c2 < R * R)
            L = np.sqrt(np.where(ch, R * R - c2, 0.0))
            l = np.maximum(lo, -L)
            u = np.minimum(hi, L)
            ch = ch & (l < u)
            l = np.where(ch, l, 0.0)
            u = np.where(ch, u, 0.0)
            rl = np.sqrt(l * l + c2)
            ru = np.sqrt(u * u + c2)
            dIm1 = np.where(ch & (c2 > 0), _dlog(l, u, rl, ru, c2), 0.0)
            dI = {-1: dIm1, 0: u - l}
            ul, uu = l.copy(), u.copy()         # s r^k at both ends
            for k in range(1, M + 1):
                ul = ul * rl
                uu = uu * ru
                prev = c2 * dI[k - 2] if k >= 2 else np.where(c2 > 0, c2 * dIm1, 0.0)
                dI[k] = (uu - ul + k * prev) / (k + 1)
            b = np.abs(w)
            sw = np.sign(w)
            th = np.where(ch, np.arctan2(b * (u - l), b * b + u * l), 0.0)
            # u r_l - l r_u without cancellation
            with np.errstate(divide="ignore", invalid="ignore"):
                same = c2 * (u - l) * (u + l) / (u * rl + l * ru)
            diff = np.where((l >= 0) | (u <= 0), np.where(u * rl + l * ru != 0, same, 0.0), u * rl - l * ru)
            ph = np.where(ch & (b > 0), np.arctan2(a * b * diff, b * b * ru * rl + a * a * u * l), 0.0)
            wK = {0: sw * th, -1: sw * ph}             # w dK_0 and w a dK_-1
            wK[1] = w * dIm1 + a * wK[-1]
            for m in range(2, M + 3):
                wK[m] = w * dI[m - 2] + a * a * wK[m - 2]
            for key in S:
                g = T[key]
                S[key] += sum(c / (m + 2) * wK[m + 2] for m, c in g.items()) - T["GR"][key] * sw * th
            zS3 += sz * sw * (-ph + (a / R) * th)
            lF = sum(c * dI[m] for m, c in T["F"].items())
            lF2 = sum(c * dI[m] for m, c in T["F2"].items())
            l1 += mu * lF[:, None]
            l2 += mu * lF2[:, None]
        for key in S:
            g = T[key]
            Ga = sum(c * a ** (m + 2) / (m + 2) for m, c in g.items())
            S[key] = np.where(act, S[key] + (T["GR"][key] - Ga) * Theta, 0.0)
        zS3 = np.where(act, zS3 + sz * (1.0 - a / R) * Theta, 0.0)
        sP += S["P"]
        sF += S["F"]
        zs0 += z * S["pP"] - T["mR"] * zS3
        zsV += z * S["pF"] - T["mRF"] * zS3
    out = np.zeros((len(V), 23))
    out[:, 0] = zs0
    out[:, 1:4] = n * sF[:, None]
    out[:, 4] = zsV
    k = 5
    for i in range(3):
        for kk in range(i, 3):
            out[:, k] = n[:, i] * (z * n[:, kk] * sF + l2[:, kk])
            k += 1
    out[:, 11:14] = -n * sP[:, None]
    k = 14
    for d in range(3):
        for j in range(3):
            out[:, k] = -n[:, j] * (z * n[:, d] * sP + l1[:, d])
            k += 1
    return out


def finalize(acc, chi, h, name):
    """accumulators [Q,23] (summed over faces) + winding chi [Q] -> the 22 channels in request order, physical."""
    T = _tables(name)
    mR = sum(t["mR"] for t in T)
    mRF = sum(t["mRF"] for t in T)
    acc = np.array(acc, dtype=np.float64)
    M0 = acc[:, 0] + 4 * np.pi * chi * mR
    V = acc[:, 4] + 4 * np.pi * chi * mRF
    out = np.zeros((len(acc), 22))
    out[:, 0] = M0
    out[:, 1:4] = acc[:, 1:4]
    M2 = acc[:, 5:11].copy()
    for k in (0, 3, 5):
        M2[:, k] -= V
    out[:, 4:10] = M2
    out[:, 10:13] = acc[:, 11:14]
    g1 = acc[:, 14:23].copy()
    for k in (0, 4, 8):
        g1[:, k] += M0
    out[:, 13:22] = g1
    out /= np.pi
    h = np.broadcast_to(np.asarray(h, dtype=np.float64), (len(acc),))
    return out * h[:, None] ** np.array(P3.H_POWER3, dtype=np.float64)[None, :]


def closed_surface22(vertices, faces, queries, h, name, chi):
    """every face for every query (the compact terms vanish outside the support)."""
    V = np.asarray(vertices, dtype=np.float64)[np.asarray(faces)]
    X = np.asarray(queries, dtype=np.float64).reshape(-1, 3)
    Q, Fn = len(X), len(V)
    hq = np.broadcast_to(np.asarray(h, dtype=np.float64), (Q,))
    pairs = face_pairs(np.broadcast_to(V, (Q,) + V.shape).reshape(-1, 3, 3), np.repeat(X, Fn, axis=0),
                       np.repeat(hq, Fn), name)
    acc = pairs.reshape(Q, Fn, 23).sum(axis=1)
    return finalize(acc, chi, hq, name)
