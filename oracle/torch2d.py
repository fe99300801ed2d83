"""Torch-CPU restatement of the reference's CPU edge backend for the twelve channels of a scene -- the CPU baseline
of ``bench.py``.  TEST / MEASUREMENT INFRASTRUCTURE (see oracle/__init__): never part of the product.

The reference's ``edge/torch2d.py`` evaluates the exact closed forms with torch tensors, vectorised over elements; its
scene layer feeds it the (query, edge) pairs of the cell-list adjacency and adds indicator x plane vector
(``scene/scene.py:384-416``).  Here the same work runs on torch CPU tensors in FP64, using every host thread torch is
given:

* the slot fill of the static cell list (``scene/fixedadj.py:67-97``), vectorised over all (row, cell item) candidates;
* the winding indicator (``scene/fixedadj.py:98-147``): the cached value of an empty cell, the background outside the
  grid, else the half-open crossing count over the row's y-bin, vectorised over all (row, bin item) pairs (rows on an
  edge take the angle sum over every edge);
* per pair the twelve edge-local channels by the block recursion of ``edge/core.py:68-212`` (the formulation of
  ``oracle.edge2d``, here with torch operations), row sums by ``index_add_``.
"""
from __future__ import annotations

import math
from math import comb

import numpy as np
import torch

from . import adjacency as oadj
from . import edge2d as oedge
from .kernels import blocks as kernel_blocks

F64 = torch.float64


class _Geo:
    def __init__(self, p, q):
        d = q - p
        ell = torch.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
        self.ok = ell > 0
        safe = torch.where(self.ok, ell, torch.ones_like(ell))
        self.t0, self.t1 = d[:, 0] / safe, d[:, 1] / safe
        self.n0, self.n1 = self.t1, -self.t0
        self.z = (p[:, 0] * q[:, 1] - p[:, 1] * q[:, 0]) / safe
        self.s0 = (d[:, 0] * p[:, 0] + d[:, 1] * p[:, 1]) / safe
        self.s1 = (d[:, 0] * q[:, 0] + d[:, 1] * q[:, 1]) / safe
        self._chords = {}

    def chord(self, R):
        v = self._chords.get(float(R))
        if v is None:
            Rf = float(R)
            az = self.z.abs()
            L = torch.sqrt(torch.clamp((Rf - az) * (Rf + az), min=0.0))
            lo = torch.maximum(self.s0, -L)
            hi = torch.minimum(self.s1, L)
            m = (az < Rf) & (lo < hi) & self.ok
            zero = torch.zeros_like(lo)
            v = (torch.where(m, lo, zero), torch.where(m, hi, zero), m)
            self._chords[float(R)] = v
        return v


def _I_all(M, s, z):
    zz = z * z
    r2 = s * s + zz
    r = torch.sqrt(r2)
    out = [s]
    if M >= 1:
        zero = z == 0
        safe = torch.where(zero, torch.ones_like(z), z.abs())
        tail = torch.where(zero, torch.zeros_like(z), zz * torch.asinh(s / safe))
        out.append((s * r + tail) / 2)
    for m in range(2, M + 1):
        rp = r2 ** (m // 2) if m % 2 == 0 else r * r2 ** ((m - 1) // 2)
        out.append((s * rp + m * zz * out[m - 2]) / (m + 1))
    return out


def _S(j, m, s, z):
    zz = z * z
    if j % 2 == 0:
        h = j // 2
        Is = _I_all(m + 2 * h, s, z)
        return sum(comb(h, i) * (-zz) ** (h - i) * Is[m + 2 * i] for i in range(h + 1))
    h = (j - 1) // 2
    r2 = s * s + zz
    tot = 0.0
    for i in range(h + 1):
        mm = m + 2 * i
        rp = r2 ** ((mm + 2) // 2) if mm % 2 == 0 else torch.sqrt(r2) * r2 ** ((mm + 1) // 2)
        tot = tot + comb(h, i) * (-zz) ** (h - i) * rp / (mm + 2)
    return tot


def _dangle(z, lo, hi):
    zero = z == 0
    safe = torch.where(zero, torch.ones_like(z), z)
    return torch.where(zero, torch.zeros_like(z), torch.atan2(safe * (hi - lo), z * z + lo * hi))


def _ypoly(g, beta):
    a, b = beta
    c1, c2 = [g.z * g.n0, g.t0], [g.z * g.n1, g.t1]
    out = [torch.ones_like(g.z)]

    def mul(u, v):
        r = [0.0] * (len(u) + len(v) - 1)
        for i, x in enumerate(u):
            for j, y in enumerate(v):
                r[i + j] = r[i + j] + x * y
        return r
    for _ in range(a):
        out = mul(out, c1)
    for _ in range(b):
        out = mul(out, c2)
    return out


class _Ctx:
    def __init__(self, g):
        self.g = g
        self.cache = {}

    def sdiff(self, j, m, R):
        key = ("S", j, m, float(R))
        if key not in self.cache:
            lo, hi, msk = self.g.chord(R)
            z = self.g.z
            self.cache[key] = torch.where(msk, _S(j, m, hi, z) - _S(j, m, lo, z), torch.zeros_like(z))
        return self.cache[key]

    def value(self, n, R):
        key = ("val", n, float(R))
        if key not in self.cache:
            lo, hi, msk = self.g.chord(R)
            z = self.g.z
            v = (z * self.sdiff(0, n, R) - float(R) ** (n + 2) * _dangle(z, lo, hi)) / (n + 2)
            self.cache[key] = torch.where(msk, v, torch.zeros_like(z))
        return self.cache[key]

    def grad(self, n, R):
        d = self.sdiff(0, n, R)
        return -self.g.n0 * d, -self.g.n1 * d

    def moment(self, alpha, n, R):
        key = ("mom", alpha, n, float(R))
        if key in self.cache:
            return self.cache[key]
        a, b = alpha
        if a + b == 0:
            return self.value(n, R)
        Rf = float(R)
        i = 0 if a > 0 else 1
        beta = (a - 1, b) if i == 0 else (a, b - 1)
        n2 = n + 2
        ni = self.g.n0 if i == 0 else self.g.n1
        acc = 0.0
        for j, cj in enumerate(_ypoly(self.g, beta)):
            acc = acc + cj * (self.sdiff(j, n2, R) - Rf ** n2 * self.sdiff(j, 0, R))
        val = ni * acc / n2
        if beta[i] >= 1:
            bm = (beta[0] - (i == 0), beta[1] - (i == 1))
            val = val - beta[i] * (self.moment(bm, n2, R) - Rf ** n2 * self.moment(bm, 0, R)) / n2
        self.cache[key] = val
        return val

    def moment_grad(self, alpha, n, R):
        acc = 0.0
        for j, cj in enumerate(_ypoly(self.g, alpha)):
            acc = acc + cj * self.sdiff(j, n, R)
        return -self.g.n0 * acc, -self.g.n1 * acc


def _sum_blocks(kname, fn):
    tot = None
    for R, coeffs in kernel_blocks(kname):
        for n, c in enumerate(coeffs):
            if c == 0:
                continue
            w = float(c) / math.pi
            v = fn(n, R)
            v = tuple(w * x for x in v) if isinstance(v, tuple) else w * v
            tot = v if tot is None else (tuple(a + b for a, b in zip(tot, v)) if isinstance(v, tuple) else tot + v)
    return tot


def edge_channels12(p, q, kname):
    """[P, 12] edge-local channels (device order, units of h) of edges p -> q relative to the query (torch CPU)."""
    c = _Ctx(_Geo(p, q))
    lam = _sum_blocks(kname, c.value)
    m10 = _sum_blocks(kname, lambda n, R: c.moment((1, 0), n, R))
    m01 = _sum_blocks(kname, lambda n, R: c.moment((0, 1), n, R))
    G = _sum_blocks(kname, c.grad)
    gx = _sum_blocks(kname, lambda n, R: c.moment_grad((1, 0), n, R))
    gy = _sum_blocks(kname, lambda n, R: c.moment_grad((0, 1), n, R))
    m20 = _sum_blocks(kname, lambda n, R: c.moment((2, 0), n, R))
    m11 = _sum_blocks(kname, lambda n, R: c.moment((1, 1), n, R))
    m02 = _sum_blocks(kname, lambda n, R: c.moment((0, 2), n, R))
    out = torch.stack([lam, m10, m01, G[0], G[1], gx[0] + lam, gx[1], gy[0], gy[1] + lam, m20, m11, m02], 1)
    return torch.where(c.g.ok[:, None], out, torch.zeros_like(out))


class Scene2D:
    """the static structures of one surface representation (built once, as the reference's cell list) as torch CPU
    tensors."""

    def __init__(self, verts, edges, background, support_max):
        st = oadj.Static(verts, edges, background, support_max)
        self.st = st
        self.V = torch.tensor(st.verts)
        self.E = torch.tensor(st.edges)
        self.cstart = torch.tensor(st.cl.start)
        self.citems = torch.tensor(st.cl.items)
        y0, dy, nb, bstart, bitems = st.ybins
        self.ybins = (float(y0), float(dy), int(nb), torch.tensor(bstart), torch.tensor(bitems))
        self.cind = torch.tensor(st.cind)
        self.cempty = torch.tensor(st.cempty)

    def _pairs(self, starts, stops):
        """(row, offset) of every item of the ragged runs [starts, stops)."""
        cnt = (stops - starts).clamp(min=0)
        rows = torch.repeat_interleave(torch.arange(len(cnt)), cnt)
        first = torch.cumsum(cnt, 0) - cnt
        off = torch.arange(int(cnt.sum())) - first[rows] + starts[rows]
        return rows, off

    def slots(self, pos, sup):
        """(rows, edges) of the kept (row, edge) pairs in slot order, and the per-row cell id / inside flag."""
        cl = self.st.cl
        nx, ny = cl.dims
        cx = torch.floor((pos[:, 0] - float(cl.lo[0])) / cl.cell).long()
        cy = torch.floor((pos[:, 1] - float(cl.lo[1])) / cl.cell).long()
        inside = (cx >= 0) & (cx < nx) & (cy >= 0) & (cy < ny)
        cid = torch.where(inside, cx * ny + cy, torch.zeros_like(cx))
        starts = torch.where(inside, self.cstart[cid], torch.zeros_like(cid))
        stops = torch.where(inside, self.cstart[cid + 1], torch.zeros_like(cid))
        rows, off = self._pairs(starts, stops)
        e = self.citems[off]
        a = self.V[self.E[e, 0]]
        d = self.V[self.E[e, 1]] - a
        x = pos[rows]
        t = torch.clamp(((x[:, 0] - a[:, 0]) * d[:, 0] + (x[:, 1] - a[:, 1]) * d[:, 1])
                        / torch.clamp(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1], min=1e-300), 0.0, 1.0)
        ex = x[:, 0] - (a[:, 0] + t * d[:, 0])
        ey = x[:, 1] - (a[:, 1] + t * d[:, 1])
        keep = torch.sqrt(ex * ex + ey * ey) < sup[rows]
        return rows[keep], e[keep], cid, inside

    def indicator(self, pos, sup, cid, inside):
        st = self.st
        y0, dy, nb, bstart, bitems = self.ybins
        ind = torch.full((len(pos),), float(st.background), dtype=F64)
        empty = inside & self.cempty[cid]
        ind = torch.where(empty, self.cind[cid], ind)
        ray = inside & ~empty
        iy = torch.floor((pos[:, 1] - y0) / dy).long()
        ray &= (iy >= 0) & (iy <= nb)
        b = torch.clamp(iy, 0, nb - 1)
        rows, off = self._pairs(torch.where(ray, bstart[b], 0), torch.where(ray, bstart[b + 1], 0))
        e = bitems[off]
        h = sup[rows, None]
        a = (self.V[self.E[e, 0]] - pos[rows]) / h
        c = (self.V[self.E[e, 1]] - pos[rows]) / h
        cross = a[:, 0] * c[:, 1] - a[:, 1] * c[:, 0]
        ay, cy = self.V[self.E[e, 0], 1], self.V[self.E[e, 1], 1]
        py = pos[rows, 1]
        w = ((ay <= py) & (py < cy) & (cross > 0)).to(F64) - ((cy <= py) & (py < ay) & (cross < 0)).to(F64)
        wind = torch.zeros(len(pos), dtype=F64).index_add_(0, rows, w)
        on = torch.zeros(len(pos), dtype=torch.bool)
        on[rows[(cross == 0) & (a[:, 0] * c[:, 0] + a[:, 1] * c[:, 1] <= 0)]] = True
        ind = torch.where(ray, wind + float(st.background), ind)
        for r in torch.nonzero(on & ray).flatten().tolist():            # on an edge: the angle sum over every edge
            ind[r] = _angle_sum(self, pos[r], sup[r]) + st.background
        return ind


def _angle_sum(sc, x, h):
    a = (sc.V[sc.E[:, 0]] - x) / h
    c = (sc.V[sc.E[:, 1]] - x) / h
    cross = a[:, 0] * c[:, 1] - a[:, 1] * c[:, 0]
    ang = torch.where(cross != 0, torch.atan2(cross, a[:, 0] * c[:, 0] + a[:, 1] * c[:, 1]), torch.zeros_like(cross))
    return float(ang.sum()) / oadj.TWO_PI_LITERAL


def twelve(scene: Scene2D, pos, sup, kernels):
    """{kernel: [N, 12]} (units of h): slot fill, indicator, edge channels of every kept pair, row sums."""
    rows, e, cid, inside = scene.slots(pos, sup)
    ind = scene.indicator(pos, sup, cid, inside)
    h = sup[rows, None]
    P = (scene.V[scene.E[e, 0]] - pos[rows]) / h
    Q = (scene.V[scene.E[e, 1]] - pos[rows]) / h
    out = {}
    for k in kernels:
        acc = torch.zeros((len(pos), 12), dtype=F64).index_add_(0, rows, edge_channels12(P, Q, k))
        out[k] = acc + ind[:, None] * torch.tensor(np.asarray(oedge.plane_vector(k), dtype=np.float64))[None]
    return out
