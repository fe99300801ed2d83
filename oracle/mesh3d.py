"""Numpy restatement of the 3-D face->edge reductions in FP64 or extended precision (``oracle.arith``; the "mp"
mode is not supported here: the formulas evaluate guarded divisions on both branches).  TEST INFRASTRUCTURE (see
oracle/__init__).

This follows ``mesh3d/reference.py`` of the reference operation by operation, vectorised over (query, face) pairs:

* kernel profile (power, cutoff, coefficient) with the 3-D normalisation : ``mesh3d/reference.py:54-61``
* the K_m primitives  int r^m/(s^2+w^2) ds                             : ``mesh3d/reference.py:74-88``
* the triangle frame (normal, z, e1, e2) and edge records
  (w, lo, hi, mu.e1, mu.e2, t.e1, t.e2)                                 : ``mesh3d/reference.py:91-125``
* face scalar  int_f r^n 1[r <= R] dA  (cuts at +-L, midpoint test)    : ``mesh3d/reference.py:127-150``
* edge line integrals of in-plane monomials                            : ``mesh3d/reference.py:152-171``
* in-plane monomial recursion                                          : ``mesh3d/reference.py:173-187``
* face block moments of y^alpha, kernel gradient                        : ``mesh3d/reference.py:189-216``
* closed surface (regular radial flux, recursion in alpha), moment
  gradient, kernel gradient, Laplacian, Hessian                        : ``mesh3d/reference.py:219-325``
* tetrahedron orientation, tetrahedral meshes                           : ``mesh3d/reference.py:328-412``
* edge primitives I_m, J_m, S_{j,m}                                     : ``edge/primitives.py:19-84``
* element adjacency predicates (host and device forms)                  : ``fluxbridge/geometry.py:97-114``,
                                                                         ``fluxbridge/device_search.py:80-85``
* compact-mode face skip and face solid angle                           : ``fluxbridge/wp_func.py:290-318``

The reference evaluates these formulas in mpmath; here they run in FP64 (or x86 extended precision) with the same
operation structure.  The
oracle is pinned by the reference's 70-digit fixtures (``tests/fixtures/fluxbridge-v1.json``,
``fluxbridge-nearplane-v1.json``), exact whole-support moments, the cavity identity and the planar limit
(``tests/test_oracle3d.py``).
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from math import comb

import numpy as np

from . import arith as A
from .kernels import blocks as kernel_blocks, c3_over_c2

ZERO = (0, 0, 0)


# ------------------------------------------------------------------ kernel profile
@lru_cache(maxsize=None)
def profile_exact(name):
    """((n, R, c)) with pi * W_3d = sum c r^n 1[r <= R] at unit support (mesh3d/reference.py:54-61, without 1/pi)."""
    ratio = c3_over_c2(name)
    return tuple((n, R, c * ratio) for R, cs in kernel_blocks(name) for n, c in enumerate(cs) if c)


def profile(name, h=1.0):
    """physical profile (n, R h, c / (pi h^(n+3))) in the working precision."""
    h = A.asarray(float(h)) if A.mode() != "f64" else float(h)
    return tuple((n, float(R) * h, A.num(c) / A.pi() / h ** (n + 3)) for n, R, c in profile_exact(name))


def _gradient_profile(prof):
    return tuple((n - 2, R, c * n) for n, R, c in prof if n)


# ------------------------------------------------------------------ primitives (edge/primitives.py)
def _I_all(M, s, z):
    zz = z * z
    r2 = s * s + zz
    r = A.sqrt(r2)
    out = [s]
    if M >= 1:
        safe = np.where(z == 0, 1.0, np.abs(z))
        tail = np.where(z == 0, 0.0, zz * A.arcsinh(s / safe))
        out.append((s * r + tail) / 2)
    for m in range(2, M + 1):
        rpow = r2 ** (m // 2) if m % 2 == 0 else r * r2 ** ((m - 1) // 2)
        out.append((s * rpow + m * zz * out[m - 2]) / (m + 1))
    return out


def Ig(m, s, z):
    """int r^m ds, r^2 = s^2 + z^2; negative m needs z != 0 (edge/primitives.py:58-72)."""
    if m >= 0:
        return _I_all(m, s, z)[m]
    zz = z * z
    safe = np.where(z == 0, 1.0, z)
    if m == -1:
        return A.arcsinh(s / np.abs(safe))
    if m == -2:
        return A.arctan(s / safe) / safe
    r2 = s * s + zz
    return ((m + 3) * Ig(m + 2, s, z) - s * r2 ** ((m + 2) / 2)) / ((m + 2) * zz)


def Jg(m, s, z):
    """int s r^m ds (edge/primitives.py:75-78)."""
    r2 = s * s + z * z
    if m == -2:
        return A.log(r2) / 2
    if m % 2 == 0:
        return r2 ** ((m + 2) // 2) / (m + 2)
    return A.sqrt(r2) * r2 ** ((m + 1) // 2) / (m + 2)


def Sg(j, m, s, z):
    """int s^j r^m ds (edge/primitives.py:81-88)."""
    zz = z * z
    if j % 2 == 0:
        h = j // 2
        return sum(comb(h, i) * (-zz) ** (h - i) * Ig(m + 2 * i, s, z) for i in range(h + 1))
    h = (j - 1) // 2
    return sum(comb(h, i) * (-zz) ** (h - i) * Jg(m + 2 * i, s, z) for i in range(h + 1))


def K(m, s, w, a):
    """int r^m/(s^2+w^2) ds with r^2 = s^2 + w^2 + a^2, w != 0 (mesh3d/reference.py:74-88)."""
    b = np.abs(w)
    c = A.sqrt(w * w + a * a)
    if m == 0:
        return A.arctan(s / b) / b
    if m == -1:
        r = A.sqrt(s * s + c * c)
        sa = np.where(a == 0, 1.0, a)
        return np.where(a != 0, A.arctan(sa * s / (b * r)) / (sa * b), s / (b * b * r))
    if m == 1:
        return A.arcsinh(s / c) + a * a * K(-1, s, w, a)
    if m >= 2:
        return Ig(m - 2, s, c) + a * a * K(m - 2, s, w, a)
    raise ValueError("K supports m >= -1")


def _dot(u, v):
    return u[..., 0] * v[..., 0] + u[..., 1] * v[..., 1] + u[..., 2] * v[..., 2]


def _cross(u, v):
    return np.stack([u[..., 1] * v[..., 2] - u[..., 2] * v[..., 1],
                     u[..., 2] * v[..., 0] - u[..., 0] * v[..., 2],
                     u[..., 0] * v[..., 1] - u[..., 1] * v[..., 0]], axis=-1)


def _poly_mul(p, q):
    out = {}
    for a, c in p.items():
        for b, d in q.items():
            k = tuple(x + y for x, y in zip(a, b))
            out[k] = out.get(k, 0) + c * d
    return out


def _lower(alpha, i, amount=1):
    out = list(alpha)
    out[i] -= amount
    return tuple(out)


# ------------------------------------------------------------------ faces (mesh3d/reference.py:91-216)
class Faces:
    """P triangles, each relative to its own query: vertices V[P,3,3], queries x[P,3] (any consistent units)."""

    def __init__(self, V, x):
        V = A.asarray(V)
        x = A.asarray(x)
        p = V - x[:, None, :]
        d = p[:, 1] - p[:, 0]
        N = _cross(d, p[:, 2] - p[:, 0])
        nlen = A.sqrt(_dot(N, N))
        if np.any(nlen == 0):
            raise ValueError("degenerate triangle")
        ell = A.sqrt(_dot(d, d))
        self.P = len(V)
        self.normal = N / nlen[:, None]
        self.z = _dot(N, p[:, 0]) / nlen
        self.a = np.abs(self.z)
        self.e1 = d / ell[:, None]
        self.e2 = _cross(self.normal, self.e1)
        self.edges = []
        for j in range(3):
            v, q = p[:, j], p[:, (j + 1) % 3]
            D = q - v
            length = A.sqrt(_dot(D, D))
            t = D / length[:, None]
            mu = _cross(t, self.normal)
            w = _dot(_cross(D, N), v) / (length * nlen)
            lo, hi = _dot(D, v) / length, _dot(D, q) / length
            self.edges.append((w, lo, hi, _dot(mu, self.e1), _dot(mu, self.e2), _dot(t, self.e1), _dot(t, self.e2)))
        self._cache = {}

    def _memo(self, key, fn):
        v = self._cache.get(key)
        if v is None:
            v = fn()
            self._cache[key] = v
        return v

    # -- int_f r^n 1[r <= R] dA, R=None: whole face
    def scalar(self, n, R):
        return self._memo(("S", n, R), lambda: self._scalar(n, R))

    def _scalar(self, n, R):
        a = self.a
        if n == -2 or n < -3:
            raise ValueError("unsupported singular face profile")
        live = np.ones(self.P, bool) if R is None else (a < R)
        if n == -3:
            live = live & (a > 0)
        total = np.zeros(self.P)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            for w, lo, hi, *_ in self.edges:
                ok = live & (w != 0)
                if not ok.any():
                    continue
                ws = np.where(ok, w, 1.0)
                if R is None:
                    pts = [lo, hi]
                else:
                    c2 = a * a + ws * ws
                    cut = R * R > c2
                    L = A.sqrt(np.where(cut, R * R - c2, 0.0))
                    x1 = np.where(cut & (lo < -L) & (-L < hi), -L, lo)
                    x2 = np.where(cut & (lo < L) & (L < hi), L, hi)
                    pts = [lo, x1, x2, hi]
                for l, u in zip(pts, pts[1:]):
                    inside = np.ones(self.P, bool) if R is None else ((l + u) / 2) ** 2 + ws * ws + a * a < R * R
                    dK0 = K(0, u, ws, a) - K(0, l, ws, a)
                    if n + 2 == -1:
                        an2 = np.where(a > 0, 1.0 / np.where(a > 0, a, 1.0), 0.0)
                    else:
                        an2 = a ** (n + 2)
                    term_in = K(n + 2, u, ws, a) - K(n + 2, l, ws, a) - an2 * dK0
                    if R is None:
                        term = term_in
                    else:
                        term = np.where(inside, term_in, (R ** (n + 2) - an2) * dK0)
                    total = total + np.where(ok, ws * term / (n + 2), 0.0)
        return total

    # -- int_edge u^b1 v^b2 r^n 1[r <= R] ds
    def line(self, e, beta, n, R):
        return self._memo(("L", e, beta, n, R), lambda: self._line(e, beta, n, R))

    def _line(self, e, beta, n, R):
        w, lo, hi, mu1, mu2, t1, t2 = self.edges[e]
        c2 = w * w + self.a * self.a
        ok = c2 < R * R
        L = A.sqrt(np.where(ok, R * R - c2, 0.0))
        l, u = np.maximum(lo, -L), np.minimum(hi, L)
        ok = ok & (l < u)
        poly = [np.ones(self.P)]
        for b, mu, t in zip(beta, (mu1, mu2), (t1, t2)):
            q = [comb(b, j) * (w * mu) ** (b - j) * t ** j for j in range(b + 1)]
            new = [0.0] * (len(poly) + b)
            for i, v in enumerate(poly):
                for j, v2 in enumerate(q):
                    new[i + j] = new[i + j] + v * v2
            poly = new
        c = A.sqrt(c2)
        tot = np.zeros(self.P)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            for j, v in enumerate(poly):
                tot = tot + v * (Sg(j, n, u, c) - Sg(j, n, l, c))
        return np.where(ok, tot, 0.0)

    # -- in-plane monomial u^b1 v^b2 r^n 1[r <= R] over the face
    def local(self, beta, n, R):
        return self._memo(("l", beta, n, R), lambda: self._local(beta, n, R))

    def _local(self, beta, n, R):
        if not any(beta):
            return self.scalar(n, R)
        i = 0 if beta[0] else 1
        b = list(beta)
        b[i] -= 1
        b = tuple(b)
        edge = sum(self.edges[e][3 + i] * (self.line(e, b, n + 2, R) - R ** (n + 2) * self.line(e, b, 0, R))
                   for e in range(3)) / (n + 2)
        if b[i]:
            lower = list(b)
            lower[i] -= 1
            lower = tuple(lower)
            edge = edge - b[i] * (self.local(lower, n + 2, R) - R ** (n + 2) * self.local(lower, 0, R)) / (n + 2)
        return edge

    def block_moment(self, alpha, n, R):
        """int_f y^alpha r^n 1[r <= R] dA, y = source - query."""
        return self._memo(("B", tuple(alpha), n, R), lambda: self._block_moment(tuple(alpha), n, R))

    def _block_moment(self, alpha, n, R):
        poly = {(0, 0): np.ones(self.P)}
        for i, power in enumerate(alpha):
            linear = {(0, 0): self.z * self.normal[:, i], (1, 0): self.e1[:, i], (0, 1): self.e2[:, i]}
            for _ in range(power):
                poly = _poly_mul(poly, linear)
        tot = np.zeros(self.P)
        for b, c in poly.items():
            tot = tot + c * self.local(b, n, R)
        return tot

    def integrate(self, prof, alpha=ZERO):
        return sum((c * self.block_moment(alpha, n, R) for n, R, c in prof), np.zeros(self.P))

    # -- triangle (surface measure) API
    def moment(self, name="w2", alpha=ZERO, h=1.0):
        return self.integrate(profile(name, h), alpha)

    def kernel_gradient(self, name="w2", alpha=ZERO, h=1.0):
        L = _gradient_profile(profile(name, h))
        return tuple(-self.integrate(L, tuple(a + int(i == j) for i, a in enumerate(alpha))) for j in range(3))


# ------------------------------------------------------------------ closed surfaces (mesh3d/reference.py:219-325)
class ClosedSurfacePairs:
    """Q queries x all F faces of one closed oriented surface (faces with outward normals)."""

    def __init__(self, vertices, faces, queries, winding=None):
        """winding=None: the reference's regular radial flux with every face's complete solid angle
        (mesh3d/reference.py:255-258).  winding=[Q]: the compact form of fluxbridge/wp_func.py:290-306 with the total
        solid angle 4 pi winding supplied (the solid angle sum of the faces, or its integer value off the surface)."""
        V = A.asarray(vertices)
        Fi = np.asarray(faces, dtype=np.int64)
        X = A.asarray(queries).reshape(-1, 3)
        self.Q, self.F = len(X), len(Fi)
        self.winding = None if winding is None else A.asarray(winding).reshape(-1)
        tri = V[Fi]                                            # [F,3,3]
        self.faces = Faces(np.broadcast_to(tri, (self.Q,) + tri.shape).reshape(-1, 3, 3),
                           np.repeat(X, self.F, axis=0))
        self._cache = {}

    def _sum(self, v):
        return v.reshape(self.Q, self.F).sum(axis=1)

    def block_moment(self, alpha, n, R):
        key = (tuple(alpha), n, R)
        v = self._cache.get(key)
        if v is None:
            v = self._block_moment(tuple(alpha), n, R)
            self._cache[key] = v
        return v

    def _block_moment(self, alpha, n, R):
        f = self.faces
        if not any(alpha):
            if self.winding is not None:
                term = f.z * (f.scalar(n, R) - R ** (n + 3) * f.scalar(-3, R)) / (n + 3)
                return self._sum(np.where(f.z != 0, term, 0.0)) + 4 * A.pi() * self.winding * R ** (n + 3) / (n + 3)
            # regular radial flux: z (S(n,R) + R^(n+3) (S(-3,inf) - S(-3,R))) / (n+3), faces with z != 0
            term = f.z * (f.scalar(n, R) + R ** (n + 3) * (f.scalar(-3, None) - f.scalar(-3, R))) / (n + 3)
            return self._sum(np.where(f.z != 0, term, 0.0))
        i = next(j for j, a in enumerate(alpha) if a)
        beta = _lower(alpha, i)
        res = self._sum(f.normal[:, i] * (f.block_moment(beta, n + 2, R) - R ** (n + 2) * f.block_moment(beta, 0, R))) / (n + 2)
        if beta[i]:
            lower = _lower(beta, i)
            res = res - beta[i] * (self.block_moment(lower, n + 2, R) - R ** (n + 2) * self.block_moment(lower, 0, R)) / (n + 2)
        return res

    def integrate(self, prof, alpha=ZERO):
        return sum((c * self.block_moment(alpha, n, R) for n, R, c in prof), np.zeros(self.Q))

    def moment(self, name="w2", alpha=ZERO, h=1.0):
        return self.integrate(profile(name, h), alpha)

    def moment_gradient(self, name="w2", alpha=ZERO, h=1.0):
        prof = profile(name, h)
        face = self.faces.integrate(prof, tuple(alpha))
        return tuple(-self._sum(self.faces.normal[:, j] * face) for j in range(3))

    def kernel_gradient(self, name="w2", alpha=ZERO, h=1.0):
        g = self.moment_gradient(name, alpha, h)
        return tuple(g[j] + (alpha[j] * self.moment(name, _lower(alpha, j), h) if alpha[j] else 0.0) for j in range(3))

    def laplacian(self, name="w2", alpha=ZERO, h=1.0):
        W = profile(name, h)
        L = _gradient_profile(W)
        f = self.faces
        value = self._sum(f.z * f.integrate(L, tuple(alpha)))
        for j, a in enumerate(alpha):
            if a:
                lower = _lower(alpha, j)
                value = value - a * self._sum(f.normal[:, j] * f.integrate(W, lower))
            if a >= 2:
                value = value + a * (a - 1) * self.integrate(W, _lower(alpha, j, 2))
        return value

    def kernel_hessian(self, name="w2", alpha=ZERO, h=1.0):
        L = _gradient_profile(profile(name, h))
        f = self.faces
        out = []
        for i in range(3):
            lower_g = self.kernel_gradient(name, _lower(alpha, i), h) if alpha[i] else (0.0, 0.0, 0.0)
            row = []
            for j in range(3):
                up = tuple(a + int(k == j) for k, a in enumerate(alpha))
                row.append(self._sum(f.normal[:, i] * f.integrate(L, up)) + alpha[i] * lower_g[j])
            out.append(tuple(row))
        return tuple(out)


def tetrahedron_faces(vertices):
    """outward faces of one tetrahedron (mesh3d/reference.py:332-347)."""
    V = [np.asarray(v, dtype=np.float64) for v in vertices]
    faces = []
    for i in range(4):
        f = [j for j in range(4) if j != i]
        a, b, c = (V[j] for j in f)
        side = np.dot(np.cross(b - a, c - a), V[i] - a)
        if side == 0:
            raise ValueError("degenerate tetrahedron")
        if side > 0:
            f[1], f[2] = f[2], f[1]
        faces.append(tuple(f))
    return faces


def check_closed(faces):
    """watertight and consistently oriented (mesh3d/reference.py:238-245)."""
    edges = {}
    for f in faces:
        for a, b in zip(f, tuple(f[1:]) + tuple(f[:1])):
            key = (min(a, b), max(a, b))
            edges.setdefault(key, []).append(1 if a < b else -1)
    return all(len(e) == 2 and sum(e) == 0 for e in edges.values())


# ------------------------------------------------------------------ the 22 stored channels (fluxbridge/contract.py:41-46)
def monomials(d, degree):
    """graded order of fluxbridge/contract.py:29-38."""
    def lattice(dd, n):
        if dd == 1:
            return [(n,)]
        return [(i,) + b for i in range(n, -1, -1) for b in lattice(dd - 1, n - i)]
    return tuple(a for n in range(degree + 1) for a in lattice(d, n))


def baseline3():
    """[(alpha, derivative)] of baseline(3): 10 moments, 12 first-gradient moments, 6 Hessian entries."""
    z = (0, 0, 0)
    return ([(a, ()) for a in monomials(3, 2)] + [(a, (j,)) for a in monomials(3, 1) for j in range(3)]
            + [(z, (i, j)) for i in range(3) for j in range(i, 3)])


def closed_surface_requests(vertices, faces, queries, kernel, h, requests, chunk=64, winding=None):
    """values [Q, len(requests)] of volume requests (alpha, derivative) over the solid a closed surface bounds
    (the evaluate() dispatch of fluxbridge/reference.py:9-30 for a ClosedSurface); winding[Q] selects the compact
    form (see ClosedSurfacePairs)."""
    X = np.asarray(queries, dtype=np.float64).reshape(-1, 3)
    out = A.zeros((len(X), len(requests)))
    W = None if winding is None else np.asarray(winding, dtype=np.float64).reshape(-1)
    for s in range(0, len(X), chunk):
        cs = ClosedSurfacePairs(vertices, faces, X[s:s + chunk], None if W is None else W[s:s + chunk])
        for k, (a, ds) in enumerate(requests):
            if not ds:
                out[s:s + chunk, k] = cs.moment(kernel, a, h)
            elif len(ds) == 1:
                out[s:s + chunk, k] = cs.kernel_gradient(kernel, a, h)[ds[0]]
            else:
                out[s:s + chunk, k] = cs.kernel_hessian(kernel, a, h)[ds[0]][ds[1]]
    return out


def triangle_requests(vertices, queries, kernel, h, requests):
    """values [Q, len(requests)] of surface-measure requests of one open triangle (first derivatives only)."""
    X = np.asarray(queries, dtype=np.float64).reshape(-1, 3)
    V = np.broadcast_to(np.asarray(vertices, dtype=np.float64), (len(X), 3, 3))
    f = Faces(V, X)
    out = A.zeros((len(X), len(requests)))
    for k, (a, ds) in enumerate(requests):
        if not ds:
            out[:, k] = f.moment(kernel, a, h)
        elif len(ds) == 1:
            out[:, k] = f.kernel_gradient(kernel, a, h)[ds[0]]
        else:
            raise NotImplementedError("surface Hessians are differentiated numerically in the reference")
    return out


# ------------------------------------------------------------------ solid angle, predicates
def face_solid_angle(V, x):
    """2 atan2(det, den) of the normalised face corners seen from x (fluxbridge/wp_func.py:296-302), V[P,3,3]."""
    a, b, c = (A.asarray(V[:, k]) - A.asarray(x) for k in range(3))
    det = _dot(a, _cross(b - a, c - a))
    la, lb, lc = (A.sqrt(_dot(u, u)) for u in (a, b, c))
    den = la * lb * lc + _dot(a, b) * lc + _dot(b, c) * la + _dot(c, a) * lb
    return np.where(det != 0, 2.0 * A.arctan2(det, den), 0.0)


def winding(vertices, faces, queries, chunk=256):
    """sum of face solid angles / 4 pi over every face."""
    V = np.asarray(vertices, dtype=np.float64)[np.asarray(faces, dtype=np.int64)]
    X = np.asarray(queries, dtype=np.float64).reshape(-1, 3)
    out = A.zeros(len(X))
    for s in range(0, len(X), chunk):
        xs = X[s:s + chunk]
        VV = np.broadcast_to(V, (len(xs),) + V.shape).reshape(-1, 3, 3)
        xx = np.repeat(xs, len(V), axis=0)
        out[s:s + chunk] = face_solid_angle(VV, xx).reshape(len(xs), len(V)).sum(axis=1) / (4 * A.pi())
    return out


def snapped_winding(vertices, faces, queries, tol=1e-9):
    """the solid-angle winding number, rounded to the nearest integer where it is within tol of one (points off the
    surface, where the winding number is an integer); unchanged on the surface."""
    w = winding(vertices, faces, queries)
    r = np.rint(w)
    return np.where(np.abs(w - r) < tol, r, w)


def element_bounds(V):
    """[E,2,3] per-element vertex min / max (fluxbridge/geometry.py:62)."""
    V = np.asarray(V, dtype=np.float64)
    return np.stack([V.min(axis=1), V.max(axis=1)], axis=1)


def device_keep(bounds, x, h):
    """device predicate: sum_a (max(0, lo-x, x-hi)/h)^2 <= 1 + 1.42e-14, accumulated in axis order
    (fluxbridge/device_search.py:80-85).  bounds[E,2,3], x[3], h scalar -> bool[E]."""
    sq = np.zeros(len(bounds))
    for a in range(3):
        delta = np.maximum(0.0, np.maximum(bounds[:, 0, a] - x[a], x[a] - bounds[:, 1, a])) / h
        sq = sq + delta * delta
    return sq <= 1.0 + 1.4210854715202004e-14


def host_keep(bounds, x, h):
    """host predicate: |max(lo-x, x-hi, 0)|^2 <= h^2 (fluxbridge/geometry.py:106-110)."""
    delta = np.maximum(np.maximum(bounds[:, 0] - x, x - bounds[:, 1]), 0)
    return (delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1] + delta[:, 2] * delta[:, 2]) <= h * h


def compact_keep(V, x, h):
    """compact-mode face test: keep unless |max(0, low, -high)|^2 >= 1 in units of h (fluxbridge/wp_func.py:307-313)."""
    p = (np.asarray(V, dtype=np.float64) - x) / h
    low = np.minimum(p[:, 0], np.minimum(p[:, 1], p[:, 2]))
    high = np.maximum(p[:, 0], np.maximum(p[:, 1], p[:, 2]))
    delta = np.maximum(0.0, np.maximum(low, -high))
    return (delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1] + delta[:, 2] * delta[:, 2]) < 1.0


def face_adjacency(V, X, H):
    """(query, face) CSR of the device predicate, faces ascending per query: brute force."""
    bounds = element_bounds(V)
    cols, offs = [], [0]
    for x, h in zip(np.asarray(X, dtype=np.float64), np.broadcast_to(np.asarray(H, dtype=np.float64), (len(X),))):
        ids = np.nonzero(device_keep(bounds, x, h))[0]
        cols.append(ids)
        offs.append(offs[-1] + len(ids))
    return np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32), np.array(offs, dtype=np.int64)
