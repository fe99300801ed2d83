"""Synthetic scenes of the benchmark configurations (seeded, numpy ``Generator(PCG64(seed))``).

* ``config1``: unit-square box domain (4 edges, solid outside) plus a closed regular 64-gon obstacle of radius 0.15
  at (0.5, 0.5) (``areaPreserving=False``); 16,384 particles uniform in the box outside the obstacle; w4, h = 0.02.
* ``config2``: one star polygon of 4,096 edges (vertex radii 0.3 + U(-0.05, 0.05) at equal angles, centre
  (0.5, 0.5), solid inside); 1,048,576 particles in [0,1]^2, 30 % within one support of the wall (an edge picked in
  proportion to its length, a point uniform on it, offset d = h U(0,1) along the fluid normal) and the rest uniform
  outside the star; h = 0.004; kernels cubic, w2, w4, w6.
* ``config3``: the fluid in the unit box [0,1]^3 around an icosphere of radius 0.25 at (0.5, 0.5, 0.5) subdivided to
  81,920 triangles: one closed surface of the 12 outward box triangles and the sphere's triangles turned into the
  cavity; 2,097,152 particles uniform in the box outside the sphere; w4, h = 0.01.
* ``config4`` (precision stress): 2-D wedge / notch polygons with interior angle U(5, 355) degrees at the tip and
  particles at d/h = 10^U(-7, 0) from an edge interior (normal offset, both sides) or from a vertex (random
  direction); 3-D sliver tetrahedra whose base triangle has aspect ratio 10^U(0, 4), particles at the same distances
  from the base face interior (both sides) or from a base vertex; kernels cubic, w2, w4, w6.
* ``config5`` (served mix): 256 scenes, a quarter of them 3-D meshes; see ``config5`` and ``served_scene_3d``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def inside_polygon(pts: np.ndarray, poly: np.ndarray) -> np.ndarray:
    """crossing-number point-in-polygon test (vectorised over points, loop over edges)."""
    x, y = pts[:, 0], pts[:, 1]
    inside = np.zeros(len(pts), dtype=bool)
    n = len(poly)
    for i in range(n):
        ax, ay = poly[i]
        bx, by = poly[(i + 1) % n]
        cond = (ay > y) != (by > y)
        with np.errstate(divide="ignore", invalid="ignore"):
            xc = ax + (y - ay) * (bx - ax) / (by - ay)
        inside ^= cond & (x < xc)
    return inside


def inside_star(pts: np.ndarray, star: np.ndarray, center=(0.5, 0.5)) -> np.ndarray:
    """point-in-polygon for a polygon star-shaped about ``center`` with vertices at equal angles 2 pi k / n (ccw):
    the point is inside iff it lies left of the edge whose angular wedge holds it.  O(N)."""
    n = len(star)
    rel = pts - np.asarray(center)
    th = np.mod(np.arctan2(rel[:, 1], rel[:, 0]), 2 * np.pi)
    k = np.minimum((th / (2 * np.pi / n)).astype(np.int64), n - 1)
    a = star[k]
    b = star[(k + 1) % n]
    cr = (b[:, 0] - a[:, 0]) * (pts[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (pts[:, 0] - a[:, 0])
    inside_wedge_tri = cr > 0
    return inside_wedge_tri


@dataclass
class Workload:
    name: str
    loops: list          # [(vertices [n,2] counter-clockwise around the solid, solid 'inside' | 'outside')]
    positions: np.ndarray
    supports: np.ndarray
    kernels: tuple
    h: float
    near: np.ndarray     # bool [N]: generated within one support of the wall


def regular_polygon(center, radius, n):
    t = 2 * np.pi * np.arange(n) / n
    return np.stack([center[0] + radius * np.cos(t), center[1] + radius * np.sin(t)], 1)


def config1(n=16384, seed=1):
    g = rng(seed)
    box = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    gon = regular_polygon((0.5, 0.5), 0.15, 64)
    pts = np.empty((0, 2))
    while len(pts) < n:
        cand = g.uniform(0.0, 1.0, (2 * n, 2))
        cand = cand[~inside_polygon(cand, gon)]
        pts = np.concatenate([pts, cand])
    pts = pts[:n]
    h = 0.02
    return Workload("box+64gon", [(box, "outside"), (gon, "inside")], pts, np.full(n, h), ("w4",), h, np.zeros(n, dtype=bool))


def star_polygon(n_edges=4096, seed=2, g=None):
    g = g or rng(seed)
    th = 2 * np.pi * np.arange(n_edges) / n_edges
    r = 0.3 + g.uniform(-0.05, 0.05, n_edges)
    return np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)], 1)


def config2(n=1048576, seed=2, n_edges=4096, h=0.004, near_frac=0.3):
    g = rng(seed)
    star = star_polygon(n_edges, g=g)
    n_near = int(round(near_frac * n))
    p = star
    q = np.roll(star, -1, axis=0)
    d = q - p
    ell = np.hypot(d[:, 0], d[:, 1])
    nf = np.stack([d[:, 1], -d[:, 0]], 1) / ell[:, None]           # outward (fluid) normal of the ccw solid
    near = np.empty((0, 2))
    while len(near) < n_near:
        m = 2 * (n_near - len(near)) + 1024
        e = g.choice(n_edges, size=m, p=ell / ell.sum())
        t = g.uniform(0.0, 1.0, m)
        off = h * g.uniform(0.0, 1.0, m)
        x = p[e] + t[:, None] * d[e] + off[:, None] * nf[e]
        ok = (x >= 0).all(1) & (x <= 1).all(1)
        x = x[ok]
        x = x[~inside_star(x, star)]
        near = np.concatenate([near, x])
    near = near[:n_near]
    far = np.empty((0, 2))
    n_far = n - n_near
    while len(far) < n_far:
        c = g.uniform(0.0, 1.0, (int(1.3 * (n_far - len(far))) + 1024, 2))
        c = c[~inside_star(c, star)]
        far = np.concatenate([far, c])
    far = far[:n_far]
    pts = np.concatenate([near, far])
    flag = np.zeros(n, dtype=bool)
    flag[:n_near] = True
    perm = g.permutation(n)
    return Workload("star4096", [(star, "inside")], pts[perm], np.full(n, h), ("cubic", "w2", "w4", "w6"), h, flag[perm])


def subsample(w: Workload, n: int, seed=7) -> Workload:
    """a stratified subsample that keeps the near-wall fraction."""
    g = rng(seed)
    near = np.nonzero(w.near)[0]
    far = np.nonzero(~w.near)[0]
    k = int(round(n * len(near) / max(len(w.near), 1)))
    idx = np.concatenate([g.choice(near, k, replace=False) if k else near[:0], g.choice(far, n - k, replace=False)])
    idx.sort()
    return Workload(w.name + f"-sub{n}", w.loops, w.positions[idx], w.supports[idx], w.kernels, w.h, w.near[idx])


@dataclass
class Workload3:
    name: str
    vertices: np.ndarray     # [nv, 3]
    faces: np.ndarray        # [nf, 3] outward (cavity faces point into the void)
    positions: np.ndarray    # [N, 3]
    supports: np.ndarray     # [N]
    kernels: tuple
    h: float


def box_with_sphere_cavity(level=6, radius=0.25, center=(0.5, 0.5, 0.5)):
    from .mesh3d import box_surface, icosphere
    Vb, Fb = box_surface((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    Vs, Fs = icosphere(level, radius, center)
    return np.vstack([Vb, Vs]), np.vstack([Fb, (Fs + len(Vb))[:, ::-1]]).astype(np.int32)


def config3(n=2097152, seed=3, level=6, h=0.01):
    g = rng(seed)
    V, F = box_with_sphere_cavity(level)
    pts = np.empty((0, 3))
    while len(pts) < n:
        c = g.uniform(0.0, 1.0, (int(1.1 * (n - len(pts))) + 1024, 3))
        c = c[np.sum((c - 0.5) ** 2, axis=1) > 0.25 ** 2]
        pts = np.concatenate([pts, c])
    pts = pts[:n]
    return Workload3(f"box+icosphere{len(F) - 12}", V, F, pts, np.full(n, h), ("w4",), h)


@dataclass
class StressPolygons:
    """precision stress, 2-D: ``polygons[i]`` counter-clockwise around the solid (the wedge); particle j belongs to
    polygon ``poly[j]``, sits d_over_h[j] * h from its anchor (``anchor[j]`` 0: an edge interior, 1: a vertex)."""
    name: str
    h: float
    kernels: tuple
    polygons: list
    angles_deg: np.ndarray
    poly: np.ndarray
    positions: np.ndarray
    d_over_h: np.ndarray
    anchor: np.ndarray


@dataclass
class StressTetrahedra:
    """precision stress, 3-D: tetrahedron i (vertices ``vertices[i]`` [4,3], outward faces ``faces`` [4,3]) has a
    sliver base triangle (vertices 0, 1, 2) of aspect ratio ``aspect[i]``; particle j belongs to tetrahedron
    ``tet[j]``, sits d_over_h[j] * h from the base face interior (anchor 0) or from a base vertex (anchor 1)."""
    name: str
    h: float
    kernels: tuple
    vertices: np.ndarray
    faces: np.ndarray
    aspect: np.ndarray
    tet: np.ndarray
    positions: np.ndarray
    d_over_h: np.ndarray
    anchor: np.ndarray


def wedge_polygon(theta, arm, center, rotation, arc_step=np.pi / 8):
    """tip at ``center`` with interior angle ``theta`` (radians, in (0, 2 pi)): the solid is the sector of radius
    ``arm`` around the direction ``rotation + pi``, closed by a polygonal arc; counter-clockwise vertices."""
    m = max(2, int(np.ceil(theta / arc_step)))
    phi = rotation + np.pi + np.linspace(-theta / 2, theta / 2, m + 1)
    arc = arm * np.stack([np.cos(phi), np.sin(phi)], 1)
    return np.vstack([[0.0, 0.0], arc]) + np.asarray(center)


def config4_2d(n=262144, n_poly=256, h=0.01, seed=4):
    """precision stress in 2-D (see the module docstring): ``n_poly`` wedges of arm 4 h, ``n // n_poly`` particles
    each, half anchored on edge interiors (an edge picked uniformly, t ~ U(0, 1), side +-1), half on vertices (the
    tip with probability 1/2, else a vertex picked uniformly; a uniform direction)."""
    g = rng(seed)
    per = n // n_poly
    polys, angles, P, D, A, I = [], [], [], [], [], []
    for i in range(n_poly):
        theta = np.deg2rad(g.uniform(5.0, 355.0))
        V = wedge_polygon(theta, 4 * h, g.uniform(0.2, 0.8, 2), g.uniform(0, 2 * np.pi))
        polys.append(V)
        angles.append(np.rad2deg(theta))
        ne = len(V)
        d = h * 10.0 ** g.uniform(-7.0, 0.0, per)
        anchor = (np.arange(per) >= per // 2).astype(np.int32)
        pts = np.empty((per, 2))
        ke = per // 2
        e = g.integers(0, ne, ke)
        a, b = V[e], V[(e + 1) % ne]
        t = g.uniform(0.0, 1.0, ke)[:, None]
        tan = (b - a) / np.linalg.norm(b - a, axis=1)[:, None]
        nrm = np.stack([tan[:, 1], -tan[:, 0]], 1)
        side = np.where(g.uniform(size=ke) < 0.5, -1.0, 1.0)[:, None]
        pts[:ke] = a + t * (b - a) + side * d[:ke, None] * nrm
        kv = per - ke
        v = np.where(g.uniform(size=kv) < 0.5, 0, g.integers(0, ne, kv))
        ang = g.uniform(0, 2 * np.pi, kv)
        pts[ke:] = V[v] + d[ke:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
        P.append(pts)
        D.append(d / h)
        A.append(anchor)
        I.append(np.full(per, i, dtype=np.int32))
    return StressPolygons("stress-2d", h, ("cubic", "w2", "w4", "w6"), polys, np.array(angles), np.concatenate(I),
                          np.concatenate(P), np.concatenate(D), np.concatenate(A))


def _rotation(g):
    q, r = np.linalg.qr(g.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def config4_3d(n=262144, n_tet=256, h=0.01, seed=44):
    """precision stress in 3-D (see the module docstring): base triangle A = 0, B = (2h, 0, 0),
    C = (2h u, 2h / aspect, 0) with u ~ U(0.1, 0.9) and aspect = 10^U(0, 4); apex 2h below the base; a random rotation
    and a centre in [0.3, 0.7]^3.  Half the particles are anchored on the base face interior (uniform barycentric
    point, offset d along the normal, either side), half on a base vertex (uniform direction)."""
    g = rng(seed)
    per = n // n_tet
    ell = 2 * h
    Vs, asp, P, D, A, I = [], [], [], [], [], []
    faces = np.array([[0, 1, 2], [0, 3, 1], [1, 3, 2], [2, 3, 0]], dtype=np.int32)
    for i in range(n_tet):
        aspect = 10.0 ** g.uniform(0.0, 4.0)
        u = g.uniform(0.1, 0.9)
        base = np.array([[0.0, 0.0, 0.0], [ell, 0.0, 0.0], [u * ell, ell / aspect, 0.0]])
        apex = np.array([[base[:, 0].mean(), base[:, 1].mean(), -ell]])
        R = _rotation(g)
        c = g.uniform(0.3, 0.7, 3)
        V = np.vstack([base, apex]) @ R.T + c
        Vs.append(V)
        asp.append(aspect)
        nrm = R[:, 2]                          # the base plane's normal (either orientation)
        d = h * 10.0 ** g.uniform(-7.0, 0.0, per)
        ke = per // 2
        w = g.dirichlet(np.ones(3), ke)
        side = np.where(g.uniform(size=ke) < 0.5, -1.0, 1.0)[:, None]
        pts = np.empty((per, 3))
        pts[:ke] = w @ V[:3] + side * d[:ke, None] * nrm
        kv = per - ke
        dirs = g.normal(size=(kv, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        pts[ke:] = V[g.integers(0, 3, kv)] + d[ke:, None] * dirs
        P.append(pts)
        D.append(d / h)
        A.append((np.arange(per) >= ke).astype(np.int32))
        I.append(np.full(per, i, dtype=np.int32))
    V = np.array(Vs)
    # outward orientation: the apex lies below the base, so (B - A) x (C - A) must point away from it
    for i in range(n_tet):
        a, b, c2, dd = V[i]
        if np.dot(np.cross(b - a, c2 - a), dd - a) >= 0:
            raise AssertionError("tetrahedron orientation")
    return StressTetrahedra("stress-3d", h, ("cubic", "w2", "w4", "w6"), V, faces, np.array(asp), np.concatenate(I),
                            np.concatenate(P), np.concatenate(D), np.concatenate(A))


@dataclass
class ServedScene:
    name: str
    loops: list                 # 2-D: [(vertices, 'inside' | 'outside')]; None for a 3-D scene
    h: float
    kernel: str
    positions: np.ndarray       # [n, 2] or [n, 3]
    mesh: tuple = None          # 3-D: (vertices [V, 3], outward faces [F, 3])

    @property
    def dim(self):
        return 2 if self.mesh is None else 3


def box_surface_subdivided(k, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """(vertices, faces) of an axis-aligned box whose six sides are each split into k x k squares of two triangles
    (12 k^2 outward faces), vertices shared along the edges."""
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    index = {}
    V, F = [], []

    def vid(ijk):
        key = tuple(int(v) for v in ijk)
        j = index.get(key)
        if j is None:
            j = index[key] = len(V)
            V.append(lo + (hi - lo) * np.array(key, dtype=np.float64) / k)
        return j
    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        for side in (0, k):
            for a in range(k):
                for b in range(k):
                    q = []
                    for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = [0, 0, 0]
                        p[axis], p[u], p[v] = side, a + da, b + db
                        q.append(vid(p))
                    # (u, v) is counter-clockwise seen from +axis: outward on the high side, reversed on the low side
                    tris = [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
                    if side == 0:
                        tris = [(t[0], t[2], t[1]) for t in tris]
                    F.extend(tris)
    return np.array(V), np.array(F, dtype=np.int32)


def served_scene_3d(i, g, n_min, n_max):
    """a served 3-D scene: the unit box around a spherical cavity of radius 0.25 (icosphere of level 4 or 5: 5,120 or
    20,480 faces) or a box subdivided into 12 k^2 faces with k = 10^U(1, log 81) (1,200 to 78,732 faces, fluid
    inside); n = 10^U(log n_min, log n_max) particles uniform in the fluid; h = dx (30/pi)^(1/3) (about 40
    neighbours, dx = n^(-1/3))."""
    n = int(10 ** g.uniform(np.log10(n_min), np.log10(n_max)))
    kind = int(g.integers(0, 3))
    if kind < 2:
        V, F = box_with_sphere_cavity(level=4 + kind)
        pts = np.empty((0, 3))
        while len(pts) < n:
            c = g.uniform(0.0, 1.0, (int(1.2 * (n - len(pts))) + 64, 3))
            c = c[np.linalg.norm(c - 0.5, axis=1) > 0.25]
            pts = np.concatenate([pts, c])
    else:
        k = int(round(10 ** g.uniform(1.0, np.log10(81.0))))
        V, F = box_surface_subdivided(k)
        pts = g.uniform(0.0, 1.0, (n, 3))
    h = float((30.0 / np.pi) ** (1.0 / 3.0) / n ** (1.0 / 3.0))
    return ServedScene(f"scene{i:03d}", None, h, ["w2", "w4", "w6"][int(g.integers(0, 3))], pts[:n], (V, F))


def config5(n_scenes=256, seed=5, n_min=1000, n_max=200000, e_min=16, e_max=8192, frac3d=0.25):
    """the served mixed-size workload: every 1/frac3d-th scene 3-D (``served_scene_3d``, its own random stream), the
    others 2-D star polygons with E = 10^U(log 16, log 8192) edges, particle counts 10^U(3, log 2e5) uniform in
    [0,1]^2 outside the star, kernel in {w2, w4, w6}, h = dx sqrt(40/pi) (about 40 neighbours, dx = 1/sqrt(n))."""
    g = rng(seed)
    g3 = rng(seed + 1000)
    every = int(round(1.0 / frac3d)) if frac3d > 0 else 0
    scenes = []
    for i in range(n_scenes):
        if every and i % every == every - 1:
            scenes.append(served_scene_3d(i, g3, n_min, n_max))
            continue
        E = int(10 ** g.uniform(np.log10(e_min), np.log10(e_max)))
        n = int(10 ** g.uniform(np.log10(n_min), np.log10(n_max)))
        star = star_polygon(max(E, 3), g=g)
        pts = np.empty((0, 2))
        while len(pts) < n:
            c = g.uniform(0.0, 1.0, (int(1.5 * (n - len(pts))) + 64, 2))
            c = c[~inside_star(c, star)]
            pts = np.concatenate([pts, c])
        h = float(np.sqrt(40.0 / np.pi) / np.sqrt(n))
        scenes.append(ServedScene(f"scene{i:03d}", [(star, "inside")], h, ["w2", "w4", "w6"][int(g.integers(0, 3))], pts[:n]))
    return scenes


def jitter(positions, h, seed):
    """one timestep: positions moved by N(0, (0.01 h)^2)."""
    g = rng(seed)
    return positions + g.normal(0.0, 0.01 * h, positions.shape)


def poisson_arrivals(n, qps, seed=6):
    g = rng(seed)
    return np.cumsum(g.exponential(1.0 / qps, n))


@dataclass
class FibreBundle:
    name: str
    centres: np.ndarray
    radii: np.ndarray
    positions: np.ndarray
    supports: np.ndarray
    h: float
    kernels: tuple


def fibre_bundle(n=1 << 20, side=64, seed=8):
    """a bundle of side^2 solid fibres (disks) on a jittered lattice in the unit square, radii 0.2-0.3 of the lattice
    spacing, n particles uniform in the square outside the fibres, h = 0.006 (about 0.38 lattice spacings): the
    disk-element workload of a DiskArrayRep body (reference scene/scene.py:808-888)."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), indexing="ij"), -1).reshape(-1, 2)
    centres = (g + 0.5 + rng.uniform(-0.15, 0.15, g.shape)) / side
    radii = rng.uniform(0.2, 0.3, len(g)) / side
    pos = np.empty((0, 2))
    while len(pos) < n:
        p = rng.uniform(0.0, 1.0, (n, 2))
        cell = np.clip(np.floor(p * side).astype(np.int64), 0, side - 1)
        near = np.full(len(p), np.inf)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                cx, cy = cell[:, 0] + dx, cell[:, 1] + dy
                ok = (cx >= 0) & (cx < side) & (cy >= 0) & (cy < side)
                k = np.where(ok, cx * side + cy, 0)
                d = np.linalg.norm(p - centres[k], axis=1) - radii[k]
                near = np.minimum(near, np.where(ok, d, np.inf))
        pos = np.concatenate([pos, p[near > 0]])
    h = 0.006
    return FibreBundle(f"fibre bundle {side * side}", centres, radii, pos[:n], np.full(n, h), h,
                       ("cubic", "w2", "w4", "w6"))
