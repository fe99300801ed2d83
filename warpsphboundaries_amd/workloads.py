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
class ServedScene:
    name: str
    loops: list
    h: float
    kernel: str
    positions: np.ndarray


def config5(n_scenes=256, seed=5, n_min=1000, n_max=200000, e_min=16, e_max=8192):
    """the served mixed-size workload (2-D scenes): star polygons with E = 10^U(log 16, log 8192) edges, particle
    counts 10^U(3, log 2e5) uniform in [0,1]^2 outside the star, kernel in {w2, w4, w6}, h = dx sqrt(40/pi)
    (about 40 neighbours, dx = 1/sqrt(n))."""
    g = rng(seed)
    scenes = []
    for i in range(n_scenes):
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
