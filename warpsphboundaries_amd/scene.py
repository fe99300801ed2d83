"""Scene layer (2-D): bodies with a pose, closed-loop surface representations, the fixed-capacity adjacency.

Mirrors the reference's ``scene/scene.py`` (``Pose`` :47-87, ``SurfaceRep`` :213-416, ``Body`` :913-1067,
``Scene`` :1227-1372) and ``scene/fixedadj.py`` (``FixedTopology``, ``fixed_adjacency`` :249-270) for the
surface representations of the wall-integral path.  Representations are stored in the body frame; every
device computation (adjacency, indicator, moments, signed distance) runs in the HIP library.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Optional

import numpy as np
import torch

from . import _lib as L
from . import triangle
from .device import Disks2D, Geom2D
from .kernels import KERNEL_ID, plane_channels

F64 = torch.float64

KERNEL_NAMES = {"CubicSpline": "cubic", "cubic": "cubic", "Wendland2": "w2", "Wendland4": "w4", "Wendland6": "w6",
                "w2": "w2", "w4": "w4", "w6": "w6",
                # the wall-operation kernels (cover, wall Laplacian, tensile) of the full aggregate layout
                "cone": "cone", "lw2": "lw2", "lw4": "lw4", "w2p5": "w2p5", "w4p5": "w4p5"}


def kernelName(kernel) -> str:
    """registered kernel name of a string or an enum-like kernel (``.name`` Wendland2/4/6, CubicSpline)
    (reference scene/boundaryOps.py:35-44)."""
    key = kernel if isinstance(kernel, str) else getattr(kernel, "name", str(kernel))
    try:
        return KERNEL_NAMES[key]
    except KeyError:
        raise NotImplementedError(f"{kernel}: not one of the exact piecewise-polynomial kernels {sorted(set(KERNEL_NAMES.values()))}") from None


@dataclass
class ParticleState:
    """the duck type the provider reads (warpSPHCore ParticleState): positions [N,2], supports [N], kinds [N] (0 fluid)."""
    positions: torch.Tensor
    supports: torch.Tensor
    masses: Optional[torch.Tensor] = None
    densities: Optional[torch.Tensor] = None
    kinds: Optional[torch.Tensor] = None


# ----------------------------------------------------------------------------------------------------------------------- frames
@dataclass
class Pose:
    center: torch.Tensor
    angle: float = 0.0

    def _cs(self):
        a = float(self.angle)
        return math.cos(a), math.sin(a)

    @property
    def R(self):
        c, s = self._cs()
        return torch.tensor([[c, -s], [s, c]], dtype=F64, device=self.center.device)

    def Rflat(self):
        return self.R.reshape(-1)

    def toLocal(self, p):
        c, s = self._cs()
        d = p - self.center
        return torch.stack([d[:, 0] * c + d[:, 1] * s, d[:, 1] * c - d[:, 0] * s], 1)

    def toWorld(self, p):
        c, s = self._cs()
        return torch.stack([p[:, 0] * c - p[:, 1] * s + self.center[0], p[:, 0] * s + p[:, 1] * c + self.center[1]], 1)

    def vecToWorld(self, v):
        c, s = self._cs()
        return torch.stack([v[:, 0] * c - v[:, 1] * s, v[:, 0] * s + v[:, 1] * c], 1)


# ----------------------------------------------------------------------------------------------------------------------- surfaces
def _orientation_sign(P: np.ndarray) -> float:
    """sign of the signed area of a closed loop, robust for tiny / translated / nearly collinear loops
    (anchor at the first vertex, long double; exact rationals when that is not conclusive -- reference
    scene/scene.py:229-244)."""
    local = P.astype(np.longdouble) - P[0].astype(np.longdouble)
    terms = local[:, 0] * np.roll(local[:, 1], -1) - np.roll(local[:, 0], -1) * local[:, 1]
    a = np.sum(terms)
    if len(P) == 3 or abs(a) <= 64 * np.finfo(np.longdouble).eps * np.sum(abs(terms)):
        v = [[Fraction(float(c)) for c in p] for p in P]
        a = sum((v[i][0] - v[0][0]) * (v[(i + 1) % len(v)][1] - v[0][1]) - (v[(i + 1) % len(v)][0] - v[0][0]) * (v[i][1] - v[0][1])
                for i in range(len(v)))
    return float(np.sign(a))


class SurfaceRep:
    """closed polylines in the body frame, counter-clockwise around the SOLID (solid on the left of every edge);
    ``background = 1`` when the solid is the unbounded outside (a tank wall / cavity)."""
    kind = "surface"

    def __init__(self, vertices, edges, background: int = 0, conditioned: bool = True):
        self.vertices = torch.as_tensor(np.asarray(vertices, dtype=np.float64), dtype=F64)
        self.edges = torch.as_tensor(np.asarray(edges, dtype=np.int32), dtype=torch.int32)
        self.background = int(background)
        self.conditioned = bool(conditioned)
        # a closed 3-vertex loop: rows whose support makes it small or a sliver take the conditioned quadrature
        # (reference scene/scene.py:228 and edge/warptriangle.py); None: the closed form everywhere
        self._triangle = triangle.prepare(self.vertices, self.edges) if conditioned else None
        self._geoms = {}

    @staticmethod
    def polygon(points, solid: str = "inside", conditioned: bool = True):
        """one closed loop; ``solid='inside'``: the polygon is the body (made counter-clockwise), ``'outside'``: a hole
        in an infinite solid (clockwise, background 1)."""
        P = np.asarray(points, dtype=np.float64)
        if (_orientation_sign(P) > 0) != (solid == "inside"):
            P = P[::-1].copy()
        n = len(P)
        E = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
        return SurfaceRep(np.ascontiguousarray(P), E, background=0 if solid == "inside" else 1, conditioned=conditioned)

    @staticmethod
    def box(lo, hi, solid: str = "inside"):
        (x0, y0), (x1, y1) = lo, hi
        return SurfaceRep.polygon([(x0, y0), (x1, y0), (x1, y1), (x0, y1)], solid)

    @staticmethod
    def regularPolygon(center, radius, n, solid="inside", areaPreserving=True):
        r = radius * np.sqrt(2 * np.pi / (n * np.sin(2 * np.pi / n))) if areaPreserving else radius
        t = 2 * np.pi * np.arange(n) / n
        return SurfaceRep.polygon(np.stack([center[0] + r * np.cos(t), center[1] + r * np.sin(t)], 1), solid)

    def to(self, device):
        self.vertices = self.vertices.to(device)
        self.edges = self.edges.to(device)
        return self

    def bounds(self):
        return self.vertices.amin(0), self.vertices.amax(0)

    # ---- device structures
    def geom(self, support_max: float, device) -> Geom2D:
        """static device structures for ``support_max`` (built once per (support_max, device))."""
        dev = torch.device(device)
        idx = dev.index if dev.index is not None else torch.cuda.current_device()
        key = (float(support_max), idx)
        g = self._geoms.get(key)
        if g is None:
            g = Geom2D(self.vertices.detach().cpu().numpy(), self.edges.detach().cpu().numpy(), self.background,
                       float(support_max), idx)
            self._geoms[key] = g
        return g

    def indicator(self, p, supports=None):
        """winding number + background (1 inside the solid); on an edge 1/2, at a vertex the angle fraction."""
        L.require_device(p)
        sup = torch.ones(len(p), dtype=F64, device=p.device) if supports is None else supports.to(F64)
        smax = float(sup.max()) if len(p) else 1.0
        return self.geom(smax, p.device).adjacency(p, sup, pad=False)[2]

    def topology(self, lpos, lsup):
        """`SurfaceTopology`: the (query, edge) pairs with the segment closer than the support, from the fixed slots
        (pair order = slot order, as the reference's sorted pair list)."""
        if len(lpos) == 0:
            return None
        smax = float(lsup.max())
        slots, counts, ind = self.geom(smax, lpos.device).adjacency(lpos, lsup, pad=True)
        K = self.geom(smax, lpos.device).K
        e = slots.view(-1, K)
        keep = e >= 0
        qi = torch.arange(len(lpos), device=lpos.device, dtype=torch.int32)[:, None].expand(-1, K)[keep]
        return SurfaceTopology(qi.contiguous(), e[keep].contiguous(), smax, ind)


@dataclass
class BoxRep:
    """an axis-aligned rectangle [lo, hi] of the body frame: the solid (``solid='inside'``) or the fluid domain
    (``'outside'``: a tank).  The reference (scene/scene.py:459-535, scene/box.py) evaluates it from Chebyshev corner
    tables (error to 3e-7 on the gradient channels); here it is lowered to its 4-edge polygon (``surface()``), whose
    closed form is exact, so a box and its polygon give the same results."""
    kind = "box"

    def __init__(self, lo, hi, solid: str = "inside"):
        if solid not in ("inside", "outside"):
            raise ValueError("solid must be 'inside' or 'outside'")
        self.lo = torch.as_tensor(lo, dtype=F64)
        self.hi = torch.as_tensor(hi, dtype=F64)
        self.lo_h = tuple(float(v) for v in self.lo)
        self.hi_h = tuple(float(v) for v in self.hi)
        if not (self.hi_h[0] > self.lo_h[0] and self.hi_h[1] > self.lo_h[1]):
            raise ValueError("BoxRep needs hi > lo on both axes")
        self.solid = solid
        self._surface = None

    def surface(self):
        if self._surface is None:
            self._surface = SurfaceRep.box(self.lo_h, self.hi_h, self.solid).to(self.lo.device)
        return self._surface

    def to(self, device):
        self.lo, self.hi = self.lo.to(device), self.hi.to(device)
        if self._surface is not None:
            self._surface.to(device)
        return self

    def bounds(self):
        return self.lo, self.hi

    def lowered(self, hmin, device):
        return self.surface().to(device)

    def indicator(self, p, supports=None):
        return self.surface().indicator(p, supports)


@dataclass
class TierPolicy:
    """the reference's tier thresholds (scene/implicitBodies.py:292-297); only the tier-2 polygon's edge bound is used
    here: implicit shapes are lowered to polygons whose edges are at most ``maxPolygonEdgeOverH`` times the smallest
    support."""
    tier4MaxRadius: float = 0.2
    tier3MinRadius: float = 2.0
    maxPolygonEdgeOverH: float = 1.0 / 16.0


@dataclass
class DiskBody:
    """a disk: ``solid='inside'`` an obstacle, ``'outside'`` a cavity wall with the fluid inside the circle
    (reference scene/implicitBodies.py:311-316)."""
    center: tuple
    radius: float
    solid: str = "inside"
    bodyId: int = 0


class ImplicitRep:
    """an analytic primitive of the body frame (reference scene/scene.py:558-596).  A SOLID disk is evaluated as a
    one-disk ``DiskArrayRep`` (the disk element, any radius; reference scene/scene.py:576-587), a cavity as the
    reference's tier-2 polygon: area preserving, ``n = max(24, ceil(2 pi R / (maxPolygonEdgeOverH hmin)))`` edges
    (``fallbackSurface``, scene/scene.py:588-596); other primitives have no lowering and raise NotImplementedError."""
    kind = "implicit"

    def __init__(self, shape, policy=None):
        if not isinstance(shape, DiskBody):
            raise NotImplementedError("implicit representations: disks only (the half plane has no polygon)")
        if shape.solid not in ("inside", "outside"):
            raise ValueError("solid must be 'inside' or 'outside'")
        self.shape = shape
        self.policy = policy or TierPolicy()
        self._fallback = {}
        self._disks = None
        self._device = "cpu"

    @property
    def solid(self):
        return self.shape.solid

    def bounds(self):
        c = torch.as_tensor(self.shape.center, dtype=F64)
        return c - self.shape.radius, c + self.shape.radius

    def to(self, device):
        self._device = device
        for rep in self._fallback.values():
            rep.to(device)
        if self._disks is not None:
            self._disks.to(device)
        return self

    def disks(self):
        """a solid disk as a one-disk ``DiskArrayRep`` (cached)."""
        if self._disks is None:
            self._disks = DiskArrayRep([self.shape.center], [self.shape.radius]).to(self._device)
        return self._disks

    def fallbackSurface(self, hmin, device):
        s = self.shape
        n = max(24, int(np.ceil(2 * np.pi * s.radius / (self.policy.maxPolygonEdgeOverH * float(hmin)))))
        if n not in self._fallback:
            self._fallback[n] = SurfaceRep.regularPolygon(s.center, s.radius, n, solid=s.solid).to(device)
        return self._fallback[n]

    def lowered(self, hmin, device):
        if self.shape.solid == "inside":
            return self.disks().to(device)
        return self.fallbackSurface(hmin, device)

    def signed(self, lpos):
        """(d, n): exact signed distance of the circle (positive in the fluid) and the unit normal into the fluid,
        body frame (reference scene/implicitBodies.py:317-330)."""
        c = torch.as_tensor(self.shape.center, dtype=F64, device=lpos.device)
        v = lpos.to(F64) - c
        r = v.norm(dim=1).clamp(min=1e-300)
        e = v / r[:, None]
        if self.shape.solid == "inside":
            return r - self.shape.radius, e
        return self.shape.radius - r, -e

    def indicator(self, p, supports=None):
        """1 in the solid, 0 in the fluid (the circle itself: 1/2)."""
        d, _ = self.signed(p)
        return torch.where(d < 0, 1.0, torch.where(d > 0, 0.0, 0.5)).to(F64)


class SdfRep:
    """a sampled signed distance of the body frame: node values ``values`` [nx, ny] (positive in the fluid) on the grid
    ``origin + spacing * (i, j)`` (reference scene/scene.py:599-670).  The fused evaluation integrates it exactly as
    a surface: its ``fallback`` SurfaceRep when given, else the zero level set of the samples (``contour``, marching
    squares).  ``signed`` is the bilinear interpolant with the gradient of the bilinear cell as normal direction."""
    kind = "sdf"

    def __init__(self, values, origin, spacing, fallback: Optional[SurfaceRep] = None, **policy):
        self.d = torch.as_tensor(values, dtype=F64)
        if self.d.dim() != 2 or min(self.d.shape) < 2:
            raise ValueError("SdfRep: values must be a [nx, ny] grid with nx, ny >= 2")
        self.origin = torch.as_tensor(origin, dtype=F64).reshape(2)
        self.spacing = float(spacing)
        if not self.spacing > 0:
            raise ValueError("SdfRep: spacing must be positive")
        self.fallback = fallback
        self.policy = dict(policy)          # the reference's tier-3 guard thresholds; the fused path does not use them
        self._contour = None

    @staticmethod
    def fromFunction(fn, lo, hi, spacing, **kw):
        """``fn(points [M, 2]) -> d`` sampled on the grid covering [lo, hi] (reference scene/scene.py:636-642)."""
        nx = int(np.ceil((hi[0] - lo[0]) / spacing)) + 1
        ny = int(np.ceil((hi[1] - lo[1]) / spacing)) + 1
        X, Y = np.meshgrid(lo[0] + spacing * np.arange(nx), lo[1] + spacing * np.arange(ny), indexing="ij")
        pts = torch.as_tensor(np.stack([X.ravel(), Y.ravel()], 1), dtype=F64)
        return SdfRep(torch.as_tensor(fn(pts), dtype=F64).reshape(nx, ny), lo, spacing, **kw)

    @property
    def solid(self):
        return "outside" if self.contour().background else "inside"

    def to(self, device):
        self.d, self.origin = self.d.to(device), self.origin.to(device)
        if self.fallback is not None:
            self.fallback.to(device)
        if self._contour is not None:
            self._contour.to(device)
        return self

    def bounds(self):
        ext = torch.as_tensor(self.d.shape, dtype=F64, device=self.d.device) - 1
        return self.origin, self.origin + ext * self.spacing

    def contour(self):
        """the zero level set of the samples as a SurfaceRep: on every cell edge whose end nodes differ in sign
        (solid = d < 0) a vertex at the linear zero, joined cell by cell with the solid on the left of every edge; a
        cell with four crossings joins them around its solid corners, or across its centre when the mean of its four
        nodes is solid.  The border nodes must be all fluid (an enclosed solid, background 0) or all solid (a tank,
        background 1).  Vertices and edges are numbered in the order the cells are visited (i outer, j inner,
        corners counter-clockwise), as the reference's marching squares (scene/scene.py:762-806)."""
        if self._contour is None:
            self._contour = _contour_of(self.d.cpu().numpy(), self.origin.cpu().numpy(), self.spacing).to(self.d.device)
        return self._contour

    def lowered(self, hmin, device):
        return (self.fallback if self.fallback is not None else self.contour()).to(device)

    def signed(self, pos):
        """(d, n): the bilinear interpolant of the samples (clamped to the grid) and the unit normal into the fluid
        (its normalised gradient), body frame."""
        d = self.d.to(pos.device)
        nx, ny = d.shape
        u = (pos.to(F64) - self.origin.to(pos.device)) / self.spacing
        uc = torch.stack([u[:, 0].clamp(0, nx - 1), u[:, 1].clamp(0, ny - 1)], 1)
        i0 = torch.stack([uc[:, 0].floor().clamp(0, nx - 2), uc[:, 1].floor().clamp(0, ny - 2)], 1).long()
        t = uc - i0
        a, b = i0[:, 0], i0[:, 1]
        d00, d10, d01, d11 = d[a, b], d[a + 1, b], d[a, b + 1], d[a + 1, b + 1]
        tx, ty = t[:, 0], t[:, 1]
        val = (1 - tx) * (1 - ty) * d00 + tx * (1 - ty) * d10 + (1 - tx) * ty * d01 + tx * ty * d11
        gx = ((1 - ty) * (d10 - d00) + ty * (d11 - d01)) / self.spacing
        gy = ((1 - tx) * (d01 - d00) + tx * (d11 - d10)) / self.spacing
        g = torch.stack([gx, gy], 1)
        return val, g / g.norm(dim=1, keepdim=True).clamp(min=1e-300)

    def indicator(self, p, supports=None):
        d, _ = self.signed(p)
        return torch.where(d < 0, 1.0, torch.where(d > 0, 0.0, 0.5)).to(F64)


def _contour_of(d, origin, spacing):
    """the SurfaceRep of the zero level set of node values d [nx, ny] (see ``SdfRep.contour``)."""
    nx, ny = d.shape
    solid = d < 0
    border = np.concatenate([solid[0], solid[-1], solid[1:-1, 0], solid[1:-1, -1]])
    if border.any() and not border.all():
        raise ValueError("SdfRep.contour: the sampled grid must cover the body with a margin (all border nodes fluid, "
                         "or all solid)")
    background = 1 if border.all() else 0
    index, verts, edges = {}, [], []

    def vertex(p, q):
        if (p, q) not in index:
            t = d[p] / (d[p] - d[q])
            index[(p, q)] = len(verts)
            verts.append((origin[0] + spacing * (p[0] + t * (q[0] - p[0])), origin[1] + spacing * (p[1] + t * (q[1] - p[1]))))
        return index[(p, q)]

    mixed = solid[:-1, :-1].astype(np.int8) + solid[1:, :-1] + solid[1:, 1:] + solid[:-1, 1:]
    for i, j in zip(*np.nonzero((mixed > 0) & (mixed < 4))):
        i, j = int(i), int(j)
        ring = ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1))
        sol = [bool(solid[c]) for c in ring]
        cross = []                                  # (leaving the solid?, vertex) counter-clockwise along the cell
        for k in range(4):
            if sol[k] != sol[(k + 1) % 4]:
                cross.append((sol[k], vertex(ring[k], ring[(k + 1) % 4])))
        m = len(cross)
        step = -1 if (m == 4 and float(d[i:i + 2, j:j + 2].mean()) < 0) else 1
        for k in range(m):
            if cross[k][0]:
                q = (k + step) % m
                while cross[q][0]:
                    q = (q + step) % m
                edges.append((cross[k][1], cross[q][1]))
    if not edges:
        raise ValueError("SdfRep.contour: no zero level set in the sampled grid")
    return SurfaceRep(np.asarray(verts, dtype=float), np.asarray(edges, dtype=np.int32), background=background)


class DiskArrayRep:
    """a set of solid disks (fibres) of the body frame: centres [M, 2], radii [M]; the disks must not overlap (the
    solid is their union and the integrals add).  Every (particle, disk) pair is one closed evaluation of the disk
    element (``Disks2D.moments``; no polygon), the disks near a particle come from a cell list over the centres with a
    fixed capacity (``Disks2D.adjacency``) (reference scene/scene.py:808-888, scene/fixedadj.py:199-246)."""
    kind = "disks"

    def __init__(self, centres, radii):
        self.centres = torch.as_tensor(centres, dtype=F64).reshape(-1, 2)
        self.radii = torch.as_tensor(radii, dtype=F64).reshape(-1)
        if len(self.centres) != len(self.radii):
            raise ValueError("DiskArrayRep: centres and radii differ in length")
        if len(self.radii) == 0:
            raise ValueError("DiskArrayRep: no disks")
        self._geoms = {}
        self._bounds = None

    solid = "inside"
    background = 0

    def to(self, device):
        self.centres, self.radii = self.centres.to(device), self.radii.to(device)
        self._bounds = None
        return self

    def bounds(self):
        if self._bounds is None:
            r = self.radii[:, None]
            self._bounds = ((self.centres - r).amin(0), (self.centres + r).amax(0))
        return self._bounds

    def geom(self, support_max: float, device) -> Disks2D:
        dev = torch.device(device)
        key = (float(support_max), dev.index or 0)
        if key not in self._geoms:
            self._geoms[key] = Disks2D(self.centres.cpu().numpy(), self.radii.cpu().numpy(), float(support_max),
                                       dev.index or 0)
        return self._geoms[key]

    def signed(self, pos):
        """(d, n): signed distance (positive in the fluid) to the nearest circle and the unit normal into the fluid
        there, body frame (reference scene/scene.py:853-867), in chunks of 4096 points."""
        N = len(pos)
        out_d = torch.full((N,), float("inf"), dtype=F64, device=pos.device)
        out_n = torch.zeros((N, 2), dtype=F64, device=pos.device)
        cen = self.centres.to(pos.device)
        rad = self.radii.to(pos.device)
        for lo in range(0, N, 4096):
            v = pos[lo:lo + 4096, None, :].to(F64) - cen[None]
            dist = v.norm(dim=2)
            d = dist - rad[None]
            k = d.argmin(1)
            ar = torch.arange(len(d), device=pos.device)
            out_d[lo:lo + 4096] = d[ar, k]
            out_n[lo:lo + 4096] = v[ar, k] / dist[ar, k].clamp(min=1e-300)[:, None]
        return out_d, out_n

    def indicator(self, p, supports=None):
        """1 in a disk, 0 in the fluid (on a circle: 1/2)."""
        d, _ = self.signed(p)
        return torch.where(d < 0, 1.0, torch.where(d > 0, 0.0, 0.5)).to(F64)

    def lowered(self, hmin, device):
        return self


class SurfaceTopology:
    qi: torch.Tensor
    e: torch.Tensor
    supportMax: float
    ind: torch.Tensor

    def indicator(self, rep=None, lpos=None, lsup=None):
        return self.ind


@dataclass
class FixedTopology:
    """K slots per row: ``e[r*K + s]`` (-1 = empty), ``counts`` [N], ``ind`` [N] (reference scene/fixedadj.py:39-64)."""
    e: torch.Tensor
    counts: torch.Tensor
    K: int
    ind: torch.Tensor
    geom: Geom2D = None

    @property
    def qi(self):
        return torch.arange(len(self.counts) * self.K, device=self.e.device, dtype=torch.int32) // self.K

    def indicator(self, rep=None, lpos=None, lsup=None):
        return self.ind


# ----------------------------------------------------------------------------------------------------------------------- bodies
@dataclass
class Body:
    bodyId: int = 0
    center: tuple = (0.0, 0.0)
    angle: float = 0.0
    linearVelocity: tuple = (0.0, 0.0)
    angularVelocity: float = 0.0
    reps: list = field(default_factory=list)

    def __post_init__(self):
        self.center = torch.as_tensor(self.center, dtype=F64)
        self.linearVelocity = torch.as_tensor(self.linearVelocity, dtype=F64)

    @property
    def pose(self):
        return Pose(self.center, float(self.angle))

    def toLocal(self, world):
        return self.pose.toLocal(world)

    def unbounded(self):
        return any((isinstance(r, SurfaceRep) and r.background) or getattr(r, "solid", "inside") == "outside"
                   for r in self.reps)

    def fusedReps(self, hmin, device):
        """the representations this body is evaluated with: a SurfaceRep or DiskArrayRep as it is, a BoxRep its
        polygon, a solid ImplicitRep disk a one-disk DiskArrayRep, a cavity its tier-2 polygon for the smallest support
        ``hmin`` (reference scene/scene.py:970-996)."""
        return [r if isinstance(r, SurfaceRep) else r.lowered(hmin, device) for r in self.reps]

    def obb(self):
        los, his = zip(*[r.bounds() for r in self.reps])
        dev = self.center.device
        return torch.stack([a.to(dev) for a in los]).amin(0), torch.stack([a.to(dev) for a in his]).amax(0)

    def velocityAt(self, world):
        w = float(self.angularVelocity)
        r = world - self.center
        return self.linearVelocity + w * torch.stack([-r[:, 1], r[:, 0]], 1)


@dataclass
class BodyAdjacency:
    body: object
    cand: torch.Tensor
    lpos: torch.Tensor
    lsup: torch.Tensor
    reps: list
    valid: Optional[torch.Tensor] = None
    repList: Optional[list] = None


@dataclass
class SceneAdjacency:
    numQueries: int
    bodies: list
    stats: dict
    queryParticles: object = None
    supportMax: float = 0.0


class Scene:
    def __init__(self, bodies: List[Body], device="cuda:0"):
        self.bodies = bodies
        self.device = device
        for b in bodies:
            b.center = b.center.to(device)
            b.linearVelocity = b.linearVelocity.to(device)
            for r in b.reps:
                r.to(device)

    def precompute(self, adjacency, operationProperties, channels=None, form=None):
        """per-representation moments of one kernel (``pairmoments.PairMoments``; reference scene/scene.py:1418)."""
        from .pairmoments import precompute
        return precompute(self, adjacency, operationProperties, channels, form)

    def pairMoments(self, queryParticles, operationProperties, channels=None, form=None):
        from .pairmoments import pair_moments
        return pair_moments(self, queryParticles, operationProperties, channels, form)

    def moments(self, source, queryParticles, operationProperties, channels=None, form=None):
        from .pairmoments import moments
        return moments(self, source, queryParticles, operationProperties, channels, form)

    def inside(self, points, body=None):
        pts = points.to(self.device, F64)
        out = torch.zeros(len(pts), dtype=torch.bool, device=pts.device)
        for bi, bd in enumerate(self.bodies):
            if body is not None and bi != body:
                continue
            lp = bd.toLocal(pts)
            for rep in bd.reps:
                out |= rep.indicator(lp) > 0.5
        return out

    def signed_distance(self, points, body=None, supportMax=None, want_body=False):
        """(d [M], n [M,2], hit [M][, body [M]]): d > 0 in the fluid, n the unit normal into the fluid (world frame);
        the nearest edge of every surface loop, sign by the winding number at ``supportMax`` normalisation
        (reference scene/scene.py:1312-1372)."""
        pts = points.to(self.device, F64)
        M = len(pts)
        bidx = torch.full((M,), -1, dtype=torch.long, device=pts.device)
        best = torch.full((M,), float("inf"), dtype=F64, device=pts.device)
        normal = torch.zeros((M, 2), dtype=F64, device=pts.device)
        hit = torch.zeros(M, dtype=torch.bool, device=pts.device)
        for bi, bd in enumerate(self.bodies):
            if body is not None and bi != body:
                continue
            lp = bd.toLocal(pts)
            for rep in bd.reps:
                if isinstance(rep, BoxRep):
                    rep = rep.surface()
                if isinstance(rep, (ImplicitRep, DiskArrayRep, SdfRep)):
                    d, n_loc = rep.signed(lp)
                elif isinstance(rep, SurfaceRep):
                    smax = float(supportMax) if supportMax is not None else _default_support(rep)
                    d, n_loc, _ = rep.geom(smax, pts.device).signed_distance(lp, smax)
                else:
                    continue
                n = bd.pose.vecToWorld(n_loc)
                better = d < best
                best = torch.where(better, d, best)
                normal = torch.where(better[:, None], n, normal)
                bidx = torch.where(better, torch.full_like(bidx, bi), bidx)
                hit |= True
        if want_body:
            return best, normal, hit, bidx
        return best, normal, hit


def _default_support(rep: SurfaceRep) -> float:
    """support normalisation of the winding number when none is given: the largest edge box extent (any positive
    value gives the same winding number away from the boundary)."""
    lo, hi = rep.bounds()
    return float(max((hi - lo).max().item(), 1e-12))


def queryAllowed(queryParticles, dev):
    """fluid queries only (BoundaryToFluid): kinds == 0, or every query when kinds are absent."""
    kinds = getattr(queryParticles, "kinds", None)
    n = queryParticles.positions.shape[0]
    if kinds is None:
        return torch.ones(n, dtype=torch.bool, device=dev)
    return kinds.to(dev) == 0


def fixed_adjacency(scene: Scene, queryParticles, operationProperties=None, supportMax: float = None):
    """`SceneAdjacency` of every query against every body with fixed shapes: rows = queries, `valid` = broadphase
    (support circle against the body box in the body frame; every row for an exterior solid) and fluid kind;
    one adjacency launch per (body, representation), no host synchronisation (reference scene/fixedadj.py:249-270)."""
    dev = torch.device(scene.device)
    pos = queryParticles.positions.to(dev, F64)
    sup = queryParticles.supports.to(dev, F64)
    if pos.shape[1] != 2:
        raise NotImplementedError("2D only")
    if supportMax is None:
        supportMax = float(sup.max())
    allowed = queryAllowed(queryParticles, dev)
    N = len(pos)
    cand = torch.arange(N, device=dev)
    bodies = []
    for body in scene.bodies:
        lpos = body.toLocal(pos)
        if body.unbounded():
            valid = allowed
        else:
            lo, hi = body.obb()
            d = (lo - lpos).clamp(min=0) + (lpos - hi).clamp(min=0)
            valid = (d.norm(dim=1) < sup) & allowed
        reps = []
        hmin = float(sup.min()) if any(isinstance(r, ImplicitRep) for r in body.reps) and N else float(supportMax)
        surfaces = body.fusedReps(hmin, dev)
        for rep in surfaces:
            g = rep.geom(float(supportMax), dev)
            if isinstance(rep, DiskArrayRep):
                slots, counts = g.adjacency(lpos, sup, valid=valid)
                ind = torch.zeros(N, dtype=F64, device=dev)            # the disk element is complete
            else:
                slots, counts, ind = g.adjacency(lpos, sup, valid=valid, pad=True)
            reps.append(FixedTopology(slots, counts, g.K, ind, g))
        bodies.append(BodyAdjacency(body, cand, lpos, sup, reps, valid, surfaces))
    return SceneAdjacency(N, bodies, {"candidates": None}, queryParticles, float(supportMax))


def plane_vector(kernel: str, device) -> torch.Tensor:
    return torch.tensor(plane_channels(kernel), dtype=F64, device=device)
