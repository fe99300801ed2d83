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
from .periodic import Periodic, min_image

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

    def vecToWorld(self, v):
        """R v of body-frame vectors v [N, 2]."""
        c, s = self._cs()
        x, y = v[:, 0], v[:, 1]
        return torch.stack((c * x - s * y, s * x + c * y), dim=1)

    def toWorld(self, p):
        """body-frame points [N, 2] to the world: R p + center."""
        return self.vecToWorld(p) + self.center


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

    def tier2Edges(self, hmin):
        """edges of the tier-2 polygon: at least 24, and no edge longer than maxPolygonEdgeOverH * hmin along the
        circle."""
        circumference = 2 * np.pi * self.shape.radius
        return max(24, int(np.ceil(circumference / (self.policy.maxPolygonEdgeOverH * float(hmin)))))

    def fallbackSurface(self, hmin, device):
        """the area-preserving regular polygon of the disk at the resolution of ``hmin`` (one per edge count)."""
        n = self.tier2Edges(hmin)
        rep = self._fallback.get(n)
        if rep is None:
            rep = self._fallback[n] = SurfaceRep.regularPolygon(self.shape.center, self.shape.radius, n,
                                                                solid=self.shape.solid)
        return rep.to(device)

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
    ``origin + spacing * (i, j)`` (reference scene/scene.py:599-670).  The fused evaluation integrates it exactly as a
    surface: its ``fallback`` SurfaceRep when given, else the zero level set of the samples (``contour``, marching
    squares).  ``signed`` blends node jets (value, central-difference gradient and Laplacian) bilinearly.  The keyword
    arguments are the reference's tier-3 guard thresholds, validated as there; the tier-3 local model itself is not
    built, so the fused path does not use them."""
    kind = "sdf"

    def __init__(self, values, origin, spacing, fallback: Optional[SurfaceRep] = None, maxKappaH: float = 0.5,
                 gradTol: float = 0.1, minCurvatureDenominator: float = 0.25, maxCurvatureVariationH2: float = 0.5,
                 backend: str = "warp"):
        if backend not in ("warp", "torch"):
            raise ValueError(f"SdfRep backend {backend!r}: 'warp' or 'torch'")
        if not 0.0 < float(minCurvatureDenominator) <= 1.0:
            raise ValueError("SdfRep: minCurvatureDenominator must lie in (0, 1]")
        if not (math.isfinite(float(maxCurvatureVariationH2)) and float(maxCurvatureVariationH2) > 0.0):
            raise ValueError("SdfRep: maxCurvatureVariationH2 must be finite and positive")
        self.d = torch.as_tensor(values, dtype=F64)
        if self.d.dim() != 2 or min(self.d.shape) < 2:
            raise ValueError("SdfRep: values must be a [nx, ny] grid with nx, ny >= 2")
        self.origin = torch.as_tensor(origin, dtype=F64).reshape(2)
        self.spacing = float(spacing)
        if not self.spacing > 0:
            raise ValueError("SdfRep: spacing must be positive")
        self.fallback = fallback
        self.backend = backend
        self.maxKappaH, self.gradTol = float(maxKappaH), float(gradTol)
        self.minCurvatureDenominator = float(minCurvatureDenominator)
        self.maxCurvatureVariationH2 = float(maxCurvatureVariationH2)
        self._contour = None
        self._jets = None

    @staticmethod
    def fromFunction(fn, lo, hi, spacing, **kw):
        """the SdfRep of ``fn(points [M, 2]) -> d`` on the smallest grid of the given spacing that starts at ``lo`` and
        reaches ``hi`` (reference scene/scene.py:636-642)."""
        counts = [int(np.ceil((b - a) / spacing)) + 1 for a, b in zip(lo, hi)]
        axes = [a + spacing * np.arange(n) for a, n in zip(lo, counts)]
        nodes = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 2)
        values = torch.as_tensor(fn(torch.as_tensor(nodes, dtype=F64)), dtype=F64).reshape(counts)
        return SdfRep(values, lo, spacing, **kw)

    @property
    def solid(self):
        """'outside' when the solid is unbounded (a tank): decided by the fallback when one is given, else by the
        border of the samples."""
        background = self.fallback.background if self.fallback is not None else self.contour().background
        return "outside" if background else "inside"

    def to(self, device):
        self.d, self.origin = self.d.to(device), self.origin.to(device)
        if self.fallback is not None:
            self.fallback.to(device)
        if self._contour is not None:
            self._contour.to(device)
        self._jets = None
        return self

    def bounds(self):
        ext = torch.as_tensor(self.d.shape, dtype=F64, device=self.d.device) - 1
        return self.origin, self.origin + ext * self.spacing

    def contour(self):
        """the zero level set of the samples as a SurfaceRep (see ``_contour_of``)."""
        if self._contour is None:
            self._contour = _contour_of(self.d.cpu().numpy(), self.origin.cpu().numpy(), self.spacing).to(self.d.device)
        return self._contour

    def lowered(self, hmin, device):
        return (self.fallback if self.fallback is not None else self.contour()).to(device)

    def _node_jets(self):
        """[4, nx, ny]: node value, central-difference gradient and five-point Laplacian; at the border the missing
        neighbour is the node itself (the half one-sided difference there)."""
        if self._jets is None:
            d, s = self.d, self.spacing
            xp, xm = torch.cat((d[1:], d[-1:]), 0), torch.cat((d[:1], d[:-1]), 0)
            yp, ym = torch.cat((d[:, 1:], d[:, -1:]), 1), torch.cat((d[:, :1], d[:, :-1]), 1)
            self._jets = torch.stack((d, (xp - xm) / (2 * s), (yp - ym) / (2 * s), (xp + xm + yp + ym - 4 * d) / (s * s)))
        return self._jets

    def signed(self, pos):
        """(d, n, kappa, gradNorm) at body-frame points [M, 2]: the node jets blended bilinearly over the cell of the
        point clamped to the grid, d grown by the distance from the grid beyond it, n the unit gradient (into the
        fluid), kappa = lap / (1 - d lap) (reference scene/scene.py:680-706)."""
        jets = self._node_jets().to(pos.device)
        nx, ny = self.d.shape
        u = (pos.to(F64) - self.origin.to(pos.device)) / self.spacing
        uc = torch.stack((u[:, 0].clamp(0, nx - 1), u[:, 1].clamp(0, ny - 1)), 1)
        cell = torch.stack((uc[:, 0].floor().clamp(0, nx - 2), uc[:, 1].floor().clamp(0, ny - 2)), 1)
        f = uc - cell
        ci = cell.long()
        w0 = (1 - f[:, 0], f[:, 0])
        w1 = (1 - f[:, 1], f[:, 1])
        blend = 0
        for a in (0, 1):
            for b in (0, 1):
                blend = blend + (w0[a] * w1[b])[None] * jets[:, ci[:, 0] + a, ci[:, 1] + b]
        d, gx, gy, lap = blend
        d = d + ((u - uc) * self.spacing).norm(dim=1)
        gn = torch.sqrt(gx * gx + gy * gy).clamp(min=1e-300)
        return d, torch.stack((gx, gy), 1) / gn[:, None], lap / (1.0 - d * lap), gn

    def indicator(self, p, supports=None):
        d = self.signed(p)[0]
        return torch.where(d < 0, 1.0, torch.where(d > 0, 0.0, 0.5)).to(F64)


# the four corners of a grid cell counter-clockwise, as offsets from its lower-left node
_CELL_CORNERS = np.array([(0, 0), (1, 0), (1, 1), (0, 1)])


def _contour_of(d, origin, spacing):
    """the SurfaceRep of the zero level set of node values d [nx, ny] (solid = d < 0), vectorised over the cells.

    A cell with both kinds of corner has a crossing on every side whose two corners differ (the linear zero of the
    node values along it); each side of each cell is one vertex, numbered cell by cell (i outer, j inner) and side by
    side counter-clockwise: the two cells that share a grid edge number its crossing once each, as the reference's
    marching squares does (scene/scene.py:762-806), so the loops close by position.  A side that leaves the solid
    (its first corner solid) is joined to the next side that enters it: the neighbouring crossing counter-clockwise,
    or clockwise in a saddle cell whose four nodes average below zero.  The border nodes must be all fluid (an enclosed
    solid, background 0) or all solid (a tank, background 1)."""
    d = np.asarray(d, dtype=np.float64)
    nx, ny = d.shape
    solid = d < 0
    ring = np.concatenate((solid[0], solid[-1], solid[1:-1, 0], solid[1:-1, -1]))
    if ring.any() and not ring.all():
        raise ValueError("SdfRep.contour: the border nodes of the grid are partly solid; sample the body with a margin "
                         "of fluid (or, for a tank, of solid) all round")
    background = int(ring.all())
    ci, cj = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    ci, cj = ci.ravel(), cj.ravel()
    corner = [solid[ci + di, cj + dj] for di, dj in _CELL_CORNERS]             # 4 x [cells]
    cuts = np.stack([corner[k] != corner[(k + 1) % 4] for k in range(4)], 1)  # [cells, 4]
    first = np.cumsum(cuts.ravel()) - 1                                       # vertex number of each (cell, side)
    vid = np.where(cuts.ravel(), first, -1).reshape(-1, 4)
    cell_k, side_k = np.nonzero(cuts)
    a = _CELL_CORNERS[side_k] + np.stack((ci[cell_k], cj[cell_k]), 1)
    b = _CELL_CORNERS[(side_k + 1) % 4] + np.stack((ci[cell_k], cj[cell_k]), 1)
    da, db = d[a[:, 0], a[:, 1]], d[b[:, 0], b[:, 1]]
    t = da / (da - db)
    verts = np.stack([origin[ax] + spacing * (a[:, ax] + t * (b[:, ax] - a[:, ax])) for ax in (0, 1)], 1)
    # joins: every side leaving the solid, in cell order and counter-clockwise within a cell
    leave_c, leave_k = np.nonzero(cuts & np.stack(corner, 1))
    saddle = cuts.sum(1) == 4
    mean_solid = (d[ci, cj] + d[ci + 1, cj] + d[ci + 1, cj + 1] + d[ci, cj + 1]) / 4 < 0
    step = np.where(saddle[leave_c] & mean_solid[leave_c], -1, 1)
    target = (leave_k + step) % 4
    two = ~saddle[leave_c]
    # two crossings: the other one, wherever it sits on the cell
    other = np.argmax(cuts[leave_c] & (np.arange(4)[None] != leave_k[:, None]), axis=1)
    target = np.where(two, other, target)
    edges = np.stack((vid[leave_c, leave_k], vid[leave_c, target]), 1)
    if len(edges) == 0:
        raise ValueError("SdfRep.contour: the samples have no zero level set")
    return SurfaceRep(verts, edges.astype(np.int32), background=background)


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

    def tiled(self, periodic: Periodic, supportMax):
        """this bundle with its periodic images (reference scene/scene.py:834-850): every disk at every box shift k,
        |k| <= 2 per periodic axis, whose image reaches the window [-L/2 - H, L/2 + H] around the body-frame origin
        (H = ``supportMax``).  A query min-imaged into the box then meets each disk it can see once (support < L / 2).
        The images are listed shift by shift, (kx, ky) in lexicographic order, each the bundle's disks in order; the
        body frame must be the world frame (no rotation) for the offsets to be box lengths."""
        L = periodic.length(self.centres)
        Lh = L.cpu().tolist()
        half = 0.5 * L
        keep_c, keep_r = [], []
        for kx in range(-2, 3):
            for ky in range(-2, 3):
                if (kx and Lh[0] == 0.0) or (ky and Lh[1] == 0.0):
                    continue
                c = self.centres + torch.stack([kx * L[0], ky * L[1]])
                over = c.abs() - half - float(supportMax) - self.radii[:, None]
                keep = ((over <= 0) | (L == 0)).all(dim=1)
                keep_c.append(c[keep])
                keep_r.append(self.radii[keep])
        out = DiskArrayRep(torch.cat(keep_c), torch.cat(keep_r)).to(self.centres.device)
        out._tileOf = (len(self.radii), len(out.radii))
        return out

    def signed(self, pos):
        """(d, n): the smallest |x - c| - R over the disks (positive in the fluid; the first disk on a tie) and the
        unit vector from that disk's centre, body frame (reference scene/scene.py:853-867).  The points go in blocks
        sized so that a block's distance table holds about 4 M entries; no host synchronisation."""
        cen = self.centres.to(pos.device)
        rad = self.radii.to(pos.device)
        p = pos.to(F64)
        block = max(1, (1 << 22) // max(len(rad), 1))
        d_out, n_out = [], []
        for s in range(0, len(p), block):
            q = p[s:s + block]
            gap = torch.linalg.vector_norm(q[:, None, :] - cen[None], dim=2) - rad[None]
            d, k = gap.min(dim=1)
            v = q - cen[k]
            d_out.append(d)
            n_out.append(v / torch.linalg.vector_norm(v, dim=1, keepdim=True).clamp(min=1e-300))
        if not d_out:
            return torch.zeros(0, dtype=F64, device=pos.device), torch.zeros((0, 2), dtype=F64, device=pos.device)
        return torch.cat(d_out), torch.cat(n_out)

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
    periodic: Optional[Periodic] = None     # set by Scene.setPeriodic: particles are seen at their nearest image
    periodicSupport: float = 0.0            # the largest kernel support of the particles (the reach of the images)

    def __post_init__(self):
        self.center = torch.as_tensor(self.center, dtype=F64)
        self.linearVelocity = torch.as_tensor(self.linearVelocity, dtype=F64)

    @property
    def pose(self):
        return Pose(self.center, float(self.angle))

    def image(self, world):
        """the world points at their nearest periodic image of the body centre, c + min_image(x - c) (the points
        themselves without a periodic box; the positions are never written)."""
        if self.periodic is None:
            return world
        return self.center + min_image(world - self.center, self.periodic)

    def relative(self, world):
        """x - c at the nearest image (the lever arms of the loads)."""
        return min_image(world - self.center, self.periodic)

    def toLocal(self, world):
        """body-frame coordinates of world points (nearest image of the centre in a periodic box)."""
        return self.pose.toLocal(self.image(world))

    def tiledRep(self, rep):
        """``rep`` with its periodic images for the reach ``periodicSupport`` (cached); ``rep`` itself without a box
        (reference scene/scene.py:998-1006)."""
        if self.periodic is None:
            return rep
        return self._tile(rep, self.periodicSupport)

    def _tile(self, rep, supportMax):
        """a disk bundle of more than one disk in a periodic box: its tiled images for the reach max(supportMax,
        periodicSupport); any other representation: checked to fit or span the box and returned as it is
        (reference scene/scene.py:1008-1018).  Images need a body that does not rotate."""
        if self.periodic is None:
            return rep
        H = max(float(supportMax), float(self.periodicSupport))
        if not isinstance(rep, DiskArrayRep) or len(rep.radii) < 2:
            self._checkCompact(rep, H)
            return rep
        cache = self.__dict__.setdefault("_tileCache", {})
        key = (id(rep), H, self.periodic)
        if key not in cache:
            t = rep.tiled(self.periodic, H)
            if len(t.radii) > len(rep.radii) and float(self.angle) != 0.0:
                raise NotImplementedError("a bundle with periodic images cannot rotate (its image offsets are fixed "
                                          "in the body frame)")
            cache[key] = t
        return cache[key]

    def _checkCompact(self, rep, supportMax):
        """on every periodic axis the representation fits its cell (extent from the centre + support below L / 2: the
        nearest image of the centre serves every row) or spans the window [-L/2 - H, L/2 + H] (a wall that is itself
        periodic, taken as translation invariant by L); anything else is a ValueError (reference
        scene/scene.py:1020-1031)."""
        lo, hi = rep.bounds()
        L = self.periodic.length(lo)
        H = float(supportMax)
        ext = torch.maximum(lo.abs(), hi.abs())
        compact = ext + H < 0.5 * L
        spanning = (lo <= -0.5 * L - H) & (hi >= 0.5 * L + H)
        if bool(((L > 0) & ~(compact | spanning)).any()):
            raise ValueError("a periodic body must fit its cell (extent + support below half the box length) or span "
                             "it (cover the box plus a support) on every periodic axis")

    def unbounded(self):
        """whether the solid of the body reaches infinity (its broadphase keeps every query): a surface loop with
        background 1, a tank box, a cavity disk, or a sampled field whose fallback surface has background 1 (the
        reference's convention, scene/scene.py:1039-1051: a field without a fallback counts as bounded)."""
        for r in self.reps:
            if isinstance(r, SurfaceRep):
                hit = bool(r.background)
            elif isinstance(r, SdfRep):
                hit = r.fallback is not None and bool(r.fallback.background)
            else:
                hit = getattr(r, "solid", "inside") == "outside"
            if hit:
                return True
        return False

    def rot4(self, device):
        """R as a row-major [4] float64 device tensor, built once per (angle, device): the fused evaluation of a body
        whose pose does not change makes no host-to-device copy (it can be captured in a graph)."""
        key = (float(self.angle), str(torch.device(device)))
        cache = self.__dict__.setdefault("_rot4", {})
        if key not in cache:
            cache.clear()
            cache[key] = self.pose.Rflat().to(device, F64).contiguous()
        return cache[key]

    def fusedReps(self, hmin, device):
        """the representations this body is evaluated with: a SurfaceRep or DiskArrayRep as it is, a BoxRep its
        polygon, a solid ImplicitRep disk a one-disk DiskArrayRep, a cavity its tier-2 polygon for the smallest support
        ``hmin`` (reference scene/scene.py:970-996); in a periodic box a bundle's tiled images, the others checked to
        fit or span the box."""
        out = []
        for r in self.reps:
            low = r if isinstance(r, SurfaceRep) else r.lowered(hmin, device)
            out.append(self._tile(low, hmin) if self.periodic is not None else low)
        return out

    def obb(self):
        los, his = zip(*[r.bounds() for r in self.reps])
        dev = self.center.device
        lo = torch.stack([a.to(dev) for a in los]).amin(0)
        hi = torch.stack([a.to(dev) for a in his]).amax(0)
        if self.periodic is not None and any(isinstance(r, DiskArrayRep) and len(r.radii) > 1 for r in self.reps):
            # a bundle sees its tiled images: every query of the box around the centre is a candidate
            L = self.periodic.length(lo)
            lo, hi = torch.where(L > 0, -0.5 * L, lo), torch.where(L > 0, 0.5 * L, hi)
        return lo, hi

    def velocityAt(self, world):
        w = float(self.angularVelocity)
        r = world - self.center
        return self.linearVelocity + w * torch.stack([-r[:, 1], r[:, 0]], 1)


@dataclass
class BodyAdjacency:
    """the queries of one body: ``cand`` their global rows (all N rows for the fixed-capacity adjacency, the
    broadphase candidates when ``compact``), their body-frame positions and supports, one topology per representation
    (``repList``), and ``valid`` the broadphase of the fixed-capacity rows."""
    body: object
    cand: torch.Tensor
    lpos: torch.Tensor
    lsup: torch.Tensor
    reps: list
    valid: Optional[torch.Tensor] = None
    repList: Optional[list] = None
    compact: bool = False


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
        self.periodic = None
        for b in bodies:
            b.center = b.center.to(device)
            b.linearVelocity = b.linearVelocity.to(device)
            for r in b.reps:
                r.to(device)

    def setPeriodic(self, periodic: Optional[Periodic], support: float = 0.0):
        """a periodic box for all bodies (reference scene/scene.py:1243-1250): every body sees the particles at the
        nearest image of its centre; ``support`` is the particles' largest support (the reach of a bundle's images)."""
        if periodic is not None:
            periodic.checkSupport(support)
        self.periodic = periodic
        for b in self.bodies:
            b.periodic, b.periodicSupport = periodic, float(support)
            b.__dict__.pop("_tileCache", None)

    def adjacency(self, queryParticles, operationProperties=None, supportMax: float = None):
        """`SceneAdjacency` over the broadphase candidates of every body (reference scene/scene.py:1400-1416): the
        fluid queries whose support circle meets the body box (every fluid query of an unbounded solid), compacted
        (this reads the candidate counts on the host), and per representation the fixed-capacity topology of the
        candidates.  The integrals over it equal those of ``fixed_adjacency`` row for row; only the fixed-capacity
        adjacency avoids the host synchronisation (and can be captured in a graph)."""
        dev = torch.device(self.device)
        pos = queryParticles.positions.to(dev, F64)
        sup = queryParticles.supports.to(dev, F64)
        if pos.shape[1] != 2:
            raise NotImplementedError("2D only")
        smax = float(sup.max()) if supportMax is None else float(supportMax)
        fluid = queryAllowed(queryParticles, dev)
        bodies, counts = [], []
        for b in self.bodies:
            full = _body_adjacency(b, pos, sup, fluid, torch.arange(len(pos), device=dev), smax, dev)
            cand = torch.nonzero(full.valid).flatten()
            lpos, lsup = full.lpos[cand].contiguous(), sup[cand].contiguous()
            topos = [_rep_topology(r, lpos, lsup, None, smax, dev) for r in full.repList]
            bodies.append(BodyAdjacency(b, cand, lpos, lsup, topos, None, full.repList, compact=True))
            counts.append(int(len(cand)))
        return SceneAdjacency(len(pos), bodies, {"candidates": counts}, queryParticles, smax)

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

    def _chosen(self, body):
        return [(bi, bd) for bi, bd in enumerate(self.bodies) if body is None or bi == body]

    def inside(self, points, body=None):
        """True where a world point lies in the solid of a body (of ``body`` only when given): the winding indicator
        above 1/2 for a surface loop, a negative signed distance for the other representations (reference
        scene/scene.py:1279-1310)."""
        pts = points.to(self.device, F64)
        flags = [_rep_inside(bd.tiledRep(rep), bd.toLocal(pts)) for _, bd in self._chosen(body) for rep in bd.reps]
        if not flags:
            return torch.zeros(len(pts), dtype=torch.bool, device=pts.device)
        return torch.stack(flags).any(dim=0)

    def signed_distance(self, points, body=None, supportMax=None, want_body=False):
        """(d [M], n [M,2], hit [M][, body [M]]): the smallest signed distance over the representations of the
        bodies (the first representation on a tie), positive in the fluid, with its unit normal into the fluid in the
        world frame and the index of its body (reference scene/scene.py:1312-1372; surface loops: the nearest edge,
        the sign from the winding number normalised by ``supportMax``)."""
        pts = points.to(self.device, F64)
        M = len(pts)
        dist, normals, owner = [], [], []
        for bi, bd in self._chosen(body):
            lp = bd.toLocal(pts)
            for rep in bd.reps:
                d, n_loc = _rep_signed(bd.tiledRep(rep), lp, supportMax)
                dist.append(torch.nan_to_num(d, nan=float("inf")))
                normals.append(bd.pose.vecToWorld(n_loc))
                owner.append(torch.full((M,), bi, dtype=torch.long, device=pts.device))
        if not dist:
            best = torch.full((M,), float("inf"), dtype=F64, device=pts.device)
            nrm = torch.zeros((M, 2), dtype=F64, device=pts.device)
            bidx = torch.full((M,), -1, dtype=torch.long, device=pts.device)
            hit = torch.zeros(M, dtype=torch.bool, device=pts.device)
        else:
            D = torch.stack(dist)
            k = D.argmin(dim=0)                                  # the first minimum: the earliest representation
            rows = torch.arange(M, device=pts.device)
            best = D[k, rows]
            nrm = torch.stack(normals)[k, rows]
            bidx = torch.stack(owner)[k, rows]
            hit = torch.ones(M, dtype=torch.bool, device=pts.device)
        if want_body:
            return best, nrm, hit, bidx
        return best, nrm, hit


def _rep_inside(rep, lp):
    if isinstance(rep, SurfaceRep):
        return rep.indicator(lp) > 0.5
    if isinstance(rep, BoxRep):
        return rep.surface().indicator(lp) > 0.5
    return rep.signed(lp)[0] < 0


def _rep_signed(rep, lp, supportMax):
    """(d, n) of one representation in its body frame."""
    if isinstance(rep, BoxRep):
        rep = rep.surface()
    if isinstance(rep, SurfaceRep):
        smax = float(supportMax) if supportMax is not None else _default_support(rep)
        d, n, _ = rep.geom(smax, lp.device).signed_distance(lp, smax)
        return d, n
    d, n = rep.signed(lp)[:2]
    return d, n


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
    """`SceneAdjacency` of the queries against every body with fixed shapes (reference scene/fixedadj.py:249-270): per
    body the queries in its frame, the broadphase (support circle against the body box; every fluid query of an
    unbounded solid) and one fixed-capacity adjacency launch per representation it is evaluated with (implicit and
    sampled shapes through their polygons at the resolution of ``supportMax``).  With ``supportMax`` given (a host
    float >= every support) no step synchronises with the host."""
    dev = torch.device(scene.device)
    pos = queryParticles.positions.to(dev, F64)
    sup = queryParticles.supports.to(dev, F64)
    if pos.shape[1] != 2:
        raise NotImplementedError("2D only")
    smax = float(sup.max()) if supportMax is None else float(supportMax)
    fluid = queryAllowed(queryParticles, dev)
    rows = torch.arange(len(pos), device=dev)
    bodies = [_body_adjacency(b, pos, sup, fluid, rows, smax, dev) for b in scene.bodies]
    return SceneAdjacency(len(pos), bodies, {"candidates": None}, queryParticles, smax)


def _body_adjacency(body, pos, sup, fluid, rows, smax, dev):
    lpos = body.toLocal(pos)
    if body.unbounded():
        valid = fluid
    else:
        lo, hi = body.obb()
        gap = torch.maximum(lo - lpos, lpos - hi).clamp(min=0)          # per axis, outside the box only
        valid = fluid & (gap.norm(dim=1) < sup)
    reps = body.fusedReps(smax, dev)
    return BodyAdjacency(body, rows, lpos, sup, [_rep_topology(r, lpos, sup, valid, smax, dev) for r in reps], valid,
                         reps)


def _rep_topology(rep, lpos, sup, valid, smax, dev):
    """the fixed-capacity slots (and the winding indicator of a surface loop) of one representation."""
    g = rep.geom(smax, dev)
    if isinstance(rep, DiskArrayRep):
        slots, counts = g.adjacency(lpos, sup, valid=valid)
        return FixedTopology(slots, counts, g.K, torch.zeros(len(lpos), dtype=F64, device=dev), g)
    slots, counts, ind = g.adjacency(lpos, sup, valid=valid, pad=True)
    return FixedTopology(slots, counts, g.K, ind, g)


def plane_vector(kernel: str, device) -> torch.Tensor:
    return torch.tensor(plane_channels(kernel), dtype=F64, device=device)
