"""Fused wall evaluation at one position set (reference ``scene/fused.py`` and ``edge/warpfused.py``).

    fw = FusedWall(scene, adjacency, groups)       # stage 1: one tile-kernel launch per (group, body, representation)
    out = fw.evaluate(outputs, a1=None)            # stage 2: {name: float64 tensor [B, N, ...]}, world frame, physical units

Stage 1 evaluates all twelve channels of every row with the row sums, the indicator x plane vector, the physical
scaling and the world rotation fused into the kernel (``WSB_OUT_PHYSICAL | WSB_OUT_WORLD``); several
representations of one body accumulate into the same rows.  A conditioned triangle (a closed 3-vertex loop) replaces
its closed form by the adaptive quadrature of ``triangle.py`` in the rows where it is small or a sliver at their
support.  Stage 2 only slices:

    "lam"  [B, N]        int_solid W                  "g0"   [B, N, 2]     int_solid grad_x W
    "cov"  [B, N, 2, 2]  int_solid y (x) grad_x W      "m1"   [B, N, 2]     int_solid y W
    "m2"   [B, N, 2, 2]  int_solid y (x) y W           "lap"  [B, N]        2 lam - tr cov
    "a1g1" [B, N, 2]     sum_d a1_d cov_dj for a per-query world vector a1
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import torch

from . import _lib as L
from . import precision, triangle
from .kernels import P5_NAMES
from .scene import BoxRep, DiskArrayRep, FixedTopology, ImplicitRep, SdfRep, SurfaceRep, kernelName

F64 = torch.float64
KINDS = {"lam": (0, 1), "g0": (1, 2), "cov": (2, 4), "a1g1": (3, 2), "lap": (4, 1), "m1": (5, 2), "m2": (6, 4)}


@dataclass(frozen=True)
class WallOutput:
    name: str
    group: int
    kind: str


@dataclass(frozen=True)
class FusedGroup:
    """one kernel of the fused evaluation; ``channels`` None = all twelve; ``form`` the arithmetic form
    ('fp64' closed form, 'mixed', or 'plain' = the plan interpreter)."""
    kernel: str
    channels: Optional[tuple] = None
    form: Optional[str] = None          # None: the process precision (warpSPHCore_PRECISION)

    def resolved(self):
        ch = None if self.channels is None else tuple(sorted(set(int(c) for c in self.channels)))
        name = kernelName(self.kernel)
        if name in P5_NAMES and (ch is None or not set(ch) <= {3, 4}):
            raise NotImplementedError(f"{name}: the W^5 kernels give the gradient channels (3, 4) only")
        return FusedGroup(name, ch, precision.default_form(self.form))


class WallAggregate:
    """marker base of the wall aggregates a scheme consumes (``.out``, ``evaluate``)."""


def rep_channels(rep, topo, ba, kernel, form, R, out=None, accumulate=False):
    """[N, 12] world-frame physical channels (device order) of one representation of a body over its fixed-capacity
    adjacency: for a surface loop the closed form, or for a conditioned triangle its quadrature on the rows where the
    triangle is small or a sliver at their support (reference scene/fused.py:556-564); for solid disks the disk
    element of every slot (FP64 whatever the form; reference scene/fused.py:567-575); added to ``out`` when
    ``accumulate``."""
    if isinstance(rep, DiskArrayRep) and isinstance(topo, FixedTopology):
        flags = L.OUT_PHYSICAL | L.OUT_WORLD | (L.OUT_ACCUMULATE if accumulate else 0)
        return topo.geom.moments(kernel, topo.e, topo.counts, ba.lpos, ba.lsup, rot4=R, flags=flags, out=out)
    if not isinstance(rep, SurfaceRep) or not isinstance(topo, FixedTopology):
        raise NotImplementedError("FusedWall: surface representations with the fixed-capacity adjacency only")
    if rep._triangle is None:
        flags = L.OUT_PHYSICAL | L.OUT_WORLD | (L.OUT_ACCUMULATE if accumulate else 0)
        return topo.geom.moments(kernel, topo.e, topo.counts, topo.ind, ba.lpos, ba.lsup, form=form, rot4=R,
                                 flags=flags, out=out)
    flags = L.OUT_PHYSICAL | L.OUT_WORLD
    own = topo.geom.moments(kernel, topo.e, topo.counts, topo.ind, ba.lpos, ba.lsup, form=form, rot4=R, flags=flags)
    triangle.override(rep.vertices, rep._triangle, kernel, rep.background, ba.lpos, ba.lsup, valid=ba.valid, rot4=R,
                      flags=flags, out=own)
    if out is None:
        return own
    if accumulate:
        out += own
    else:
        out.copy_(own)
    return out


class FusedWall(WallAggregate):
    def __init__(self, scene, adjacency, groups):
        self.scene = scene
        self.adj = adjacency
        self.groups = tuple((g if isinstance(g, FusedGroup) else FusedGroup(kernelName(g))).resolved() for g in groups)
        self.N = adjacency.numQueries
        dev = torch.device(scene.device)
        self.dev = dev
        B = len(adjacency.bodies)
        # per group: [B, N, 12] world-frame physical channels (device order)
        self.channels = torch.zeros((len(self.groups), B, self.N, 12), dtype=F64, device=dev)
        self._scratch = None
        for gi, grp in enumerate(self.groups):
            for bi, ba in enumerate(adjacency.bodies):
                R = ba.body.rot4(dev)
                reps = ba.repList if ba.repList is not None else ba.body.reps
                # the rows of a compacted adjacency are its candidates: evaluated apart, then put at their places
                dst = torch.zeros((len(ba.cand), 12), dtype=F64, device=dev) if ba.compact else self.channels[gi, bi]
                for rep, topo in zip(reps, ba.reps):
                    rep_channels(rep, topo, ba, grp.kernel, grp.form, R, dst, accumulate=True)
                if ba.compact:
                    self.channels[gi, bi].index_copy_(0, ba.cand, dst)
            if grp.channels is not None:
                keep = torch.zeros(12, dtype=F64, device=dev)
                keep[list(grp.channels)] = 1.0
                self.channels[gi] *= keep

    def evaluate(self, outputs, a1=None):
        """{name: tensor [B, N, ...]} float64, world frame, physical units."""
        res = {}
        for o in outputs:
            c = self.channels[o.group]
            B, N = c.shape[0], c.shape[1]
            if o.kind == "lam":
                res[o.name] = c[..., 0].clone()
            elif o.kind == "g0":
                res[o.name] = c[..., 3:5].clone()
            elif o.kind == "m1":
                res[o.name] = c[..., 1:3].clone()
            elif o.kind == "cov":
                res[o.name] = c[..., 5:9].reshape(B, N, 2, 2).clone()
            elif o.kind == "m2":
                res[o.name] = torch.stack([c[..., 9], c[..., 10], c[..., 10], c[..., 11]], -1).reshape(B, N, 2, 2)
            elif o.kind == "lap":
                res[o.name] = 2.0 * c[..., 0] - c[..., 5] - c[..., 8]
            elif o.kind == "a1g1":
                if a1 is None:
                    raise ValueError("FusedWall.evaluate: a1 [B, N, 2] is required for 'a1g1' outputs")
                cov = c[..., 5:9].reshape(B, N, 2, 2)
                res[o.name] = torch.einsum("bnd,bndj->bnj", a1.to(F64), cov)
            else:
                raise KeyError(f"unknown output kind {o.kind!r}; expected one of {sorted(KINDS)}")
        return res

    def copy_from(self, other):
        """overwrite this wall's channels with those of ``other`` (same scene, groups and query count) in place."""
        if other.channels.shape != self.channels.shape:
            raise ValueError("FusedWall.copy_from: different structure")
        self.channels.copy_(other.channels)

    def reset_derived(self):
        """nothing is derived lazily: the channels are complete after stage 1."""
        return None

    # ---- geometric queries of the solid within one support (reference scene/fused.py:741-835), FP64
    def _reps(self):
        for bi, ba in enumerate(self.adj.bodies):
            c, s = ba.body.pose._cs()
            reps = ba.repList if ba.repList is not None else ba.body.reps
            for rep, topo in zip(reps, ba.reps):
                yield bi, ba, rep, topo, c, s

    def cone_area(self, axes, half_angle):
        """[2, N] float64: area(solid cap disk(x, h) cap wedge(x, axis, half_angle)) (row 0, units length^2;
        ``half_angle >= pi``: the full disk) and the full-disk area (row 1), for the world ``axes`` [N, 2] (only the
        direction matters), summed over the bodies; h = the query's support."""
        N, dev = self.N, self.dev
        out = torch.zeros(2 * N, dtype=F64, device=dev)
        ax = axes.to(dev, F64).contiguous()
        lib = L.load()
        for bi, ba, rep, topo, c, s in self._reps():
            n = len(ba.cand)
            part = torch.zeros(2 * n, dtype=F64, device=dev) if ba.compact else out
            axp = ax[ba.cand].contiguous() if ba.compact else ax
            if isinstance(rep, DiskArrayRep):
                L.check(lib.wsb_disk_cone_area2d(topo.geom.handle, L.ptr(topo.e), L.ptr(topo.counts), topo.K,
                                                 L.ptr(ba.lpos), L.ptr(ba.lsup), n, L.ptr(axp), c, s, float(half_angle),
                                                 L.ptr(part), L.stream_ptr(dev)))
            else:
                L.check(lib.wsb_cone_area2d(topo.geom.handle, L.ptr(topo.e), L.ptr(topo.counts), topo.K,
                                            L.ptr(topo.ind), L.ptr(ba.lpos), L.ptr(ba.lsup), n, L.ptr(axp), c, s,
                                            float(half_angle), L.ptr(part), L.stream_ptr(dev)))
            if ba.compact:
                out[:N].index_add_(0, ba.cand, part[:n])
                out[N:].index_add_(0, ba.cand, part[n:])
            excluded = self._excluded(ba)
            if getattr(rep, "background", 0) and excluded is not None:
                # a query the adjacency left out of a background representation is deep in its solid: the full sector
                out_of = excluded.to(F64)
                h2 = self._supports() ** 2
                wedge = 2.0 * float(half_angle) if float(half_angle) < math.pi else 2.0 * math.pi
                out[:N] += out_of * 0.5 * h2 * wedge
                out[N:] += out_of * math.pi * h2
        return out.reshape(2, N)

    def _supports(self):
        return self.adj.queryParticles.supports.to(self.dev, F64)

    def _excluded(self, ba):
        """bool [N]: the queries a body's adjacency left out (None: none)."""
        if ba.compact:
            kept = torch.zeros(self.N, dtype=torch.bool, device=self.dev)
            kept[ba.cand] = True
            return ~kept
        return None if ba.valid is None else ~ba.valid

    def dir_extreme(self, vecs):
        """[B, N] float64: for the world vectors ``vecs`` [B, N, 2] the largest |cos| between the vector and a
        direction from the query to a solid point of that body within one support (1 where the vector's line meets
        the solid, 0 where there is no solid in support or the vector vanishes)."""
        N, dev = self.N, self.dev
        B = len(self.adj.bodies)
        out = torch.zeros(B * N, dtype=F64, device=dev)
        v = vecs.to(dev, F64).contiguous()
        lib = L.load()
        for bi, ba, rep, topo, c, s in self._reps():
            seg = out[bi * N:(bi + 1) * N]
            n = len(ba.cand)
            part = torch.zeros(n, dtype=F64, device=dev) if ba.compact else seg
            vp = v[bi][ba.cand].contiguous() if ba.compact else v[bi]
            if isinstance(rep, DiskArrayRep):
                L.check(lib.wsb_disk_dir_extreme2d(topo.geom.handle, L.ptr(topo.e), L.ptr(topo.counts), topo.K,
                                                   L.ptr(ba.lpos), L.ptr(ba.lsup), n, L.ptr(vp), c, s, L.ptr(part),
                                                   L.stream_ptr(dev)))
            else:
                L.check(lib.wsb_dir_extreme2d(topo.geom.handle, L.ptr(topo.e), L.ptr(topo.counts), topo.K,
                                              L.ptr(topo.ind), L.ptr(ba.lpos), L.ptr(ba.lsup), n, L.ptr(vp), c, s,
                                              L.ptr(part), L.stream_ptr(dev)))
            if ba.compact:
                seg.index_copy_(0, ba.cand, torch.maximum(seg[ba.cand], part))
            excluded = self._excluded(ba)
            if getattr(rep, "background", 0) and excluded is not None:
                seg[excluded & (v[bi].norm(dim=1) > 0)] = 1.0
        return out.reshape(B, N)

    @staticmethod
    def supported(scene, lowering=True):
        """every representation evaluates exactly: surface loops and disk arrays, and (with ``lowering``) boxes and
        sampled distance fields through their polygons and disks through the disk element (a cavity through its
        polygon)."""
        kinds = (SurfaceRep, DiskArrayRep, BoxRep, ImplicitRep, SdfRep) if lowering else (SurfaceRep, DiskArrayRep)
        return all(isinstance(r, kinds) for b in scene.bodies for r in b.reps)
