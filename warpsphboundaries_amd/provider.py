"""Boundary provider over the analytic bodies of a scene (the protocol of the reference's ``scene/provider.py``).

A particle scheme hands the provider its particle state once per position set and reads back, per body and particle,
the wall integrals it needs; the physics of the boundary conditions stays in the scheme.

    aggregate(ps, support, kernel, laplacian, fixedAdjacency, lean)
        -> a FusedWall; ``.out`` maps  lam [B,N] (int W),  G [B,N,2] (int grad W),  Cov [B,N,2,2] (int y (x) grad W)
           and, unless ``lean``, the raw cover [B,N,2], lap [B,N] (with ``laplacian``) and tens [B,N,2] integrals
           (``wallops`` holds their factors); further outputs through ``evaluate``, ``cone_area``, ``dir_extreme``
    signed_distance(x, body, supportMax, want_body)
        -> (d, n, hit[, body]): distance to the nearest solid, positive in the fluid, normal pointing into the fluid
    bodies, supported(fixedAdjacency)
"""
from __future__ import annotations

from typing import Protocol, runtime_checkable

from . import precision
from .fused import FusedGroup, FusedWall, WallAggregate, WallOutput
from .scene import Scene, fixed_adjacency, kernelName

_FAMILY = {"w2": "w2", "w4": "w4"}


@runtime_checkable
class BoundaryProvider(Protocol):
    bodies: list

    def aggregate(self, ps, support, **kwargs): ...

    def signed_distance(self, x, body=None, supportMax=None, want_body=False): ...


class AnalyticBoundary:
    """serves a ``Scene`` through the ``BoundaryProvider`` protocol: every integral is the exact closed form of the
    bodies' edges, evaluated by the HIP library in the arithmetic ``form`` (the process precision when None)."""

    def __init__(self, scene: Scene, form: str = None):
        self.scene = scene
        self.form = precision.default_form(form)      # None: the process precision (warpSPHCore_PRECISION)
        self._layouts = {}

    @property
    def bodies(self):
        return self.scene.bodies

    def family(self, kernel):
        """the Wendland family ('w2' or 'w4') whose wall Laplacian and tensile kernels exist; other kernels raise
        NotImplementedError."""
        fam = _FAMILY.get(kernelName(kernel))
        if fam is None:
            raise NotImplementedError("exact wall operations: Wendland C2 and C4 only")
        return fam

    def supported(self, fixedAdjacency=True):
        return FusedWall.supported(self.scene, fixedAdjacency)

    def properties(self, kernel, operation=None):
        return {"kernel": kernelName(kernel), "operation": operation, "operationMode": "BoundaryToFluid"}

    def layout(self, kernel, laplacian, fixedAdjacency=True):
        """(groups, roles) of the full layout (reference ``scene/provider.py:73-89``): a tuple of FusedGroups and the
        group index of each role -- 'w' the scheme kernel, 'cone' the cover kernel's gradient, 'lap' the wall Laplacian
        kernel ``l<fam>`` (only with ``laplacian``), 'tens' the gradient of the tensile kernel ``<fam>p5`` -- cached per
        (kernel, laplacian, fixedAdjacency)."""
        key = (kernelName(kernel), bool(laplacian), bool(fixedAdjacency))
        if key not in self._layouts:
            fam = self.family(kernel)
            groups, idx = [FusedGroup(kernelName(kernel), None, self.form), FusedGroup("cone", (3, 4), self.form)], \
                {"w": 0, "cone": 1}
            if laplacian:
                idx["lap"] = len(groups)
                groups.append(FusedGroup("l" + fam, None, self.form))
            idx["tens"] = len(groups)
            groups.append(FusedGroup(fam + "p5", (3, 4), self.form))
            self._layouts[key] = (tuple(groups), idx)
        return self._layouts[key]

    def aggregate(self, ps, support, kernel="w2", laplacian=False, fixedAdjacency=True, lean=False):
        """a FusedWall at the positions of ``ps`` (``support``: a host float no smaller than any support of ``ps``).
        The adjacency is the fixed-capacity one (graph-capturable) or, with ``fixedAdjacency=False``, the broadphase
        candidates of ``Scene.adjacency``; both give the same integrals.  ``.out`` holds lam, G, Cov of the scheme
        kernel and, unless ``lean``, the raw cover, lap (2 lam - tr Cov of ``l<fam>``, with ``laplacian``) and tens
        integrals, whose factors ``wallops`` supplies."""
        if lean:
            groups, idx = (FusedGroup(kernelName(kernel), None, self.form),), {"w": 0}
        else:
            groups, idx = self.layout(kernel, laplacian, fixedAdjacency)
        adj = fixed_adjacency(self.scene, ps, None, support) if fixedAdjacency else self.scene.adjacency(ps, None, support)
        fw = FusedWall(self.scene, adj, groups)
        outs = [WallOutput("lam", 0, "lam"), WallOutput("G", 0, "g0"), WallOutput("Cov", 0, "cov")]
        if not lean:
            outs.append(WallOutput("cover", idx["cone"], "g0"))
            if "lap" in idx:
                outs.append(WallOutput("lap", idx["lap"], "lap"))
            outs.append(WallOutput("tens", idx["tens"], "g0"))
        fw.out = fw.evaluate(outs)
        fw.roles = idx
        return fw

    def signed_distance(self, x, body=None, supportMax=None, want_body=False):
        return self.scene.signed_distance(x, body=body, supportMax=supportMax, want_body=want_body)
