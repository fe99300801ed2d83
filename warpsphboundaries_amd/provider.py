"""The analytic boundary as a PROVIDER for an SPH scheme (reference ``scene/provider.py``).

    aggregate(ps, support, ...)   the integrals of every body at the particle positions of ``ps``: a ``FusedWall``
                                  whose ``.out`` holds lam [B,N], G [B,N,2], Cov [B,N,2,2] (world frame, float64)
    signed_distance(x, ...)       (d, n, hit[, body]) of points to the nearest body: d > 0 in the fluid, n into the fluid
    bodies                        the rigid bodies of the scene
    supported(...)                whether the fused path serves the scene
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
    """the analytic bodies of a ``Scene`` as a boundary provider (exact edge-reduction integrals on the MI355X)."""

    def __init__(self, scene: Scene, form: str = None):
        self.scene = scene
        self.form = precision.default_form(form)      # None: the process precision (warpSPHCore_PRECISION)
        self._layouts = {}

    @property
    def bodies(self):
        return self.scene.bodies

    def family(self, kernel):
        """Wendland family of the exact wall operations (tensile, wall Laplacian): C2 and C4 only."""
        fam = _FAMILY.get(kernelName(kernel))
        if fam is None:
            raise NotImplementedError("exact wall operations: Wendland C2 and C4 only")
        return fam

    def supported(self, fixedAdjacency=True):
        return FusedWall.supported(self.scene, fixedAdjacency)

    def properties(self, kernel, operation=None):
        return {"kernel": kernelName(kernel), "operation": operation, "operationMode": "BoundaryToFluid"}

    def layout(self, kernel, laplacian, fixedAdjacency=True):
        """(groups, roles) of the full layout: the scheme kernel plus the cover ('cone'), wall-Laplacian ('l<fam>') and
        tensile ('<fam>p5') kernel groups of the reference; only the scheme kernel group is built here."""
        self.family(kernel)
        raise NotImplementedError("the cover / wall-Laplacian / tensile kernel groups (cone, l<fam>, <fam>p5) are not built: "
                                  "use aggregate(..., lean=True)")

    def aggregate(self, ps, support, kernel="w2", laplacian=False, fixedAdjacency=True, lean=False):
        """the fused wall evaluation at the positions of ``ps``; ``support``: host float >= every support of ``ps``."""
        if not lean:
            self.layout(kernel, laplacian, fixedAdjacency)
        if not fixedAdjacency:
            raise NotImplementedError("only the fixed-capacity (graph-capturable) adjacency is built")
        groups = (FusedGroup(kernelName(kernel), None, self.form),)
        adj = fixed_adjacency(self.scene, ps, None, support)
        fw = FusedWall(self.scene, adj, groups)
        fw.out = fw.evaluate((WallOutput("lam", 0, "lam"), WallOutput("G", 0, "g0"), WallOutput("Cov", 0, "cov")))
        return fw

    def signed_distance(self, x, body=None, supportMax=None, want_body=False):
        return self.scene.signed_distance(x, body=body, supportMax=supportMax, want_body=want_body)
