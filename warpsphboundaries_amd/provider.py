"""The analytic boundary as a PROVIDER for an SPH scheme (reference ``scene/provider.py``).

    aggregate(ps, support, ...)   the integrals of every body at the particle positions of ``ps``: a ``FusedWall``
                                  whose ``.out`` holds lam [B,N], G [B,N,2], Cov [B,N,2,2] (world frame, float64)
                                  and, in the full layout, cover [B,N,2], lap [B,N], tens [B,N,2] (raw, see
                                  ``wallops`` for the factors)
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
        """(groups, roles): the fused kernel groups the scheme needs and the group index of each role (reference
        ``scene/provider.py:73-89``): the scheme kernel 'w', the cover kernel 'cone' (gradient channels), the wall
        Laplacian 'lap' (``l<fam>``) when ``laplacian``, and the tensile kernel 'tens' (``<fam>p5``, gradient
        channels); built once per configuration."""
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
        """the fused wall evaluation at the positions of ``ps``; ``support``: host float >= every support of ``ps``.
        ``.out``: lam, G, Cov of the scheme kernel, and unless ``lean`` the raw cover (gradient of the cone kernel),
        lap (2 lam - tr Cov of ``l<fam>``, with ``laplacian``) and tens (gradient of ``<fam>p5``): the scheme applies
        the factors (``wallops.cover_factor``, ``lap_factor``, ``tensile_factor``)."""
        if not fixedAdjacency:
            raise NotImplementedError("only the fixed-capacity (graph-capturable) adjacency is built")
        if lean:
            groups, idx = (FusedGroup(kernelName(kernel), None, self.form),), {"w": 0}
        else:
            groups, idx = self.layout(kernel, laplacian, fixedAdjacency)
        adj = fixed_adjacency(self.scene, ps, None, support)
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
