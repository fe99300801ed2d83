"""MI355X-native wall-integral hot path of warpSPHBoundaries: exact SPH boundary integrals (twelve moments) of
closed polygonal walls at fluid particles, on hand-written HIP kernels for gfx950 behind a C ABI (include/wsb.h).

Public surface (the reference's names):
    AnalyticBoundary, BoundaryProvider                     provider protocol (scene/provider.py)
    Scene, Body, SurfaceRep, ParticleState, fixed_adjacency scene layer
    BoxRep, ImplicitRep, DiskBody, TierPolicy, DiskArrayRep, SdfRep  representations (boxes and sampled distance
                                                           fields as polygons, solid disks as the disk element)
    FusedWall, FusedGroup, WallOutput                      fused evaluation
    sceneOperation, BodyField, SceneReaction,              operations over PairMoments (Density, Interpolate,
    OperationProperties, WarpOperation, GradientScheme     Gradient, Divergence, Curl, Covariance; the reaction)
    edge_channels, second_channels, indicator_vector       pair-list evaluators
    value, gradient, moment, moment_gradient               polygon element API
    KERNELS                                                exact kernel table
"""
from .kernels import KERNELS, kernel
from .plan import CHANNEL_NAMES, REPORT_NAMES, REPORT_ORDER

__all__ = ["KERNELS", "kernel", "CHANNEL_NAMES", "REPORT_NAMES", "REPORT_ORDER"]


def __getattr__(name):
    # the device layers import torch and load the HIP library lazily
    if name in ("AnalyticBoundary", "BoundaryProvider"):
        from . import provider
        return getattr(provider, name)
    if name in ("Scene", "Body", "SurfaceRep", "ParticleState", "fixed_adjacency", "Pose", "BoxRep", "ImplicitRep", "DiskArrayRep", "SdfRep",
                "DiskBody", "TierPolicy"):
        from . import scene
        return getattr(scene, name)
    if name in ("sceneOperation", "BodyField", "SceneReaction", "OperationProperties", "WarpOperation", "GradientScheme"):
        from . import operation
        return getattr(operation, name)
    if name in ("FusedWall", "FusedGroup", "WallOutput"):
        from . import fused
        return getattr(fused, name)
    if name in ("edge_channels", "second_channels", "indicator_vector", "value", "gradient", "moment", "moment_gradient"):
        from . import edge
        return getattr(edge, name)
    raise AttributeError(name)
