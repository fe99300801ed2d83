"""Exact wall operations of the full aggregate layout: the cover vector, the wall Laplacian and the tensile vector
(reference ``scene/cover.py:119-147``, ``scene/viscosity.py:52-110``, ``scene/tensile.py:31-90``), each one kernel
group of ``FusedWall`` times a constant:

* cover   C = -(pi H^3 / 3) grad_x int_solid W_cone dA'      W_cone = 3 (1 - q) / (pi H^2) (the kernel ``cone``)
                                                           (points away from the wall; units length^2)
* lap     Delta-lambda = f (2 lam - tr Cov)[L],  L = W'/r as the normalised kernel ``l<fam>``,
          f = P c / (C_l H^2), P = -20 (w2) / -56/3 (w4), c and C_l the normalisations (c2 * pi) of the family and
          of ``l<fam>``  (int_solid lap W dA'; positive on the fluid side, 0 beyond H and deep inside the solid)
* tens    T = (1/5) c^5 / (pi^4 c5 H^8) grad_x int_solid W5 dA',  W5 = W^5 normalised as ``<fam>p5`` (c5)
          (= int_solid W^4 grad W dA'; points into the wall)

``AnalyticBoundary.aggregate`` returns the raw kernel-group results (``cover``, ``lap``, ``tens`` in ``.out``); the
functions below apply the factors, with one uniform support ``H`` (the reference's scene routes).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .fused import FusedGroup, FusedWall, WallOutput
from .kernels import kernel as get_kernel
from .scene import ParticleState, fixed_adjacency

F64 = torch.float64
_PRE = {"w2": -20.0, "w4": -56.0 / 3.0}


def _family(family):
    if family not in ("w2", "w4"):
        raise NotImplementedError("exact wall operations: Wendland C2 and C4 only")
    return family


def cover_factor(H):
    """-(pi H^3 / 3): the raw cover (gradient of the ``cone`` group) to the cover vector (reference cover.py:147)."""
    return -(math.pi * float(H) ** 3 / 3.0)


def lap_factor(H, family="w2"):
    """f = P c / (C_l H^2): 2 lam - tr Cov of ``l<fam>`` to the wall Laplacian (reference viscosity.py:58-66)."""
    _family(family)
    c = float(get_kernel(family).c2_pi)
    cl = float(get_kernel("l" + family).c2_pi)
    return _PRE[family] * c / (cl * float(H) ** 2)


def tensile_factor(H, family="w2"):
    """(1/5) c^5 / (pi^4 c5 H^8): the raw tensile gradient of ``<fam>p5`` to T (reference tensile.py:43-50)."""
    _family(family)
    c = float(get_kernel(family).c2_pi)
    c5 = float(get_kernel(family + "p5").c2_pi)
    return 0.2 * c ** 5 / (math.pi ** 4 * c5 * float(H) ** 8)


def _wall(scene, positions, H, group, outputs):
    dev = torch.device(scene.device)
    pos = torch.as_tensor(np.asarray(positions) if not torch.is_tensor(positions) else positions, dtype=F64).to(dev)
    ps = ParticleState(pos, torch.full((len(pos),), float(H), dtype=F64, device=dev))
    return FusedWall(scene, fixed_adjacency(scene, ps, None, float(H)), (group,)).evaluate(outputs)


def cover_vector_scene(scene, positions, H):
    """the cover vector [N, 2] (summed over bodies; reference cover.py:119-147)."""
    g = _wall(scene, positions, H, FusedGroup("cone", (3, 4)), (WallOutput("g", 0, "g0"),))["g"]
    return cover_factor(H) * g.sum(0)


def lap_lambda_scene(scene, positions, H, family="w2"):
    """the wall Laplacian [B, N] per body (reference viscosity.py:69-110)."""
    lap = _wall(scene, positions, H, FusedGroup("l" + _family(family)), (WallOutput("l", 0, "lap"),))["l"]
    return lap_factor(H, family) * lap


def tensile_vector_scene(scene, positions, H, family="w2"):
    """the tensile vector T [N, 2] (summed over bodies; reference tensile.py:53-90)."""
    g = _wall(scene, positions, H, FusedGroup(_family(family) + "p5", (3, 4)), (WallOutput("g", 0, "g0"),))["g"]
    return tensile_factor(H, family) * g.sum(0)
