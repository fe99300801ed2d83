"""The precompute format of the reference's oracle path (scene/scene.py:181-210, 1111-1119, 1418-1445): per body and
representation the moments of one kernel at every query that sees the representation, world frame, physical units,
for consumers that build them once and contract them many times.

    lam = int W,  m1 = int y W,  g0 = int grad_x W,  g1[d, j] = int y_d d_j W,  m2[d, e] = int y_d y_e W   (y = x' - x)

Each row holds one query's sum over the representation's edges (slot order) plus its indicator term, from the same
device closed form as the fused evaluation; the reference's rows are per (query, element) pair, which a consumer
adds per query (``index_add_`` over ``q``) to the same values.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch

from .scene import SceneAdjacency, kernelName

F64 = torch.float64


@dataclass
class RepMoments:
    q: torch.Tensor                       # [P] global query index
    lam: torch.Tensor                     # [P]
    g0: torch.Tensor                      # [P, 2]
    m1: Optional[torch.Tensor] = None     # [P, 2]
    g1: Optional[torch.Tensor] = None     # [P, 2, 2]
    m2: Optional[torch.Tensor] = None     # [P, 2, 2]
    quadrature_estimate: Optional[torch.Tensor] = None
    quadrature_status: Optional[torch.Tensor] = None


@dataclass
class PairMoments:
    numQueries: int
    kernel: str
    entries: list                         # per body: dict(body, cand, surface=[RepMoments], implicit=[], volume=[])
    stats: dict
    channels: Optional[frozenset] = None  # None: all twelve; else the evaluated channels (the others are 0)


def _kernel_of(props):
    return kernelName(getattr(props, "kernel", props))


def precompute(scene, adjacency: SceneAdjacency, operationProperties, channels=None, form=None) -> PairMoments:
    """``Scene.precompute``: the moments of ``operationProperties.kernel`` (or a kernel name) of every surface
    representation (boxes and disks through their polygons) at the queries with a pair or a nonzero indicator."""
    from .fused import rep_channels
    from . import precision
    name = _kernel_of(operationProperties)
    form = precision.default_form(form)
    dev = torch.device(scene.device)
    keep = None
    if channels is not None:
        keep = torch.zeros(12, dtype=F64, device=dev)
        keep[[int(c) for c in channels]] = 1.0
    entries, pairs = [], 0
    for ba in adjacency.bodies:
        ent = {"body": ba.body, "cand": ba.cand, "surface": [], "implicit": [], "volume": []}
        R = ba.body.rot4(dev)
        reps = ba.repList if ba.repList is not None else ba.body.reps
        for rep, topo in zip(reps, ba.reps):
            c = rep_channels(rep, topo, ba, name, form, R)
            if keep is not None:
                c = c * keep
            seen = (topo.counts > 0) | (topo.ind != 0)
            if ba.valid is not None:
                seen &= ba.valid
            rows = torch.nonzero(seen).flatten()
            pairs += int(topo.counts[rows].sum())
            c = c[rows]
            ent["surface"].append(RepMoments(
                q=ba.cand[rows], lam=c[:, 0].clone(), g0=c[:, 3:5].clone(), m1=c[:, 1:3].clone(),
                g1=c[:, 5:9].reshape(-1, 2, 2).clone(),
                m2=torch.stack([c[:, 9], c[:, 10], c[:, 10], c[:, 11]], 1).reshape(-1, 2, 2)))
        entries.append(ent)
    stats = {"candidates": [len(ba.cand) for ba in adjacency.bodies], "pairs": pairs, "tier": {}}
    return PairMoments(adjacency.numQueries, name, entries, stats,
                       None if channels is None else frozenset(int(c) for c in channels))


def pair_moments(scene, queryParticles, operationProperties, channels=None, form=None) -> PairMoments:
    """``Scene.pairMoments``: the broadphase adjacency (``Scene.adjacency``) + precompute in one call."""
    return precompute(scene, scene.adjacency(queryParticles, operationProperties), operationProperties, channels, form)


def moments(scene, source, queryParticles, operationProperties, channels=None, form=None) -> PairMoments:
    """``Scene.moments``: None builds adjacency + precompute, a SceneAdjacency is precomputed, a PairMoments is
    returned as it is."""
    if source is None:
        return pair_moments(scene, queryParticles, operationProperties, channels, form)
    if isinstance(source, SceneAdjacency):
        return precompute(scene, source, operationProperties, channels, form)
    if isinstance(source, PairMoments):
        return source
    raise TypeError("moments: source must be None, a SceneAdjacency or a PairMoments")
