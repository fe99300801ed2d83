"""Boundary operations evaluated from PairMoments: the consumer of ``Scene.precompute`` (reference
``scene/scene.py:1640-1806``: ``_apply``, ``SceneReaction``, ``sceneOperation``; ``BodyField`` :894-910).

A body carries a boundary field A(x') = a0 + a1 (x' - c) (c its centre; scalar a0 [] with a1 [2], or C components
a0 [C] with a1 [C, 2]), or, with ``perQuery``, the value a0[i] of the field AT each query and its gradient a1.  With
the per-query moments of one kernel (lam = int W, m1 = int y W, g0 = int grad_x W, g1[d, j] = int y_d d_j W, world
frame, y = x' - x) every operation is linear in the moments:

    Density       rho lam
    Covariance    rho g1
    Interpolate   A(x) lam + a1 m1
    Gradient      (s A(x) + f) (x) g0 + s a1 g1      (s, f from the gradient scheme and the query values)
    Divergence    trace of the Gradient of a vector field;  Curl  its antisymmetric part out[1, 0] - out[0, 1]

The reaction of a scalar Gradient on the body is -sum_i m_i out_i, its torque about c from g1's antisymmetric part
and the lever r = x - c (exact for a constant field).  Everything here is a handful of batched tensor operations
on the device per representation; the moments themselves come from the HIP library.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import List, Optional

import torch

from .scene import kernelName

F64 = torch.float64


class WarpOperation(Enum):
    Density = "Density"
    Interpolate = "Interpolate"
    Gradient = "Gradient"
    Divergence = "Divergence"
    Curl = "Curl"
    Covariance = "Covariance"


class GradientScheme(Enum):
    Naive = "Naive"
    Symmetric = "Symmetric"
    Difference = "Difference"
    Summation = "Summation"


@dataclass
class OperationProperties:
    """the fields sceneOperation reads of warpSPHCore's OperationProperties (any object with ``kernel``,
    ``operation`` and ``gradientMode`` whose enums carry these names serves as well)."""
    kernel: object
    operation: object
    gradientMode: object = GradientScheme.Naive


def _name(x) -> str:
    return x if isinstance(x, str) else getattr(x, "name", str(x))


@dataclass
class BodyField:
    """boundary data of one body in world components (reference scene/scene.py:894-910): A(x') = a0 + a1 (x' - c);
    ``rho`` the boundary density (Density, the Symmetric scheme); ``perQuery``: a0 holds the field at each query
    (leading dimension N) and a1 its gradient (one for all queries or one per query)."""
    a0: Optional[torch.Tensor] = None
    a1: Optional[torch.Tensor] = None
    rho: float = 1.0
    perQuery: bool = False

    @staticmethod
    def rigid(body, rho: float = 1.0):
        """the velocity v(x') = v_c + omega x (x' - c) of a rigidly moving body: a0 = v_c, a1 = [[0, -w], [w, 0]]."""
        w = float(body.angularVelocity)
        dev = body.center.device
        return BodyField(torch.as_tensor(body.linearVelocity, dtype=F64, device=dev),
                         torch.tensor([[0.0, -w], [w, 0.0]], dtype=F64, device=dev), rho)


@dataclass
class SceneReaction:
    force: torch.Tensor                 # [B, 2] force of the fluid on each body (Gradient of a scalar field)
    torque: torch.Tensor                # [B] about the body centre; NaN where the moments do not determine it
    torqueExact: list                   # per body: every contribution had first moments and the field was constant


class _Field:
    """one body's field on the device: ``at(q, x)`` the value at the queries [P, C] and ``grad(q)`` its gradient
    [P, C, 2] (None for a constant field); ``scalar`` whether the caller passed a scalar field."""

    def __init__(self, fld: BodyField, center, dev):
        self.constant = fld.a1 is None
        self.per_query = bool(fld.perQuery)
        self.center = center.to(dev, F64)
        a0 = torch.as_tensor(fld.a0, dtype=F64, device=dev)
        a1 = None if fld.a1 is None else torch.as_tensor(fld.a1, dtype=F64, device=dev)
        if self.per_query:
            self.scalar = a0.ndim == 1
            self.a0 = a0.reshape(a0.shape[0], -1)                               # [N, C]
            C = self.a0.shape[1]
            if a1 is not None:
                if self.scalar:
                    a1 = a1.reshape(1, 1, 2) if a1.ndim == 1 else a1.reshape(-1, 1, 2)
                elif a1.ndim == 2:
                    a1 = a1.reshape(1, C, 2)
            self.a1 = a1                                                        # [1 or N, C, 2]
        else:
            self.scalar = a0.ndim == 0
            self.a0 = a0.reshape(-1)                                            # [C]
            self.a1 = None if a1 is None else a1.reshape(self.a0.shape[0], 2)   # [C, 2]
        self.C = self.a0.shape[-1]

    def at(self, q, x):
        if self.per_query:
            return self.a0[q]
        A = self.a0[None].expand(len(q), -1)
        if self.a1 is not None:
            A = A + (x - self.center) @ self.a1.T
        return A

    def grad(self, q):
        if self.a1 is None:
            return None
        if self.per_query:
            return self.a1.expand(len(q), -1, -1) if self.a1.shape[0] == 1 else self.a1[q]
        return self.a1[None].expand(len(q), -1, -1)


def _contribution(op, scheme, r, body, fld: Optional[_Field], rho, pos, rho_q, qv, masses, dev):
    """(rows [P, ...], force [2] or None, torque or None) of one representation's moment rows ``r`` (RepMoments)."""
    q = r.q
    if op == "Density":
        return rho * r.lam, None, None
    if op == "Covariance":
        if r.g1 is None:
            raise NotImplementedError("Covariance needs the first moments of the representation")
        return rho * r.g1, None, None
    if fld is None:
        raise ValueError("BodyField with a0 is required for this operation")
    x = pos[q]
    A = fld.at(q, x)                                                            # [P, C]
    G = fld.grad(q)                                                             # [P, C, 2] or None
    if G is not None and r.g1 is None:
        raise NotImplementedError("a position-dependent boundary field needs the first moments of the representation")
    if op == "Interpolate":
        val = A * r.lam[:, None]
        if G is not None:
            val = val + torch.bmm(G, r.m1[:, :, None].contiguous())[:, :, 0]
        return (val[:, 0] if fld.scalar else val), None, None
    P = len(q)
    if scheme == "Naive":
        s, f = torch.ones(P, dtype=F64, device=dev), None
    else:
        if qv is None:
            raise ValueError("queryValues required for this gradient mode")
        f = qv[q].reshape(P, -1)
        if scheme == "Symmetric":
            s = rho_q[q] / rho
            f = (rho / rho_q[q])[:, None] * f
        elif scheme == "Difference":
            s, f = torch.ones(P, dtype=F64, device=dev), -f
        elif scheme == "Summation":
            s = torch.ones(P, dtype=F64, device=dev)
        else:
            raise NotImplementedError(f"gradient scheme {scheme}")
    A0 = s[:, None] * A if f is None else s[:, None] * A + f                    # the effective field at the query
    out = A0[:, :, None] * r.g0[:, None, :]                                     # [P, C, 2]
    if G is not None:
        out = out + s[:, None, None] * torch.bmm(G.contiguous(), r.g1.contiguous())
    force = torque = None
    if fld.scalar and masses is not None:
        m = masses[q]
        force = -(m[:, None] * out[:, 0, :]).sum(0)
        if G is None and r.g1 is not None:
            arm = x - body.center.to(dev, F64)
            lever = (r.g1[:, 0, 1] - r.g1[:, 1, 0]) + (arm[:, 0] * r.g0[:, 1] - arm[:, 1] * r.g0[:, 0])
            torque = -(m * A0[:, 0] * lever).sum()
    if op == "Gradient":
        return (out[:, 0, :] if fld.scalar else out), force, torque
    if op == "Divergence":
        return out[:, 0, 0] + out[:, 1, 1], force, torque
    if op == "Curl":
        return out[:, 1, 0] - out[:, 0, 1], force, torque
    raise NotImplementedError(f"{op}: not an operation of the scene boundaries")


def _empty(op, fields, N, dev):
    C, scalar = 1, True
    for f in fields:
        if f is not None:
            C, scalar = f.C, f.scalar
    if op == "Covariance":
        return torch.zeros((N, 2, 2), dtype=F64, device=dev)
    if op in ("Density", "Divergence", "Curl"):
        return torch.zeros(N, dtype=F64, device=dev)
    if op == "Gradient":
        return torch.zeros((N, 2) if scalar else (N, C, 2), dtype=F64, device=dev)
    return torch.zeros(N if scalar else (N, C), dtype=F64, device=dev)


def sceneOperation(queryParticles, operationProperties, scene, moments=None, queryValues: Optional[torch.Tensor] = None,
                   bodyFields: Optional[List[BodyField]] = None, returnReaction: bool = False, perBody: bool = False,
                   form: Optional[str] = None):
    """The boundary contribution of ``operationProperties.operation`` at every query particle, summed over the bodies
    and their representations (``perBody``: [B, N, ...], one slice per body), and with ``returnReaction`` the
    ``SceneReaction`` of a scalar Gradient.  ``moments``: a PairMoments (shared by several operations at one position
    set), a SceneAdjacency (precomputed for this call) or None (both built here; ``form`` the arithmetic form of
    that build).  ``bodyFields[b]`` is the field of ``scene.bodies[b]`` (default: a unit-density wall, no field)."""
    dev = torch.device(scene.device)
    pm = scene.moments(moments, queryParticles, operationProperties, form=form)
    want = kernelName(operationProperties.kernel)
    if pm.kernel != want:
        raise ValueError(f"sceneOperation: the pair moments are those of kernel {pm.kernel!r}, the operation asks for "
                         f"{want!r} (a PairMoments holds the integrals of its own kernel)")
    op, scheme = _name(operationProperties.operation), _name(getattr(operationProperties, "gradientMode", "Naive"))
    nb = len(scene.bodies)
    bodyFields = list(bodyFields) if bodyFields is not None else [BodyField() for _ in range(nb)]
    if len(bodyFields) != nb:
        raise ValueError(f"sceneOperation: {len(bodyFields)} body fields for {nb} bodies")
    if pm.channels is not None:
        ok = (op == "Gradient" and scheme == "Naive" and not returnReaction and {3, 4} <= set(pm.channels)
              and not any(f.perQuery or f.a1 is not None for f in bodyFields))
        if not ok:
            raise ValueError(f"sceneOperation: the pair moments were built with the channels {sorted(pm.channels)} "
                             "only; they serve the Naive Gradient of a constant scalar field (no reaction)")
    N = pm.numQueries
    pos = queryParticles.positions.to(dev, F64)
    dens = getattr(queryParticles, "densities", None)
    rho_q = dens.to(dev, F64) if dens is not None else torch.ones(N, dtype=F64, device=dev)
    masses = getattr(queryParticles, "masses", None)
    masses = None if masses is None else masses.to(dev, F64)
    qv = None if queryValues is None else queryValues.to(dev, F64)
    fields = [None if f.a0 is None else _Field(f, b.center, dev) for f, b in zip(bodyFields, scene.bodies)]
    if returnReaction and op == "Gradient" and masses is None:
        raise ValueError("sceneOperation: the reaction needs the particle masses")
    force = torch.zeros((nb, 2), dtype=F64, device=dev)
    torque = torch.zeros(nb, dtype=F64, device=dev)
    exact = [True] * nb
    per = [None] * nb
    for ent in pm.entries:
        body = ent["body"]
        bidx = next(i for i, b in enumerate(scene.bodies) if b is body)
        for r in list(ent["surface"]) + list(ent["implicit"]):
            if r is None or len(r.q) == 0:
                continue
            rows, f, t = _contribution(op, scheme, r, body, fields[bidx], float(bodyFields[bidx].rho), pos, rho_q, qv,
                                       masses, dev)
            if per[bidx] is None:
                per[bidx] = torch.zeros((N,) + tuple(rows.shape[1:]), dtype=F64, device=dev)
            per[bidx].index_add_(0, r.q, rows)
            if f is not None:
                force[bidx] += f
                if t is None:
                    exact[bidx] = False
                else:
                    torque[bidx] += t
        if ent.get("volume"):
            raise NotImplementedError("volume representations are not part of the wall-integral path")
    shape_of = next((p for p in per if p is not None), None)
    if shape_of is None:
        shape_of = _empty(op, fields, N, dev)
    if perBody:
        return torch.stack([p if p is not None else torch.zeros_like(shape_of) for p in per])
    out = torch.zeros_like(shape_of)
    for p in per:
        if p is not None:
            out += p
    if returnReaction:
        tq = torque.clone()
        for bi in range(nb):
            if not exact[bi]:
                tq[bi] = float("nan")
        return out, SceneReaction(force, tq, exact)
    return out
