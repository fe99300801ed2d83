"""Conditioned small / sliver triangle route (reference ``edge/warptriangle.py``): for a closed 3-vertex loop, the rows
whose support makes the triangle small (longest edge <= 0.02 h) or a sliver (twice the area <= 0.01 longest edge
squared) take an adaptive area quadrature on the device (``wsb_triangle2d``) in place of the edge-local closed form,
whose edge terms are then far larger than the result.

    meta = prepare(vertices, edges)                       # static: (order, area2, sign) or None
    values, estimate, status = channels(vertices, positions, supports, meta, kernel, background)

``values`` / ``estimate`` [N, 12] (units of h, device channel order ``lam, m10, m01, Gx, Gy, g1xx, g1xy, g1yx, g1yy,
m20, m11, m02``), ``status`` [N] int32: 0 the closed-form route (values 0), 1 the quadrature converged, -1 its budget
ran out (values NaN).  Estimates are empirical (|16 - 8 outer nodes| + |32 - 16 inner nodes| per converged panel),
not certified bounds.  ``SurfaceRep`` applies the route to its rows by ``override``.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import torch

from . import _lib as L
from .kernels import KERNEL_ID, plane_channels as _plane_channels, plane_second_moment as _plane_second

F64 = torch.float64


def plane_channels(kernel):
    """the whole-support vector of the twelve channels (units of h; reference ``warptriangle.plane_channels``)."""
    return np.asarray(_plane_channels(kernel), dtype=np.float64)


def plane_second_moment(kernel):
    """int_plane y_x^2 W = int_plane y_y^2 W at h = 1 (reference ``warptriangle.plane_second_moment``)."""
    return float(_plane_second(kernel))


def prepare(vertices, edges):
    """static metadata of a closed 3-vertex loop, or None (reference ``edge/warptriangle.py:225-240``): the vertex
    order with the shortest edge first (its direction is the inner, transverse, one of the quadrature), |twice the
    signed area| and its sign (+1 counter-clockwise along the edges), the area in exact rational arithmetic of the
    float64 vertices (so sub-rounding determinants keep their sign)."""
    v = vertices.detach().cpu().numpy() if torch.is_tensor(vertices) else np.asarray(vertices, dtype=np.float64)
    e = edges.detach().cpu().tolist() if torch.is_tensor(edges) else np.asarray(edges).tolist()
    if len(v) != 3 or len(e) != 3:
        return None
    es = {tuple(int(x) for x in pair) for pair in e}
    if es == {(0, 1), (1, 2), (2, 0)}:
        orient = [0, 1, 2]
    elif es == {(0, 2), (2, 1), (1, 0)}:
        orient = [0, 2, 1]
    else:
        return None
    vf = [[Fraction(float(c)) for c in p] for p in v]
    a, b, c = (vf[i] for i in orient)
    signed = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
    if not signed:
        return None
    short = int(np.argmin([np.dot(v[(i + 1) % 3] - v[i], v[(i + 1) % 3] - v[i]) for i in range(3)]))
    return ([short, (short + 2) % 3, (short + 1) % 3], float(abs(signed)), float(1 if signed > 0 else -1))


def record(vertices, meta):
    """the 8-double triangle record of ``wsb_triangle2d``: v0, v1 - v0, v2 - v0 in ``meta``'s order, area2, sign."""
    v = vertices.detach().cpu().numpy() if torch.is_tensor(vertices) else np.asarray(vertices, dtype=np.float64)
    order, area2, sign = meta
    v0, v1, v2 = (np.asarray(v[i], dtype=np.float64) for i in order)
    e0, e1 = v1 - v0, v2 - v0
    return np.array([v0[0], v0[1], e0[0], e0[1], e1[0], e1[1], area2, sign], dtype=np.float64)


def _kid(kernel):
    try:
        return KERNEL_ID[kernel]
    except KeyError:
        raise KeyError(f"unknown kernel {kernel!r}; expected one of {sorted(KERNEL_ID)}") from None


def channels(vertices, positions, supports, metadata, kernel, background=0, valid=None, tolerance=2e-12,
             max_panels=4096):
    """device tensors (values [N, 12], estimate [N, 12], status [N]) in units of h (reference
    ``warptriangle.channels``); ``metadata`` None (not a conditioned triangle) gives zeros and status 0."""
    if tolerance <= 0 or max_panels < 1:
        raise ValueError("positive tolerance and max_panels required")
    L.require_device(positions, supports)
    dev = positions.device
    n = len(positions)
    values = torch.zeros((n, 12), dtype=F64, device=dev)
    estimate = torch.zeros_like(values)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    if metadata is None or n == 0:
        return values, estimate, status
    override(vertices, metadata, kernel, background, positions, supports, valid=valid, tolerance=tolerance,
             max_panels=max_panels, values=values, estimate=estimate, status=status)
    return values, estimate, status


def override(vertices, metadata, kernel, background, lpos, lsup, valid=None, rot4=None, flags=0, out=None, values=None,
             estimate=None, status=None, tolerance=2e-12, max_panels=4096):
    """one ``wsb_triangle2d`` launch: the rows taking the quadrature overwrite their rows of ``out`` (this
    representation's own contribution, finalised by ``flags``: ``OUT_PHYSICAL`` / ``OUT_WORLD``)."""
    n = len(lpos)
    if n == 0:
        return
    lpos = lpos.to(F64).contiguous()
    lsup = lsup.to(F64).contiguous()
    if valid is not None:
        valid = valid.to(torch.uint8).contiguous()
    rec = record(vertices, metadata)
    L.check(L.load().wsb_triangle2d(L.nptr(rec), int(background), _kid(kernel), L.ptr(lpos), L.ptr(lsup), L.ptr(valid), n,
                                    float(tolerance), int(max_panels), L.ptr(values), L.ptr(estimate), L.ptr(status),
                                    L.ptr(rot4), int(flags), L.ptr(out), L.stream_ptr(lpos.device)))


def channels_host(vertices, positions, supports, metadata, kernel, background=0, valid=None, tolerance=2e-12,
                  max_panels=4096):
    """the same rows evaluated on the host (numpy in and out): the device arithmetic without a GPU."""
    P = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 2))
    H = np.ascontiguousarray(np.broadcast_to(np.asarray(supports, dtype=np.float64), (len(P),)))
    n = len(P)
    values = np.zeros((n, 12))
    estimate = np.zeros((n, 12))
    status = np.zeros(n, dtype=np.int32)
    if metadata is None or n == 0:
        return values, estimate, status
    if tolerance <= 0 or max_panels < 1:
        raise ValueError("positive tolerance and max_panels required")
    V = None if valid is None else np.ascontiguousarray(np.asarray(valid, dtype=np.uint8))
    rec = record(vertices, metadata)
    L.check(L.load().wsb_triangle2d_host(L.nptr(rec), int(background), _kid(kernel), L.nptr(P), L.nptr(H), L.nptr(V), n,
                                         float(tolerance), int(max_panels), L.nptr(values), L.nptr(estimate),
                                         L.nptr(status)))
    return values, estimate, status
