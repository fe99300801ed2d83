"""Precision stress workload helpers (CPU side of tests/test_gpu_precision_stress.py and of the mpmath fixture script
tests/golden/make_stress2d_mp.py): oracle values of the wedge polygons and sliver tetrahedra of
``workloads.config4_2d`` / ``config4_3d`` in float64, long double or mpmath, and the error summary.

Errors are signal-relative per channel and case (``max_j |v - r| / max_j |r|`` over one polygon's or one
tetrahedron's particles, in units of h), then the maximum over cases.
"""
from __future__ import annotations

import numpy as np

from oracle import adjacency as oadj
from oracle import arith as A
from oracle import edge2d
from oracle import mesh3d as OM


def near_edges(V, pts, h, slack=1e-6):
    """(rows, edges) of every edge closer than h (1 + slack) to each point (float64 segment distance): the others
    contribute exactly zero."""
    n = len(V)
    a, b = V, np.roll(V, -1, axis=0)
    d = b - a
    rel = pts[:, None, :] - a[None]
    t = np.clip((rel * d[None]).sum(-1) / (d * d).sum(-1)[None], 0.0, 1.0)
    dist = np.linalg.norm(rel - t[..., None] * d[None], axis=-1)
    rows, edges = np.nonzero(dist < h * (1 + slack))
    return rows, edges


def polygon_oracle(V, pts, h, kernels, mode="ld", dps=50):
    """{kernel: [n, 12]} oracle values in units of h (channels in device order) for points ``pts`` of polygon V
    (counter-clockwise around the solid): edge-local channels of the edges within the support plus indicator x the
    whole-support vector; the indicator is the float64 crossing count (the points lie off the boundary)."""
    n = len(V)
    edges = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
    ind = oadj.indicator_points(V, edges, 0, pts, np.full(len(pts), h))
    rows, e = near_edges(V, pts, h)
    P = (V[e] - pts[rows]) / h
    Q = (V[(e + 1) % n] - pts[rows]) / h
    out = {}
    with A.precision(mode, dps):
        for k in kernels:
            c = A.to_float(edge2d.edge_channels12(A.asarray(P), A.asarray(Q), k)) if len(rows) else np.zeros((0, 12))
            tot = np.zeros((len(pts), 12))
            np.add.at(tot, rows, c)
            out[k] = tot + ind[:, None] * np.asarray(edge2d.plane_vector(k), dtype=np.float64)[None]
    return out


def tetra_oracle(V, faces, pts, h, kernels, requests, mode="ld", chunk=64):
    """{kernel: [n, len(requests)]} values in physical units over the solid tetrahedron (the reference's regular
    radial-flux form over all four faces) in the working precision ``mode``, rounded to float64."""
    out = {}
    with A.precision(mode):
        for k in kernels:
            out[k] = A.to_float(OM.closed_surface_requests(V, faces, pts, k, h, requests, chunk=chunk))
    return out


def case_errors(values, reference, case):
    """{case id: [channels] signal-relative error} and the maximum over cases, per channel."""
    worst = np.zeros(values.shape[1])
    for c in np.unique(case):
        m = case == c
        err = np.abs(values[m] - reference[m]).max(0)
        sig = np.abs(reference[m]).max(0)
        rel = np.where(sig > 0, err / np.where(sig > 0, sig, 1.0), err)
        worst = np.maximum(worst, rel)
    return worst


KERNELS = ("cubic", "w2", "w4", "w6")


def oracle_2d_job(V, pts, h):
    """process-pool job: the long double oracle of one wedge."""
    return polygon_oracle(V, pts, h, KERNELS, mode="ld")


def oracle_3d_job(V, faces, pts, h, requests):
    """process-pool job: the long double oracle of one tetrahedron."""
    return tetra_oracle(V, faces, pts, h, KERNELS, requests, mode="ld", chunk=256)
