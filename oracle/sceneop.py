"""numpy restatement of the boundary operations over per-query moments -- TEST INFRASTRUCTURE (see oracle/__init__).

* ``world_moments``: the per-query moments of one surface loop of a posed body, world frame, physical units (lam,
  m1, g0, g1), from the oracle's slot fill and indicator (``adjacency.fixed_topology``) and its edge channels
  (``edge2d.edge_channels12``), i.e. what ``Scene.precompute`` returns summed per query (reference
  ``scene/scene.py:1418-1445``).
* ``apply``: one body's contribution of an operation (reference ``scene/scene.py:1640-1707``, ``_apply``): Density,
  Covariance, Interpolate, Gradient in the four schemes, Divergence, Curl, and the reaction (force, torque) of a
  scalar Gradient.

Pinned in ``tests/test_sceneop_cpu.py`` by dense Gauss quadrature of the defining integrals (int A(x') W dA',
int (s A(x') + f) grad_x W dA', int (x' - c) x grad_x W dA') over a triangulation of the body, which shares no step
with the moment closed forms.
"""
from __future__ import annotations

import numpy as np

from . import adjacency as oadj
from . import edge2d


def world_moments(V, E, background, center, angle, pts, sup, kname, smax):
    """{lam [N], m1 [N,2], g0 [N,2], g1 [N,2,2]} of the loop (V [nv,2] body frame, edges E [ne,2]) of a body at
    (center, angle) at world points ``pts`` with supports ``sup``."""
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[c, -s], [s, c]])
    d = pts - np.asarray(center, dtype=np.float64)
    local = np.stack([d[:, 0] * c + d[:, 1] * s, d[:, 1] * c - d[:, 0] * s], 1)
    st = oadj.Static(np.asarray(V, np.float64), np.asarray(E, np.int64), background, smax)
    e, cnt, ind = oadj.fixed_topology(st, local, sup)
    K = st.K
    rows = np.repeat(np.arange(len(pts)), cnt)
    sl = np.concatenate([e[r * K:r * K + k] for r, k in enumerate(cnt)]) if len(rows) else np.zeros(0, np.int64)
    ch = np.zeros((len(pts), 12))
    if len(rows):
        P = (st.verts[st.edges[sl, 0]] - local[rows]) / sup[rows, None]
        Q = (st.verts[st.edges[sl, 1]] - local[rows]) / sup[rows, None]
        np.add.at(ch, rows, edge2d.edge_channels12(P, Q, kname))
    ch = ch + ind[:, None] * edge2d.plane_vector(kname)[None]
    h = sup[:, None]
    return {"lam": ch[:, 0],
            "m1": (ch[:, 1:3] * h) @ R.T,
            "g0": (ch[:, 3:5] / h) @ R.T,
            "g1": R @ ch[:, 5:9].reshape(-1, 2, 2) @ R.T}


def apply(op, scheme, mom, a0, a1, rho, per_query, center, pos, rho_q=None, qv=None, masses=None):
    """(out, force, torque) of one body: ``mom`` from world_moments (every query a row); a0/a1 the BodyField
    (a0 scalar or [C], a1 [2] or [C,2]; per_query: a0 [N] or [N,C], a1 [2] / [N,2] (scalar) or [C,2] / [N,C,2])."""
    N = len(pos)
    if op == "Density":
        return rho * mom["lam"], None, None
    if op == "Covariance":
        return rho * mom["g1"], None, None
    a0 = np.asarray(a0, np.float64)
    if per_query:
        scalar = a0.ndim == 1
        A = a0.reshape(N, -1)
        C = A.shape[1]
        G = None
        if a1 is not None:
            a1 = np.asarray(a1, np.float64)
            if scalar:
                G = np.broadcast_to(a1.reshape(-1, 1, 2), (N, 1, 2))
            else:
                G = np.broadcast_to(a1.reshape(-1, C, 2), (N, C, 2))
    else:
        scalar = a0.ndim == 0
        a0 = a0.reshape(-1)
        C = len(a0)
        A = np.broadcast_to(a0[None], (N, C)).copy()
        G = None
        if a1 is not None:
            a1 = np.asarray(a1, np.float64).reshape(C, 2)
            A = A + (pos - np.asarray(center)) @ a1.T
            G = np.broadcast_to(a1[None], (N, C, 2))
    if op == "Interpolate":
        out = A * mom["lam"][:, None]
        if G is not None:
            out = out + np.einsum("pcd,pd->pc", G, mom["m1"])
        return (out[:, 0] if scalar else out), None, None
    if scheme == "Naive":
        s, f = np.ones(N), np.zeros((N, C))
    else:
        f = np.asarray(qv, np.float64).reshape(N, -1)
        if scheme == "Symmetric":
            s = rho_q / rho
            f = (rho / rho_q)[:, None] * f
        elif scheme == "Difference":
            s, f = np.ones(N), -f
        else:
            s = np.ones(N)
    A0 = s[:, None] * A + f
    out = A0[:, :, None] * mom["g0"][:, None, :]
    if G is not None:
        out = out + s[:, None, None] * np.einsum("pcd,pdj->pcj", G, mom["g1"])
    force = torque = None
    if scalar and masses is not None:
        force = -(masses[:, None] * out[:, 0, :]).sum(0)
        if G is None:
            r = pos - np.asarray(center)
            g1, g0 = mom["g1"], mom["g0"]
            lever = (g1[:, 0, 1] - g1[:, 1, 0]) + (r[:, 0] * g0[:, 1] - r[:, 1] * g0[:, 0])
            torque = -(masses * A0[:, 0] * lever).sum()
    if op == "Gradient":
        return (out[:, 0, :] if scalar else out), force, torque
    if op == "Divergence":
        return out[:, 0, 0] + out[:, 1, 1], force, torque
    if op == "Curl":
        return out[:, 1, 0] - out[:, 0, 1], force, torque
    raise NotImplementedError(op)
