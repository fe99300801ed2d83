"""Pin of oracle/sceneop.py (the boundary operations over per-query moments, reference scene/scene.py:1640-1707)
by dense Gauss quadrature of the defining integrals over the posed L-shaped body (three unit squares, each cut into
16 x 16 cells with 6 x 6 Gauss points), which shares no step with the moment closed forms:

    Density      rho int W dA'                     Covariance   rho int y (x) grad_x W dA'
    Interpolate  int A(x') W dA'                   Gradient     int (s A(x') + f) (x) grad_x W dA'
    Divergence / Curl of the vector Gradient       torque       -sum_i m_i A int (x' - c) x grad_x W dA'

with y = x' - x, grad_x W(|x - x'|) = -W'(r) y / r, and the Wendland C4 kernel written out here (the reference's
own test of the torque uses the same form, tests/scene/test_scene.py:79-118)."""
import numpy as np
import pytest

from oracle import sceneop as SO

LSHAPE = np.array([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], dtype=float)
SQUARES = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)]
CENTER, ANGLE = np.array([0.3, -0.2]), 0.7


def _quadrature_points(m=16, n=6):
    gx, gw = np.polynomial.legendre.leggauss(n)
    gx, gw = (gx + 1) / 2, gw / 2
    u, v = np.meshgrid(gx, gx, indexing="ij")
    w = np.outer(gw, gw)
    pts, wts = [], []
    d = 1.0 / m
    for x0, y0 in SQUARES:
        for i in range(m):
            for j in range(m):
                pts.append(np.stack([x0 + (i + u.ravel()) * d, y0 + (j + v.ravel()) * d], 1))
                wts.append(w.ravel() * d * d)
    pts, wts = np.concatenate(pts), np.concatenate(wts)
    c, s = np.cos(ANGLE), np.sin(ANGLE)
    return pts @ np.array([[c, -s], [s, c]]).T + CENTER, wts


def _w4(r, h):
    q = r / h
    inside = q < 1
    W = 9 / np.pi * (1 - q) ** 6 * (1 + 6 * q + 35 / 3 * q ** 2) * inside / h ** 2
    dW = 9 / np.pi * (-6 * (1 - q) ** 5 * (1 + 6 * q + 35 / 3 * q ** 2) + (1 - q) ** 6 * (6 + 70 / 3 * q)) * inside / h ** 3
    return W, dW


@pytest.fixture(scope="module")
def setup():
    g = np.random.default_rng(1)
    n = 40
    pos = g.uniform(-1.2, 2.4, (n, 2))
    sup = g.uniform(0.5, 0.9, n)
    X, wq = _quadrature_points()
    lam = np.zeros(n)
    grad = np.zeros((n, 2))
    quad = {"W": [], "gradW": []}
    for i in range(n):
        y = X - pos[i]
        r = np.linalg.norm(y, axis=1)
        W, dW = _w4(r, sup[i])
        gx = -(dW / np.maximum(r, 1e-300))[:, None] * y
        quad["W"].append(W * wq)
        quad["gradW"].append(gx * wq[:, None])
    E = np.stack([np.arange(6), (np.arange(6) + 1) % 6], 1)
    mom = SO.world_moments(LSHAPE, E, 0, CENTER, ANGLE, pos, sup, "w4", sup.max())
    state = {"pos": pos, "sup": sup, "masses": g.uniform(0.5, 1.5, n), "rho_q": g.uniform(0.9, 1.1, n),
             "qv": g.normal(size=n), "qv2": g.normal(size=(n, 2))}
    return X, quad, mom, state


def _field(X, a0, a1):
    a0 = np.asarray(a0, dtype=float)
    if a1 is None:
        return np.broadcast_to(a0.reshape(1, -1), (len(X), a0.size))
    return a0.reshape(1, -1) + (X - CENTER) @ np.asarray(a1, dtype=float).reshape(a0.size, 2).T


def _scale(ref):
    return max(1.0, float(np.abs(ref).max()))


def test_density_covariance_and_interpolate(setup):
    X, quad, mom, st = setup
    n = len(st["pos"])
    rho = 1.3
    dens = SO.apply("Density", "Naive", mom, None, None, rho, False, CENTER, st["pos"])[0]
    ref = rho * np.array([quad["W"][i].sum() for i in range(n)])
    assert np.abs(dens - ref).max() <= 1e-7 * _scale(ref)
    assert (ref > 1e-3).sum() > 10 and (ref < 1e-12).sum() > 3
    cov = SO.apply("Covariance", "Naive", mom, None, None, rho, False, CENTER, st["pos"])[0]
    ref = rho * np.stack([np.einsum("pd,pj->dj", X - st["pos"][i], quad["gradW"][i]) for i in range(n)])
    assert np.abs(cov - ref).max() <= 1e-7 * _scale(ref)
    # a scalar linear field and a rigid velocity field
    for a0, a1 in [(0.7, [0.3, -0.5]), ([0.4, -0.1], [[0.0, -0.9], [0.9, 0.0]])]:
        out = SO.apply("Interpolate", "Naive", mom, a0, a1, rho, False, CENTER, st["pos"])[0]
        A = _field(X, a0, a1)
        ref = np.stack([(A * quad["W"][i][:, None]).sum(0) for i in range(n)])
        ref = ref[:, 0] if np.ndim(a0) == 0 else ref
        assert np.abs(out - ref).max() <= 1e-7 * _scale(ref), (a0, np.abs(out - ref).max())


@pytest.mark.parametrize("scheme", ["Naive", "Symmetric", "Difference", "Summation"])
def test_gradient_schemes_of_a_linear_scalar_field(setup, scheme):
    X, quad, mom, st = setup
    n = len(st["pos"])
    rho, a0, a1 = 1.2, 0.7, [0.3, -0.5]
    out, force, torque = SO.apply("Gradient", scheme, mom, a0, a1, rho, False, CENTER, st["pos"], st["rho_q"],
                                  st["qv"], st["masses"])
    A = _field(X, a0, a1)[:, 0]
    ref = np.zeros((n, 2))
    for i in range(n):
        s, f = 1.0, 0.0
        if scheme == "Symmetric":
            s, f = st["rho_q"][i] / rho, rho / st["rho_q"][i] * st["qv"][i]
        elif scheme == "Difference":
            f = -st["qv"][i]
        elif scheme == "Summation":
            f = st["qv"][i]
        ref[i] = ((s * A + f)[:, None] * quad["gradW"][i]).sum(0)
    assert np.abs(out - ref).max() <= 1e-7 * _scale(ref), np.abs(out - ref).max()
    np.testing.assert_allclose(force, -(st["masses"][:, None] * out).sum(0), rtol=0, atol=1e-14)
    assert torque is None               # a position-dependent field: the moments do not give the torque


def test_vector_gradient_divergence_and_curl(setup):
    X, quad, mom, st = setup
    n = len(st["pos"])
    a0, a1 = [0.4, -0.1], [[0.2, -0.9], [0.7, 0.3]]
    A = _field(X, a0, a1)
    ref = np.stack([np.einsum("pc,pj->cj", A, quad["gradW"][i]) for i in range(n)])
    out = SO.apply("Gradient", "Naive", mom, a0, a1, 1.0, False, CENTER, st["pos"])[0]
    assert out.shape == (n, 2, 2)
    assert np.abs(out - ref).max() <= 1e-7 * _scale(ref)
    div = SO.apply("Divergence", "Naive", mom, a0, a1, 1.0, False, CENTER, st["pos"])[0]
    curl = SO.apply("Curl", "Naive", mom, a0, a1, 1.0, False, CENTER, st["pos"])[0]
    assert np.abs(div - (ref[:, 0, 0] + ref[:, 1, 1])).max() <= 2e-7 * _scale(ref)
    assert np.abs(curl - (ref[:, 1, 0] - ref[:, 0, 1])).max() <= 2e-7 * _scale(ref)
    # the field given at each query (its value there and the same gradient) is the same operation
    Ax = np.asarray(a0)[None] + (st["pos"] - CENTER) @ np.asarray(a1).T
    pq = SO.apply("Gradient", "Naive", mom, Ax, a1, 1.0, True, CENTER, st["pos"])[0]
    np.testing.assert_allclose(pq, out, rtol=0, atol=1e-14 * _scale(out))


def test_torque_of_a_constant_pressure(setup):
    X, quad, mom, st = setup
    n = len(st["pos"])
    p = 1.7
    out, force, torque = SO.apply("Gradient", "Naive", mom, p, None, 1.0, False, CENTER, st["pos"], masses=st["masses"])
    arm = X - CENTER
    tau = 0.0
    for i in range(n):
        gw = quad["gradW"][i]
        tau += -st["masses"][i] * p * np.sum(arm[:, 0] * gw[:, 1] - arm[:, 1] * gw[:, 0])
    assert abs(torque - tau) <= 1e-7 * max(1.0, abs(tau))
    assert abs(tau) > 1e-3
