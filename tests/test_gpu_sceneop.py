"""sceneOperation over PairMoments (reference scene/scene.py:1717-1806) on the device against the oracle
(oracle/sceneop.py: the oracle's per-query moments and the operation restated in numpy, pinned by quadrature in
tests/test_sceneop_cpu.py): two posed bodies (an L-shape and a 96-edge star), every operation, the four gradient
schemes, scalar and vector fields, per-query fields, perBody, the reaction, and the guards.  FP64 form; the moment
gates of the single launches (1e-12 of the signal) carried through the linear operations."""
import numpy as np
import pytest
import torch

from oracle import sceneop as SO
from warpsphboundaries_amd import workloads as WL
from warpsphboundaries_amd.operation import (BodyField, GradientScheme, OperationProperties, SceneReaction,
                                             WarpOperation, sceneOperation)
from warpsphboundaries_amd.scene import Body, ParticleState, Scene, SurfaceRep

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
LSHAPE = np.array([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], dtype=float)
N = 3000


@pytest.fixture(scope="module")
def world():
    star = WL.star_polygon(96, g=WL.rng(2)) * 1.2 + np.array([-2.6, 0.4])
    bodies = [Body(bodyId=0, reps=[SurfaceRep.polygon(LSHAPE)], center=(0.3, -0.2), angle=0.7,
                   linearVelocity=(0.4, -0.1), angularVelocity=0.9),
              Body(bodyId=1, reps=[SurfaceRep.polygon(star)], center=(0.1, 0.2), angle=-0.3,
                   linearVelocity=(0.2, 0.3), angularVelocity=-0.5)]
    sc = Scene(bodies, DEV)
    g = np.random.default_rng(5)
    pos = g.uniform(-3.5, 2.6, (N, 2))
    sup = g.uniform(0.3, 0.6, N)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=DEV)
    ps = ParticleState(t(pos), t(sup), masses=t(g.uniform(0.5, 1.5, N)), densities=t(g.uniform(0.9, 1.1, N)))
    mom = []
    for b in bodies:
        rep = b.reps[0]
        mom.append(SO.world_moments(rep.vertices.cpu().numpy(), rep.edges.cpu().numpy(), rep.background,
                                    b.center.cpu().numpy(), b.angle, pos, sup, "w4", sup.max()))
    host = {"pos": pos, "sup": sup, "masses": ps.masses.cpu().numpy(), "rho_q": ps.densities.cpu().numpy(),
            "qv": g.normal(size=N), "qv2": g.normal(size=(N, 2))}
    return sc, ps, mom, host


def _props(op, mode=GradientScheme.Naive, kernel="w4"):
    return OperationProperties(kernel=kernel, operation=op, gradientMode=mode)


def _check(got, ref, tol=1e-12):
    got = got.cpu().numpy() if torch.is_tensor(got) else got
    assert got.shape == ref.shape, (got.shape, ref.shape)
    scale = max(1.0, float(np.abs(ref).max()))
    err = float(np.abs(got - ref).max())
    assert err <= tol * scale, (err, scale)


def _oracle(sc, mom, host, op, scheme, fields, qv=None):
    outs, force, torque = [], [], []
    for b, m, f in zip(sc.bodies, mom, fields):
        a0 = None if f.a0 is None else torch.as_tensor(f.a0).cpu().numpy()
        a1 = None if f.a1 is None else torch.as_tensor(f.a1).cpu().numpy()
        o, fo, to = SO.apply(op, scheme, m, a0, a1, f.rho, f.perQuery, b.center.cpu().numpy(), host["pos"], host["rho_q"],
                             qv, host["masses"])
        outs.append(o)
        force.append(fo)
        torque.append(to)
    return outs, force, torque


@pytest.mark.parametrize("op", ["Density", "Covariance", "Interpolate", "Gradient", "Divergence", "Curl"])
def test_operations_of_rigid_fields_match_the_oracle(world, op):
    sc, ps, mom, host = world
    fields = [BodyField.rigid(b, rho=1.3) for b in sc.bodies]
    got = sceneOperation(ps, _props(getattr(WarpOperation, op)), sc, bodyFields=fields)
    per = sceneOperation(ps, _props(op), sc, bodyFields=fields, perBody=True)
    ref, _, _ = _oracle(sc, mom, host, op, "Naive", fields)
    _check(got, ref[0] + ref[1])
    assert per.shape[0] == 2
    for b in range(2):
        _check(per[b], ref[b])
    assert np.abs(ref[0]).max() > 1e-3 and np.abs(ref[1]).max() > 1e-3


@pytest.mark.parametrize("mode", ["Naive", "Symmetric", "Difference", "Summation"])
def test_gradient_schemes_scalar_and_vector_fields(world, mode):
    sc, ps, mom, host = world
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=DEV)
    fields = [BodyField(t(0.7), t([0.3, -0.5]), rho=1.2), BodyField(t(-0.4), t([0.1, 0.2]), rho=0.8)]
    qv = t(host["qv"])
    got = sceneOperation(ps, _props(WarpOperation.Gradient, getattr(GradientScheme, mode)), sc, queryValues=qv,
                         bodyFields=fields)
    ref, _, _ = _oracle(sc, mom, host, "Gradient", mode, fields, host["qv"])
    _check(got, ref[0] + ref[1])
    vfields = [BodyField(t([0.4, -0.1]), t([[0.2, -0.9], [0.7, 0.3]]), rho=1.1), BodyField(t([0.0, 0.5]), None)]
    got = sceneOperation(ps, _props("Gradient", mode), sc, queryValues=t(host["qv2"]), bodyFields=vfields)
    ref, _, _ = _oracle(sc, mom, host, "Gradient", mode, vfields, host["qv2"])
    _check(got, ref[0] + ref[1])


def test_per_query_fields_equal_the_body_field_at_the_queries(world):
    sc, ps, mom, host = world
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=DEV)
    a0, a1 = np.array([0.4, -0.1]), np.array([[0.2, -0.9], [0.7, 0.3]])
    fields, pq = [], []
    for b in sc.bodies:
        Ax = a0[None] + (host["pos"] - b.center.cpu().numpy()) @ a1.T
        fields.append(BodyField(t(a0), t(a1)))
        pq.append(BodyField(t(Ax), t(a1), perQuery=True))
    for op in ("Interpolate", "Gradient", "Divergence", "Curl"):
        a = sceneOperation(ps, _props(op), sc, bodyFields=fields)
        b = sceneOperation(ps, _props(op), sc, bodyFields=pq)
        ref, _, _ = _oracle(sc, mom, host, op, "Naive", pq)
        _check(b, ref[0] + ref[1])
        _check(a, b.cpu().numpy(), 1e-13)


def test_reaction_force_and_torque(world):
    sc, ps, mom, host = world
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=DEV)
    fields = [BodyField(t(1.7)), BodyField(t(0.9))]
    out, rea = sceneOperation(ps, _props("Gradient"), sc, bodyFields=fields, returnReaction=True)
    assert isinstance(rea, SceneReaction) and rea.torqueExact == [True, True]
    ref, force, torque = _oracle(sc, mom, host, "Gradient", "Naive", fields)
    _check(out, ref[0] + ref[1])
    _check(rea.force, np.stack(force))
    _check(rea.torque, np.array(torque))
    # a position-dependent field: the torque is not determined by the moments
    lin = [BodyField(t(1.7), t([0.3, 0.1])), BodyField(t(0.9))]
    _, rea = sceneOperation(ps, _props("Gradient"), sc, bodyFields=lin, returnReaction=True)
    assert rea.torqueExact == [False, True] and torch.isnan(rea.torque[0]) and not torch.isnan(rea.torque[1])


def test_shared_moments_pruned_channels_and_guards(world):
    sc, ps, mom, host = world
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=DEV)
    pm = sc.pairMoments(ps, "w4", form="fp64")
    one = [BodyField(t(1.0)), BodyField(t(1.0))]
    a = sceneOperation(ps, _props("Gradient"), sc, pm, bodyFields=one)
    b = sceneOperation(ps, _props("Gradient"), sc, bodyFields=one)
    assert torch.equal(a, b)
    pruned = sc.pairMoments(ps, "w4", channels=(3, 4), form="fp64")
    c = sceneOperation(ps, _props("Gradient"), sc, pruned, bodyFields=one)
    assert torch.equal(c, a)
    with pytest.raises(ValueError):
        sceneOperation(ps, _props("Density"), sc, pruned)
    with pytest.raises(ValueError):
        sceneOperation(ps, _props("Gradient", kernel="w2"), sc, pm, bodyFields=one)
    with pytest.raises(ValueError):
        sceneOperation(ps, _props("Gradient", "Symmetric"), sc, pm, bodyFields=one)
    with pytest.raises(ValueError):
        sceneOperation(ps, _props("Gradient"), sc, pm)
    # the adjacency of a scene evaluated for one operation
    adj = sc.adjacency(ps, "w4")
    d = sceneOperation(ps, _props("Density"), sc, adj)
    ref, _, _ = _oracle(sc, mom, host, "Density", "Naive", [BodyField(), BodyField()])
    _check(d, ref[0] + ref[1])
