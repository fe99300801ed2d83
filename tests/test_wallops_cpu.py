"""Wall-operation kernels of the full aggregate layout on the CPU: the product's truncated-power tables (cone, lw2,
lw4, w2p5, w4p5) equal the oracle's independent sympy derivation (shape, W'/r divided by -20 / -56/3, W^5) exactly,
each with mass 1; the factors of the reference's scene routes; the layout's groups and roles; the W^5 kernels only
give their gradient channels."""
import math
from fractions import Fraction as F

import pytest

from oracle import kernels as OK
from warpsphboundaries_amd import kernels as K
from warpsphboundaries_amd import wallops as W
from warpsphboundaries_amd.fused import FusedGroup
from warpsphboundaries_amd.provider import AnalyticBoundary


@pytest.mark.parametrize("name", ["cone", "lw2", "lw4", "w2p5", "w4p5"])
def test_tables_equal_the_independent_derivation(name):
    ob, kb = OK.blocks(name), K.kernel(name).blocks
    assert len(ob) == len(kb) == 1
    (R, oc), kblk = ob[0], kb[0]
    assert F(R) == kblk.R == 1
    kc = list(kblk.coeffs) + [F(0)] * (len(oc) - len(kblk.coeffs))
    assert [F(c) for c in oc] == kc
    assert K.plane_mass(name) == 1
    assert F(OK._SPEC[name][0]) == K.kernel(name).c2_pi


def test_normalisations_and_factors():
    assert K.kernel("cone").c2_pi == 3 and K.kernel("lw2").c2_pi == 10 and K.kernel("lw4").c2_pi == F(28, 3)
    H = 0.6
    assert W.cover_factor(H) == pytest.approx(-math.pi * H ** 3 / 3, rel=1e-15)
    assert W.lap_factor(H, "w2") == pytest.approx(-20 * 7 / (10 * H * H), rel=1e-15)
    assert W.lap_factor(H, "w4") == pytest.approx((-56 / 3) * 9 / ((28 / 3) * H * H), rel=1e-15)
    c5 = float(K.kernel("w2p5").c2_pi)
    assert W.tensile_factor(H, "w2") == pytest.approx(0.2 * 7 ** 5 / (math.pi ** 4 * c5 * H ** 8), rel=1e-15)
    with pytest.raises(NotImplementedError):
        W.lap_factor(H, "w6")


def test_layout_groups_and_roles():
    ab = AnalyticBoundary(scene=None, form="fp64")
    groups, idx = ab.layout("w2", True)
    assert [g.kernel for g in groups] == ["w2", "cone", "lw2", "w2p5"]
    assert idx == {"w": 0, "cone": 1, "lap": 2, "tens": 3}
    assert groups[1].channels == (3, 4) and groups[3].channels == (3, 4)
    groups, idx = ab.layout("Wendland4", False)
    assert [g.kernel for g in groups] == ["w4", "cone", "w4p5"] and idx == {"w": 0, "cone": 1, "tens": 2}
    with pytest.raises(NotImplementedError):
        ab.layout("w6", False)
    with pytest.raises(NotImplementedError):
        FusedGroup("w4p5").resolved()
    assert FusedGroup("w4p5", (4, 3)).resolved().channels == (3, 4)
