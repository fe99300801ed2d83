"""Device solid-sphere closed form (FP64) against the exact rational evaluation."""
from fractions import Fraction as F

import numpy as np
import pytest
import torch

from warpsphboundaries_amd import sphere as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", ["cubic", "w2", "w4", "w6"])
def test_device_sphere_matches_exact(k):
    R, d = S.sample_pairs(4096, seed=3)
    R = np.concatenate([R, [0.05, 1e3, 0.5, 0.5, 2.0]])
    d = np.concatenate([d, [-0.05, 0.0, -0.5, 1.0, -1.0]])
    got = S.closed_form(torch.tensor(R, device="cuda:0"), torch.tensor(d, device="cuda:0"), k).cpu().numpy()
    ref = np.array([float(S.lambda_exact(k, F(float(r)), F(float(x)))) for r, x in zip(R, d)])
    err = np.abs(got - ref)
    # double-double evaluation: FP64-exact up to the final rounding
    assert (err <= 2.3e-16 * np.abs(ref) + 1e-26).all(), err.max()
