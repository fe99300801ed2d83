"""CPU oracle jobs of the configuration parity tests (importable by spawned worker processes, which never touch the
GPU)."""
import multiprocessing
import os
from concurrent.futures import ProcessPoolExecutor
from fractions import Fraction

import numpy as np


def workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def pool_map(fn, jobs):
    """fn(*job) for every job, in a pool of spawned CPU-only processes."""
    with ProcessPoolExecutor(workers(), mp_context=multiprocessing.get_context("spawn")) as ex:
        return list(ex.map(fn, *zip(*jobs)))


def sphere_exact(kernel, R, d):
    """the exact rational closed form of the solid sphere at each (R, d) (warpsphboundaries_amd.sphere.lambda_exact),
    rounded to float64."""
    from warpsphboundaries_amd import sphere as S
    return np.array([float(S.lambda_exact(kernel, Fraction(float(r)), Fraction(float(x)))) for r, x in zip(R, d)])


def compact_rows(V, F, P, face_lists, kernel, h, winding):
    """[len(P), 22] stored moments of the compact form: each query's own face list (oracle.mesh3d in FP64) plus its
    winding number times the whole-ball vector (a query with no face: that term alone)."""
    from oracle import mesh3d as OM
    from warpsphboundaries_amd.plan3d import H_POWER3, REQUESTS22, whole_vector3
    out = np.zeros((len(P), 22))
    u = h ** np.array(H_POWER3, dtype=np.float64)
    for j, fl in enumerate(face_lists):
        if len(fl):
            out[j] = OM.closed_surface_requests(V, F[fl], P[j:j + 1], kernel, h, list(REQUESTS22), chunk=1,
                                                winding=np.asarray(winding[j:j + 1]))[0]
        else:
            out[j] = winding[j] * np.asarray(whole_vector3(kernel), dtype=np.float64) * u
    return out
