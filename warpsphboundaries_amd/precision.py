"""Process-wide precision of the wall integrals (reference ``edge/precision.py:1-68``).

The reference follows the scalar type of the SPH solver it is coupled to: the environment variable
``warpSPHCore_PRECISION`` (float16 | float32 | float64), read once per process, with float64 as this package's
default (its exactness contracts are float64).  Here the same variable selects the default arithmetic form of every
evaluation that is not given one explicitly (``AnalyticBoundary(scene)``, ``FusedGroup(kernel)``,
``edge.edge_channels(...)``):

* float64 -> ``"fp64"``: the FP64 closed forms (1e-12 of the FP64 oracle);
* float32 or float16 -> ``"mixed"``: the same closed forms with FP32 hardware angles and square-root seeds (at most
  1e-6 signal-relative per channel), the role of the reference's float32 Chebyshev route.

A per-call ``form`` always wins.  The returned tensors are float64 in every precision, as in the reference.
"""
from __future__ import annotations

import os
import warnings

import torch

_NAMES = ("float16", "float32", "float64")
real = os.environ.get("warpSPHCore_PRECISION", "float64")
if real not in _NAMES:
    raise ValueError(f"warpSPHCore_PRECISION={real!r}: expected one of {_NAMES}")
IS_F64 = real == "float64"
IS_F32 = real == "float32"
DEFAULT_FORM = "fp64" if IS_F64 else "mixed"
torch_real = {"float16": torch.float16, "float32": torch.float32, "float64": torch.float64}[real]

if not IS_F64 and "warpSPHBoundaries_precision_warned" not in os.environ:
    os.environ["warpSPHBoundaries_precision_warned"] = "1"
    warnings.warn(f"wall integrals default to the mixed form ({real}); the exactness contracts are float64 "
                  "(set warpSPHCore_PRECISION=float64)", RuntimeWarning, stacklevel=2)


def default_form(form=None) -> str:
    """``form`` if given, else the process default."""
    return DEFAULT_FORM if form is None else form


def require_f64(what):
    if not IS_F64:
        raise NotImplementedError(f"{what} is float64 only (active precision {real}): set warpSPHCore_PRECISION=float64")


def sync(device):
    """synchronise ``device`` unless a graph is being captured on the current stream (a synchronisation is illegal
    inside a capture)."""
    dev = torch.device(device)
    if dev.type != "cpu" and not torch.cuda.is_current_stream_capturing():
        torch.cuda.synchronize(dev)
