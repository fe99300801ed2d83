"""The process-wide precision switch (reference edge/precision.py): warpSPHCore_PRECISION selects the default form of
the evaluations that are not given one; float64 is the default; an explicit form wins; a bad name is an error."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = ("import warnings; warnings.simplefilter('ignore');"
         "from warpsphboundaries_amd import precision as P;"
         "from warpsphboundaries_amd.fused import FusedGroup;"
         "print(P.real, P.DEFAULT_FORM, FusedGroup('w4').resolved().form, FusedGroup('w4', None, 'plain').resolved().form)")


def _run(value):
    env = {k: v for k, v in os.environ.items() if k not in ("warpSPHCore_PRECISION", "warpSPHBoundaries_precision_warned")}
    if value is not None:
        env["warpSPHCore_PRECISION"] = value
    return subprocess.run([sys.executable, "-c", PROBE], cwd=ROOT, env=env, capture_output=True, text=True)


@pytest.mark.parametrize("value,expect", [(None, "float64 fp64 fp64 plain"), ("float64", "float64 fp64 fp64 plain"),
                                          ("float32", "float32 mixed mixed plain"), ("float16", "float16 mixed mixed plain")])
def test_default_form_follows_the_process_precision(value, expect):
    r = _run(value)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == expect


def test_unknown_precision_is_rejected():
    r = _run("float128")
    assert r.returncode != 0 and "warpSPHCore_PRECISION" in r.stderr
