"""The C ABI library loads on the CPU host and exports every function include/wsb.h declares (no compute calls)."""
import ctypes
import os
import re

from warpsphboundaries_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"^\s*(?:int64_t|int|size_t|void|double)\s+\**(wsb_\w+)\s*\(", text, flags=re.M))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared()
    assert names, "no declarations found"
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_version_and_error_channel_without_a_device():
    lib = _lib.load()
    assert lib.wsb_version() >= 1
    # argument validation fails before any device call
    h = ctypes.c_void_p()
    code = lib.wsb_plan_create(9, None, 1, None, None, 0, None, 0, None, None, 0, ctypes.byref(h))
    assert code == 2 and "kernel" in _lib.last_error()
