"""ctypes binding of the HIP library ``libwsb.so`` (declarations in include/wsb.h).

The product path has no CPU fallback: every entry point needs the built library and device (HIP) tensors and
raises when either is missing.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_size_t, c_void_p

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WSB_LIBRARY", os.path.join(HERE, "libwsb.so"))

_ERRORS = {1: ValueError, 2: NotImplementedError, 3: ValueError, 4: MemoryError, 5: RuntimeError}

# every exported symbol with (restype, argtypes); tests check this list against include/wsb.h
SIGNATURES = {
    "wsb_last_error": (c_int, [c_char_p, c_size_t]),
    "wsb_version": (c_int, []),
    "wsb_plan_create": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                c_int, POINTER(c_void_p)]),
    "wsb_plan_destroy": (c_int, [c_void_p]),
    "wsb_geom2d_create": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_double, POINTER(c_void_p), POINTER(c_int)]),
    "wsb_geom2d_destroy": (c_int, [c_void_p]),
    "wsb_geom2d_info": (c_int, [c_void_p, POINTER(c_int64)]),
    "wsb_geom2d_arrays": (c_int, [c_void_p, POINTER(c_void_p), POINTER(c_void_p)]),
    "wsb_sort2d_scratch_bytes": (c_size_t, [c_void_p, c_int]),
    "wsb_sort2d": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_count_order2d_scratch_bytes": (c_size_t, []),
    "wsb_count_order2d": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_adjacency2d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                c_void_p]),
    "wsb_moments2d_scratch_bytes": (c_size_t, [c_int, c_int]),
    "wsb_moments2d": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                              c_int, c_void_p, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_edge_channels2d": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                    c_void_p, c_int, c_void_p, c_void_p]),
    "wsb_signed_distance2d": (c_int, [c_void_p, c_void_p, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "wsb_sphere_closed_form": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "wsb_moments2d_multi": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                    c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "wsb_adjacency2d_sorted": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_int, c_void_p]),
    "wsb_moments2d_multi_sorted": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int,
                                           c_void_p, c_int, c_void_p, c_void_p]),
    "wsb_cone_area2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_double,
                                c_double, c_double, c_void_p, c_void_p]),
    "wsb_dir_extreme2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                  c_double, c_double, c_void_p, c_void_p]),
    "wsb_triangle2d": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_double, c_int, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "wsb_triangle2d_host": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_double, c_int,
                                    c_void_p, c_void_p, c_void_p]),
    "wsb_geom2d_desc_bytes": (c_size_t, []),
    "wsb_geom2d_desc": (c_int, [c_void_p, c_void_p]),
    "wsb_batch2d_adjacency": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p]),
    "wsb_desc2d_cell_bases": (c_int, [c_void_p, c_int, c_void_p]),
    "wsb_batch2d_cellsort_scratch_bytes": (c_size_t, [c_int, c_int]),
    "wsb_batch2d_cellsort": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
    "wsb_batch2d_moments": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                    c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "wsb_batch2d_plain_moments": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                          c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "wsb_row_segment_bytes": (c_size_t, []),
    "wsb_gather_rows": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "wsb_scatter_rows": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "wsb_fill_rows_i32": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "wsb_disks2d_create": (c_int, [c_void_p, c_void_p, c_int, c_double, POINTER(c_void_p), POINTER(c_int)]),
    "wsb_disks2d_destroy": (c_int, [c_void_p]),
    "wsb_disks2d_info": (c_int, [c_void_p, POINTER(c_int64)]),
    "wsb_disk_adjacency2d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "wsb_disk_moments2d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int,
                                   c_void_p, c_void_p]),
    "wsb_disk_moments2d_multi": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int,
                                         c_void_p, c_void_p]),
    "wsb_disk_channels_host": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "wsb_disk_cone_area2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_double,
                                     c_double, c_double, c_void_p, c_void_p]),
    "wsb_disk_dir_extreme2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                       c_double, c_double, c_void_p, c_void_p]),
    "wsb_mesh3d_create": (c_int, [c_void_p, c_int, c_void_p, c_int, c_double, c_int, POINTER(c_void_p), POINTER(c_int64)]),
    "wsb_mesh3d_destroy": (c_int, [c_void_p]),
    "wsb_mesh3d_info": (c_int, [c_void_p, POINTER(c_int64)]),
    "wsb_mesh3d_grid": (c_int, [c_void_p, POINTER(c_double)]),
    "wsb_adjacency3d_scratch_bytes": (c_size_t, [c_int]),
    "wsb_adjacency3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                c_size_t, c_void_p]),
    "wsb_adjacency3d_sorted": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int64, c_void_p,
                                       c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_adjacency3d_bvh": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                c_size_t, c_void_p]),
    "wsb_sort3d_scratch_bytes": (c_size_t, [c_void_p, c_int]),
    "wsb_sort3d": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_moments3d_scratch_bytes": (c_size_t, [c_int]),
    "wsb_moments3d": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int, c_int,
                              c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_moments3d_sorted": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int,
                                     c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_hessian3d": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int, c_int,
                              c_void_p, c_void_p, c_size_t, c_void_p]),
    "wsb_moments3d_static_max_rows": (c_int64, []),
    "wsb_moments3d_static": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int,
                                     c_int, c_void_p, c_void_p]),
    "wsb_adjacency3d_host": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int64, c_void_p]),
    "wsb_winding3d_exact_host": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "wsb_face_terms3d_host": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "wsb_face_hessian3d_host": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "wsb_finalize22_host": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
}

# C enums
FORMS = {"plain": 0, "fp64": 1, "mixed": 2}
OUT_ACCUMULATE, OUT_PHYSICAL, OUT_WORLD, OUT_NO_INDICATOR = 1, 2, 4, 8
MESH_HOST_ONLY, MESH_SOUP = 1, 2

_lib = None


def load(path: str = LIB_PATH):
    """the loaded library (declared signatures); raises when it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load().wsb_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code: int):
    if code != 0:
        raise _ERRORS.get(code, RuntimeError)(last_error())


def form_id(form) -> int:
    if isinstance(form, int):
        return form
    try:
        return FORMS[str(form)]
    except KeyError:
        raise ValueError(f"unknown form {form!r}; expected one of {sorted(FORMS)}") from None


def require_device(*tensors):
    """the product path computes on the HIP device only."""
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError("warpsphboundaries_amd computes on the MI355X (HIP) device: got a CPU tensor; no CPU fallback exists")


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def nptr(a):
    """host pointer of a C-contiguous numpy array (host evaluation entry points)."""
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
