"""Device-resident objects of the library: uploaded plans and static 2-D geometries (one per device)."""
from __future__ import annotations

import ctypes
from functools import lru_cache

import numpy as np
import torch

from . import _lib as L
from .kernels import KERNEL_ID
from .plan import surface_plan, general_plan


class DevicePlan:
    """an uploaded flattened plan (the plain form's term table)."""

    def __init__(self, pl, kernel_id, device_index):
        a = pl.arrays()
        self.kernel_id = kernel_id
        self.device_index = device_index
        self.nE, self.nV = a["nE"], a["nV"]
        h = ctypes.c_void_p()
        with torch.cuda.device(device_index):
            L.check(L.load().wsb_plan_create(kernel_id, a["radii"].ctypes.data, len(a["radii"]), a["cn"].ctypes.data,
                                             a["cc"].ctypes.data, len(a["cc"]), a["E"].ctypes.data, a["nE"],
                                             a["V"].ctypes.data, a["mR"].ctypes.data, a["nV"], ctypes.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                L.load().wsb_plan_destroy(self.handle)
        except Exception:
            pass


@lru_cache(maxsize=None)
def surface_device_plan(kernel: str, device_index: int) -> DevicePlan:
    return DevicePlan(surface_plan(kernel), KERNEL_ID[kernel], device_index)


@lru_cache(maxsize=None)
def general_device_plan(kernel: str, alpha, grad: bool, device_index: int) -> DevicePlan:
    return DevicePlan(general_plan(kernel, tuple(alpha), grad), KERNEL_ID[kernel], device_index)


class Geom2D:
    """static device structures of one closed-loop surface for one support_max (cell list, cell indicator, y-bins)."""

    def __init__(self, vertices: np.ndarray, edges: np.ndarray, background: int, support_max: float, device_index: int):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        e = np.ascontiguousarray(edges, dtype=np.int32)
        self.device_index = device_index
        self.support_max = float(support_max)
        self.background = int(background)
        h = ctypes.c_void_p()
        K = ctypes.c_int()
        with torch.cuda.device(device_index):
            L.check(L.load().wsb_geom2d_create(v.ctypes.data, len(v), e.ctypes.data, len(e), int(background),
                                               float(support_max), ctypes.byref(h), ctypes.byref(K)))
        self.handle = h
        self.K = int(K.value)
        info = (ctypes.c_int64 * 10)()
        L.check(L.load().wsb_geom2d_info(self.handle, info))
        self.info = dict(K=info[0], nx=info[1], ny=info[2], nbins=info[3], ne=info[4], nv=info[5], cell_items=info[6],
                         device_bytes=info[7], local_winding=info[8], cells_without_reference=info[9])
        pv, pe = ctypes.c_void_p(), ctypes.c_void_p()
        L.check(L.load().wsb_geom2d_arrays(self.handle, ctypes.byref(pv), ctypes.byref(pe)))
        self.verts_ptr, self.edges_ptr = pv, pe
        self.ne, self.nv = len(e), len(v)

    def __del__(self):
        try:
            if self.handle:
                L.load().wsb_geom2d_destroy(self.handle)
        except Exception:
            pass

    # ---- hot path entry points (device tensors, current stream, no host synchronisation)
    def sort(self, lpos, order=None, scratch=None):
        """order [N] int32: the rows grouped by static cell (device radix sort of the cell keys)."""
        L.require_device(lpos)
        n = lpos.shape[0]
        dev = lpos.device
        lpos = lpos.to(torch.float64).contiguous()
        if order is None:
            order = torch.empty(n, dtype=torch.int32, device=dev)
        nbytes = int(L.load().wsb_sort2d_scratch_bytes(self.handle, n))
        if scratch is None or scratch.numel() < nbytes:
            scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        L.check(L.load().wsb_sort2d(self.handle, L.ptr(lpos), n, L.ptr(order), L.ptr(scratch), nbytes, L.stream_ptr(dev)))
        return order

    def count_order(self, counts, base=None, order=None, nwork=None, scratch=None):
        """(order [N] int32, nwork [1] int32): the rows grouped by slot count, largest first, and the number of rows
        with slots (``count_order``): the serving order of ``moments_multi``, one row per thread."""
        return count_order(counts, self.K, base, order, nwork, scratch)

    def adjacency(self, lpos, lsup, valid=None, pad=True, slots=None, counts=None, ind=None, order=None):
        """(slots [N*K] int32, counts [N] int32, ind [N] f64) -- the bit-exact fixed-capacity adjacency."""
        L.require_device(lpos, lsup)
        n = lpos.shape[0]
        dev = lpos.device
        lpos = lpos.to(torch.float64).contiguous()
        lsup = lsup.to(torch.float64).contiguous()
        if slots is None:
            slots = torch.empty(n * self.K, dtype=torch.int32, device=dev)      # the kernel writes every slot when pad
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=dev)
        if ind is None:
            ind = torch.empty(n, dtype=torch.float64, device=dev)
        v = None if valid is None else valid.to(torch.uint8).contiguous()
        L.check(L.load().wsb_adjacency2d(self.handle, L.ptr(lpos), L.ptr(lsup), L.ptr(v), n, L.ptr(order), L.ptr(slots), L.ptr(counts),
                                         L.ptr(ind), 1 if pad else 0, L.stream_ptr(dev)))
        return slots, counts, ind

    def moments(self, kernel, slots, counts, ind, lpos, lsup, form="fp64", rot4=None, flags=0, out=None, scratch=None, order=None):
        """[N, 12] twelve channels per row (device order), see include/wsb.h for the flags."""
        L.require_device(lpos, lsup, slots, counts, ind)
        n = lpos.shape[0]
        dev = lpos.device
        f = L.form_id(form)
        kid = KERNEL_ID[kernel]
        if out is None:
            out = torch.zeros((n, 12), dtype=torch.float64, device=dev) if flags & L.OUT_ACCUMULATE \
                else torch.empty((n, 12), dtype=torch.float64, device=dev)
        plan = None
        nbytes = 0
        if f == 0:
            plan = surface_device_plan(kernel, dev.index or 0).handle
            nbytes = int(L.load().wsb_moments2d_scratch_bytes(n, self.K))
            if scratch is None or scratch.numel() < nbytes:
                scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        lpos = lpos.to(torch.float64).contiguous()
        lsup = lsup.to(torch.float64).contiguous()
        L.check(L.load().wsb_moments2d(self.handle, plan, kid, f, L.ptr(order), L.ptr(slots), L.ptr(counts), self.K, L.ptr(ind), L.ptr(lpos),
                                       L.ptr(lsup), n, L.ptr(rot4), int(flags), L.ptr(out), L.ptr(scratch), nbytes,
                                       L.stream_ptr(dev)))
        return out

    def moments_multi(self, kernels, slots, counts, ind, lpos, lsup, form="fp64", rot4=None, flags=0, outs=None, order=None,
                      nwork=None):
        """{kernel: [N, 12]} for several of the scheme kernels (cubic, w2, w4, w6) in one pass: each pair's chord
        primitives are shared by the kernels (``wsb_moments2d_multi``); the same values as ``moments`` per kernel.
        ``order`` / ``nwork``: the serving order and its count of rows with slots (``count_order``)."""
        L.require_device(lpos, lsup, slots, counts, ind)
        n = lpos.shape[0]
        dev = lpos.device
        ptrs = (ctypes.c_void_p * 4)()
        res = {}
        for k in kernels:
            kid = KERNEL_ID[k]
            if kid > 3:
                raise NotImplementedError(f"{k}: several kernels in one pass are the scheme kernels cubic, w2, w4, w6")
            o = outs[k] if outs is not None and k in outs else (
                torch.zeros((n, 12), dtype=torch.float64, device=dev) if flags & L.OUT_ACCUMULATE
                else torch.empty((n, 12), dtype=torch.float64, device=dev))
            res[k] = o
            ptrs[kid] = o.data_ptr()
        lpos = lpos.to(torch.float64).contiguous()
        lsup = lsup.to(torch.float64).contiguous()
        if nwork is not None and (order is None or nwork.dtype != torch.int32 or nwork.numel() < 1):
            raise ValueError("nwork: the int32 device count of count_order, with the order it counts")
        L.check(L.load().wsb_moments2d_multi(self.handle, L.form_id(form), L.ptr(order), L.ptr(nwork), L.ptr(slots), L.ptr(counts), self.K,
                                             L.ptr(ind), L.ptr(lpos), L.ptr(lsup), n, L.ptr(rot4), int(flags), ptrs,
                                             L.stream_ptr(dev)))
        return res

    def adjacency_sorted(self, lpos, lsup, order, valid=None, pad=False, slots=None, counts=None, prow=None):
        """(slots [N*K] int32, counts [N] int32, prow [N, 4] f64) in the sorted-position layout: position i holds the
        adjacency of row order[i] (``order`` from ``sort``) and prow[i] = (x, y, h, indicator) of that row; the same
        values as ``adjacency`` (``wsb_adjacency2d_sorted``)."""
        L.require_device(lpos, lsup, order)
        n = lpos.shape[0]
        dev = lpos.device
        lpos = lpos.to(torch.float64).contiguous()
        lsup = lsup.to(torch.float64).contiguous()
        _require_buffer(order, "order", torch.int32, n, dev)
        if slots is None:
            slots = torch.empty(n * self.K, dtype=torch.int32, device=dev)
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=dev)
        if prow is None:
            prow = torch.empty((n, 4), dtype=torch.float64, device=dev)
        _require_buffer(slots, "slots", torch.int32, n * self.K, dev)
        _require_buffer(counts, "counts", torch.int32, n, dev)
        _require_buffer(prow, "prow", torch.float64, 4 * n, dev)
        v = None if valid is None else valid.to(torch.uint8).contiguous()
        L.check(L.load().wsb_adjacency2d_sorted(self.handle, L.ptr(lpos), L.ptr(lsup), L.ptr(v), n, L.ptr(order),
                                                L.ptr(slots), L.ptr(counts), L.ptr(prow), 1 if pad else 0,
                                                L.stream_ptr(dev)))
        return slots, counts, prow

    def moments_multi_sorted(self, kernels, order, slots, counts, prow, form="fp64", rot4=None, flags=0, outs=None):
        """{kernel: [N, 12]} in input order from the sorted-position layout of ``adjacency_sorted``: the values of
        ``moments_multi`` (``wsb_moments2d_multi_sorted``), each row's record read contiguously in cell order."""
        L.require_device(order, slots, counts, prow)
        n = order.shape[0]
        dev = order.device
        _require_buffer(slots, "slots", torch.int32, n * self.K, dev)
        _require_buffer(counts, "counts", torch.int32, n, dev)
        _require_buffer(prow, "prow", torch.float64, 4 * n, dev)
        ptrs = (ctypes.c_void_p * 4)()
        res = {}
        for k in kernels:
            kid = KERNEL_ID[k]
            if kid > 3:
                raise NotImplementedError(f"{k}: several kernels in one pass are the scheme kernels cubic, w2, w4, w6")
            o = outs[k] if outs is not None and k in outs else (
                torch.zeros((n, 12), dtype=torch.float64, device=dev) if flags & L.OUT_ACCUMULATE
                else torch.empty((n, 12), dtype=torch.float64, device=dev))
            _require_buffer(o, f"outs[{k}]", torch.float64, 12 * n, dev)
            res[k] = o
            ptrs[kid] = o.data_ptr()
        L.check(L.load().wsb_moments2d_multi_sorted(self.handle, L.form_id(form), L.ptr(order), L.ptr(slots),
                                                    L.ptr(counts), self.K, L.ptr(prow), n, L.ptr(rot4), int(flags), ptrs,
                                                    L.stream_ptr(dev)))
        return res

    def signed_distance(self, lpos, support_max=None):
        L.require_device(lpos)
        m = lpos.shape[0]
        dev = lpos.device
        lpos = lpos.to(torch.float64).contiguous()
        d = torch.empty(m, dtype=torch.float64, device=dev)
        nrm = torch.empty((m, 2), dtype=torch.float64, device=dev)
        edge = torch.empty(m, dtype=torch.int32, device=dev)
        L.check(L.load().wsb_signed_distance2d(self.handle, L.ptr(lpos), m, float(support_max or self.support_max), L.ptr(d),
                                               L.ptr(nrm), L.ptr(edge), L.stream_ptr(dev)))
        return d, nrm, edge


def _require_buffer(t, name, dtype, numel, dev):
    """a caller-supplied buffer the kernels write or read ``numel`` elements of: dtype, size, contiguity and device
    checked on the host, so that a wrong buffer is an error instead of an out-of-bounds device access."""
    if not torch.is_tensor(t):
        raise ValueError(f"{name}: expected a torch tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if t.numel() < numel:
        raise ValueError(f"{name}: {t.numel()} elements, at least {numel} needed")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if t.device != dev:
        raise ValueError(f"{name}: on {t.device}, the rows are on {dev}")


def _rows(lpos, lsup):
    if lpos.dim() != 2 or lpos.shape[1] != 2:
        raise ValueError(f"positions: expected [N, 2], got {tuple(lpos.shape)}")
    n = lpos.shape[0]
    if lsup.shape != (n,):
        raise ValueError(f"supports: expected [{n}], got {tuple(lsup.shape)}")
    if lsup.device != lpos.device:
        raise ValueError("positions and supports on different devices")
    return n


class Disks2D:
    """static device structures of a set of solid disks for one support_max (the cell list over the centres)."""

    def __init__(self, centres: np.ndarray, radii: np.ndarray, support_max: float, device_index: int):
        c = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 2)
        r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        if len(c) != len(r):
            raise ValueError("DiskArrayRep: centres and radii differ in length")
        self.device_index = device_index
        self.support_max = float(support_max)
        h = ctypes.c_void_p()
        K = ctypes.c_int()
        with torch.cuda.device(device_index):
            L.check(L.load().wsb_disks2d_create(c.ctypes.data, r.ctypes.data, len(r), float(support_max), ctypes.byref(h),
                                                ctypes.byref(K)))
        self.handle = h
        self.K = int(K.value)
        info = (ctypes.c_int64 * 5)()
        L.check(L.load().wsb_disks2d_info(self.handle, info))
        self.info = dict(K=info[0], nx=info[1], ny=info[2], m=info[3], device_bytes=info[4])

    def __del__(self):
        try:
            if self.handle:
                L.load().wsb_disks2d_destroy(self.handle)
        except Exception:
            pass

    def adjacency(self, lpos, lsup, valid=None, slots=None, counts=None):
        """(slots [N*K] int32 (-1 padded), counts [N] int32): the disks within one support plus radius of every row."""
        L.require_device(lpos, lsup)
        n = _rows(lpos, lsup)
        dev = lpos.device
        lpos = lpos.to(torch.float64).contiguous()
        lsup = lsup.to(torch.float64).contiguous()
        if slots is None:
            slots = torch.empty(n * self.K, dtype=torch.int32, device=dev)
        if counts is None:
            counts = torch.empty(n, dtype=torch.int32, device=dev)
        _require_buffer(slots, "slots", torch.int32, n * self.K, dev)
        _require_buffer(counts, "counts", torch.int32, n, dev)
        if valid is not None and (valid.numel() != n or valid.device != dev):
            raise ValueError(f"valid: expected [{n}] on {dev}")
        v = None if valid is None else valid.to(torch.uint8).contiguous()
        L.check(L.load().wsb_disk_adjacency2d(self.handle, L.ptr(lpos), L.ptr(lsup), L.ptr(v), n, L.ptr(slots),
                                              L.ptr(counts), L.stream_ptr(dev)))
        return slots, counts

    def moments(self, kernel, slots, counts, lpos, lsup, rot4=None, flags=0, out=None):
        """[N, 12] twelve channels per row (device order) of the row's disks, see include/wsb.h for the flags."""
        L.require_device(lpos, lsup, slots)
        n = _rows(lpos, lsup)
        dev = lpos.device
        if out is None:
            out = torch.zeros((n, 12), dtype=torch.float64, device=dev) if flags & L.OUT_ACCUMULATE \
                else torch.empty((n, 12), dtype=torch.float64, device=dev)
        _require_buffer(slots, "slots", torch.int32, n * self.K, dev)
        _require_buffer(counts, "counts", torch.int32, n, dev)
        _require_buffer(out, "out", torch.float64, n * 12, dev)
        if rot4 is not None:
            _require_buffer(rot4, "rot4", torch.float64, 4, dev)
        lpos = lpos.to(torch.float64).contiguous()
        lsup = lsup.to(torch.float64).contiguous()
        L.check(L.load().wsb_disk_moments2d(self.handle, KERNEL_ID[kernel], L.ptr(slots), L.ptr(counts), self.K,
                                            L.ptr(lpos), L.ptr(lsup), n, L.ptr(rot4), int(flags), L.ptr(out),
                                            L.stream_ptr(dev)))
        return out

    def moments_multi(self, kernels, slots, counts, lpos, lsup, rot4=None, flags=0, outs=None):
        """{kernel: [N, 12]} for several of the scheme kernels (cubic, w2, w4, w6) in one pass: every Gauss node's
        geometry is shared by the kernels (``wsb_disk_moments2d_multi``); the values of ``moments`` to rounding (with
        the cubic every kernel takes the cubic's panels, cut at r = h / 2 as well)."""
        L.require_device(lpos, lsup, slots)
        n = _rows(lpos, lsup)
        dev = lpos.device
        _require_buffer(slots, "slots", torch.int32, n * self.K, dev)
        _require_buffer(counts, "counts", torch.int32, n, dev)
        if rot4 is not None:
            _require_buffer(rot4, "rot4", torch.float64, 4, dev)
        ptrs = (ctypes.c_void_p * 4)()
        res = {}
        for k in kernels:
            kid = KERNEL_ID[k]
            if kid > 3:
                raise NotImplementedError(f"{k}: several kernels in one pass are the scheme kernels cubic, w2, w4, w6")
            o = outs[k] if outs is not None and k in outs else (
                torch.zeros((n, 12), dtype=torch.float64, device=dev) if flags & L.OUT_ACCUMULATE
                else torch.empty((n, 12), dtype=torch.float64, device=dev))
            _require_buffer(o, f"outs[{k!r}]", torch.float64, n * 12, dev)
            res[k] = o
            ptrs[kid] = o.data_ptr()
        lpos = lpos.to(torch.float64).contiguous()
        lsup = lsup.to(torch.float64).contiguous()
        L.check(L.load().wsb_disk_moments2d_multi(self.handle, L.ptr(slots), L.ptr(counts), self.K, L.ptr(lpos),
                                                  L.ptr(lsup), n, L.ptr(rot4), int(flags), ptrs, L.stream_ptr(dev)))
        return res


def count_order(counts, K, base=None, order=None, nwork=None, scratch=None):
    """(order [N] int32, nwork [1] int32): the rows (of ``base`` [N] when given, else 0..N-1) grouped by slot count
    (counts clamped to K), largest first, and the number of rows with at least one slot (``wsb_count_order2d``)."""
    L.require_device(counts)
    n = counts.shape[0]
    dev = counts.device
    if counts.dtype != torch.int32 or not counts.is_contiguous():
        raise ValueError("counts: a contiguous int32 device tensor")
    if base is not None and (base.dtype != torch.int32 or base.numel() != n or base.device != dev):
        raise ValueError("base: an int32 permutation of the N rows on the device of counts")
    if order is None:
        order = torch.empty(n, dtype=torch.int32, device=dev)
    if nwork is None:
        nwork = torch.empty(1, dtype=torch.int32, device=dev)
    nbytes = int(L.load().wsb_count_order2d_scratch_bytes())
    if scratch is None or scratch.numel() < nbytes:
        scratch = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    L.check(L.load().wsb_count_order2d(L.ptr(counts), L.ptr(base), n, int(K), L.ptr(order), L.ptr(nwork), L.ptr(scratch),
                                       nbytes, L.stream_ptr(dev)))
    return order, nwork


def disk_channels_host(kernel, D, R):
    """[n, 7] (lam, m1, g0 along the centre direction; g1, m2 along / across it) of single disks, units of h, computed
    on the host by the library's own arithmetic (validation)."""
    D = np.ascontiguousarray(D, dtype=np.float64).reshape(-1)
    R = np.ascontiguousarray(np.broadcast_to(R, D.shape), dtype=np.float64)
    out = np.zeros((len(D), 7))
    L.check(L.load().wsb_disk_channels_host(KERNEL_ID[kernel], L.nptr(D), L.nptr(R), len(D), L.nptr(out)))
    return out


def edge_channels12(pair_q, pair_e, positions, supports, vertices, edges, kernel, form="fp64"):
    """[P, 12] edge-local channels (units of h) of an explicit pair list on the device."""
    L.require_device(pair_q, pair_e, positions, supports, vertices, edges)
    dev = positions.device
    P = pair_q.shape[0]
    out = torch.empty((P, 12), dtype=torch.float64, device=dev)
    if P == 0:
        return out
    f = L.form_id(form)
    plan = surface_device_plan(kernel, dev.index or 0).handle if f == 0 else None
    pq = pair_q.to(torch.int32).contiguous()
    pe = pair_e.to(torch.int32).contiguous()
    pos = positions.to(torch.float64).contiguous()
    sup = supports.to(torch.float64).contiguous()
    V = vertices.to(torch.float64).contiguous()
    E = edges.to(torch.int32).contiguous()
    L.check(L.load().wsb_edge_channels2d(plan, KERNEL_ID[kernel], f, L.ptr(pq), L.ptr(pe), P, L.ptr(pos), L.ptr(sup),
                                         pos.shape[0], L.ptr(V), L.ptr(E), E.shape[0], L.ptr(out), L.stream_ptr(dev)))
    return out


def plan_channels(pair_q, pair_e, positions, supports, vertices, edges, kernel, alpha, grad=False):
    """edge-local m_alpha (channel 0) or grad_x m_alpha (channels 0, 1) of arbitrary alpha through the plan
    interpreter: [P, 12] with the unused channels zero."""
    L.require_device(pair_q, pair_e, positions, supports, vertices, edges)
    dev = positions.device
    P = pair_q.shape[0]
    out = torch.empty((P, 12), dtype=torch.float64, device=dev)
    if P == 0:
        return out
    plan = general_device_plan(kernel, tuple(int(a) for a in alpha), bool(grad), dev.index or 0).handle
    pq = pair_q.to(torch.int32).contiguous()
    pe = pair_e.to(torch.int32).contiguous()
    pos = positions.to(torch.float64).contiguous()
    sup = supports.to(torch.float64).contiguous()
    V = vertices.to(torch.float64).contiguous()
    E = edges.to(torch.int32).contiguous()
    L.check(L.load().wsb_edge_channels2d(plan, KERNEL_ID[kernel], 0, L.ptr(pq), L.ptr(pe), P, L.ptr(pos), L.ptr(sup),
                                         pos.shape[0], L.ptr(V), L.ptr(E), E.shape[0], L.ptr(out), L.stream_ptr(dev)))
    return out
