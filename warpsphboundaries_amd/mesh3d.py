"""3-D closed triangle surfaces on the device: the reference's ``fluxbridge`` evaluator surface for the stored moments,
and the ``mesh3d`` element signatures, over the HIP library.

* ``Mesh3D``        a resident closed surface (``wsb_mesh3d_create``): face records and the static grid on the device.
* ``Geometry``      ``Geometry.closed_surface(vertices, faces)`` and tetrahedra (reference ``fluxbridge/geometry.py``).
* ``Evaluator``     ``Evaluator(geometry, requests, kernel=, precision=).evaluate(queries, supports)`` -> numpy
                    ``[Q, len(requests)]`` in physical units (reference ``fluxbridge/evaluate.py:11-117``), for requests
                    of ``baseline(3)[:22]``.
* ``ClosedSurface``, ``Tetrahedron``, ``TetrahedralMesh``: ``moment(name, alpha, h)``, ``kernel_gradient``,
                    ``moment_gradient`` (reference ``mesh3d/reference.py:219-412``) for |alpha| <= 2 (moments) and
                    |alpha| <= 1 (gradients).

Every evaluation runs on the HIP device; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from numbers import Integral
from typing import Sequence

import numpy as np
import torch

from . import _lib as L
from .kernels import KERNEL_ID
from .plan3d import CHANNEL_NAMES3, NCH3, REQUESTS22

__all__ = ["Mesh3D", "Adjacency3D", "Request", "baseline", "monomials", "Geometry", "Evaluator", "ClosedSurface",
           "Tetrahedron", "TetrahedralMesh", "REQUESTS22", "CHANNEL_NAMES3", "icosphere", "box_surface"]

_PRECISION = {"float64": "fp64", "fp64": "fp64", "mixed": "mixed"}


def _kernel_name(kernel) -> str:
    name = getattr(kernel, "name", kernel)
    name = {"Wendland2": "w2", "Wendland4": "w4", "Wendland6": "w6", "CubicSpline": "cubic"}.get(name, name)
    if name not in KERNEL_ID:
        raise ValueError(f"unknown kernel {kernel!r}")
    return name


@dataclass
class Adjacency3D:
    """CSR face adjacency of a particle snapshot: ``offsets[n+1]`` (int64), ``faces[offsets[n]]`` (int32, ascending per
    particle), ``chi[n]`` winding numbers; device tensors.  The sorted-position layout (``prow`` set): entry i belongs
    to particle ``order[i]`` (cell order) and ``prow[i]`` holds its (x, y, z, h)."""
    offsets: torch.Tensor
    faces: torch.Tensor
    chi: torch.Tensor
    count: int
    order: torch.Tensor | None = None
    prow: torch.Tensor | None = None

    def numpy(self):
        """(query, element, offsets) as numpy arrays, the layout of the reference's Adjacency."""
        off = self.offsets.cpu().numpy()
        el = self.faces[: self.count].cpu().numpy()
        q = np.repeat(np.arange(len(off) - 1, dtype=np.int32), np.diff(off))
        if self.prow is not None:
            # positions -> particles, ascending by particle (a stable sort keeps each particle's ascending faces)
            order = self.order.cpu().numpy()
            q = order[q].astype(np.int32)
            k = np.argsort(q, kind="stable")
            q, el = q[k], el[k]
            off = np.zeros(len(order) + 1, dtype=np.int64)
            np.cumsum(np.bincount(q, minlength=len(order)), out=off[1:])
        return q, el, off

    def by_particle(self):
        """(offsets, faces, chi) in the particle (row) layout, whatever the layout of this adjacency (device tensors)"""
        if self.prow is None:
            return self.offsets, self.faces[: self.count], self.chi
        o = self.order.long()
        n = o.numel()
        cnt_pos = self.offsets[1:] - self.offsets[:-1]
        cnt = torch.empty_like(cnt_pos)
        cnt[o] = cnt_pos
        off = torch.zeros(n + 1, dtype=torch.int64, device=cnt.device)
        torch.cumsum(cnt, 0, out=off[1:])
        p = torch.repeat_interleave(torch.arange(n, device=o.device), cnt_pos)        # the position of every slot
        rank = torch.arange(self.count, device=o.device) - self.offsets[p]
        faces = torch.empty(self.count, dtype=self.faces.dtype, device=o.device)
        faces[off[o[p]] + rank] = self.faces[: self.count]
        chi = torch.empty_like(self.chi)
        chi[o] = self.chi
        return off, faces, chi


class Mesh3D:
    """A closed, consistently oriented triangle surface (outward normals; cavities point into the void), resident on
    the device with its static grid for ``support_max``.  ``soup=True`` accepts any triangle set (adjacency only)."""

    def __init__(self, vertices, faces, support_max: float, device="cuda:0", host_only: bool = False, soup: bool = False):
        V = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
        Fc = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1, 3))
        self.vertices, self.faces = V, Fc
        self.support_max = float(support_max)
        self.host_only = bool(host_only)
        self.soup = bool(soup)
        self.device = torch.device(device) if not host_only else None
        lib = L.load()
        h = ctypes.c_void_p()
        info = (ctypes.c_int64 * 10)()
        flags = (L.MESH_HOST_ONLY if host_only else 0) | (L.MESH_SOUP if soup else 0)
        if not host_only:
            torch.cuda.set_device(self.device)
        L.check(lib.wsb_mesh3d_create(L.nptr(V), len(V), L.nptr(Fc), len(Fc), self.support_max, flags, ctypes.byref(h),
                                      info))
        self._h = h
        (self.K, nx, ny, nz, self.nf, self.items, self.bytes, self.components, self.dirty,
         self.longest_walk) = (int(v) for v in info)
        self.grid_shape = (nx, ny, nz)
        g = (ctypes.c_double * 4)()
        L.check(lib.wsb_mesh3d_grid(h, g))
        self.grid_lo, self.cell = tuple(g[:3]), g[3]
        self._scratch = None
        self._sort_scratch = None
        self._mom_scratch = None
        self._cap = 0
        self.search = "grid"         # the default face search of adjacency()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.load().wsb_mesh3d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------------------------------------ device path
    def _check_inputs(self, pos, sup):
        if self.host_only:
            raise ValueError("host-only mesh: use the *_host methods")
        L.require_device(pos, sup)
        if pos.dim() != 2 or pos.shape[1] != 3:
            raise ValueError("positions must be [N, 3]")
        if sup.shape != (pos.shape[0],):
            raise ValueError("supports must be [N]")
        if pos.dtype != torch.float64 or sup.dtype != torch.float64:
            raise ValueError("positions and supports must be float64")

    def scratch_bytes(self, n: int) -> int:
        return int(L.load().wsb_adjacency3d_scratch_bytes(n))

    def _buf(self, name, nbytes, dev):
        b = getattr(self, name)
        if b is None or b.numel() < nbytes or b.device != dev:
            b = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)
            setattr(self, name, b)
        return b

    def sort(self, pos: torch.Tensor, order: torch.Tensor | None = None) -> torch.Tensor:
        """particles grouped by grid cell (device counting sort): a permutation for ``adjacency(order=...)``."""
        self._check_inputs(pos, pos[:, 0])
        n = pos.shape[0]
        if order is None:
            order = torch.empty(n, dtype=torch.int32, device=pos.device)
        lib = L.load()
        scratch = self._buf("_sort_scratch", lib.wsb_sort3d_scratch_bytes(self._h, n), pos.device)
        L.check(lib.wsb_sort3d(self._h, L.ptr(pos), n, L.ptr(order), L.ptr(scratch), scratch.numel(), L.stream_ptr(pos.device)))
        return order

    def adjacency(self, pos: torch.Tensor, sup: torch.Tensor, capacity: int | None = None, scratch=None,
                  order: torch.Tensor | None = None, sort: bool = True, search: str | None = None,
                  layout: str = "rows") -> Adjacency3D:
        """face adjacency + winding numbers.  Particles are grouped by cell first (``sort``; or pass ``order``), which
        changes only the thread schedule.  ``capacity`` (pairs) defaults to a remembered bound; the call retries once
        with the exact count when it was too small (one host read of the pair count).  ``search``: "grid" (the cell
        lists) or "bvh" (the face tree); both give the same face lists and winding numbers.  ``layout="sorted"``
        (cell lists, sorted particles): the entries at the positions of the cell order (``wsb_adjacency3d_sorted``),
        which ``moments`` reads contiguously."""
        search = self.search if search is None else search
        if search not in ("grid", "bvh"):
            raise ValueError(f"search {search!r}: expected 'grid' or 'bvh'")
        if layout not in ("rows", "sorted"):
            raise ValueError(f"layout {layout!r}: expected 'rows' or 'sorted'")
        if layout == "sorted" and (search != "grid" or not (sort or order is not None)):
            raise ValueError("the sorted layout takes the cell-list search and a cell order")
        self._check_inputs(pos, sup)
        n = pos.shape[0]
        dev = pos.device
        lib = L.load()
        if order is None and sort and n:
            order = self.sort(pos)
        if scratch is None:
            scratch = self._buf("_scratch", self.scratch_bytes(n), dev)
        offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
        chi = torch.empty(n, dtype=torch.float64, device=dev)
        cap = int(capacity if capacity is not None else max(self._cap, 4 * n, 1024))
        faces = torch.empty(cap, dtype=torch.int32, device=dev)
        st = L.stream_ptr(dev)
        prow = torch.empty((n, 4), dtype=torch.float64, device=dev) if layout == "sorted" else None
        if prow is not None:
            def run():
                L.check(lib.wsb_adjacency3d_sorted(self._h, L.ptr(pos), L.ptr(sup), L.ptr(order), n, L.ptr(offsets),
                                                   L.ptr(faces), cap, L.ptr(chi), L.ptr(prow), L.ptr(scratch),
                                                   scratch.numel(), st))
        else:
            fn = lib.wsb_adjacency3d if search == "grid" else lib.wsb_adjacency3d_bvh

            def run():
                L.check(fn(self._h, L.ptr(pos), L.ptr(sup), L.ptr(order), n, L.ptr(offsets), L.ptr(faces), cap,
                           L.ptr(chi), L.ptr(scratch), scratch.numel(), st))
        run()
        total = int(offsets[n]) if n else 0
        if total < 0:           # the device's check of the supports
            raise ValueError("support exceeds the mesh's support_max")
        if total > cap:
            cap = total
            faces = torch.empty(cap, dtype=torch.int32, device=dev)
            run()
        self._cap = max(self._cap, total)
        return Adjacency3D(offsets, faces, chi, total, order if prow is not None else None, prow)

    def moments(self, adj: Adjacency3D, pos: torch.Tensor, sup: torch.Tensor, kernel="w2", form="fp64",
                physical: bool = True, indicator: bool = True, out: torch.Tensor | None = None) -> torch.Tensor:
        """[N, 22] stored moments (request order of baseline(3)[:22])."""
        self._check_inputs(pos, sup)
        n = pos.shape[0]
        if out is None:
            out = torch.empty((n, NCH3), dtype=torch.float64, device=pos.device)
        flags = (L.OUT_PHYSICAL if physical else 0) | (0 if indicator else L.OUT_NO_INDICATOR)
        fid = L.form_id(form)
        if fid == 0:
            raise NotImplementedError("the 3-D path provides the fp64 and mixed forms")
        lib = L.load()
        scratch = self._buf("_mom_scratch", lib.wsb_moments3d_scratch_bytes(n), pos.device)
        if getattr(adj, "prow", None) is not None:
            if adj.prow.shape[0] != n:
                raise ValueError("the adjacency's particle count differs from the positions'")
            L.check(lib.wsb_moments3d_sorted(self._h, KERNEL_ID[_kernel_name(kernel)], fid, L.ptr(adj.order),
                                             L.ptr(adj.prow), L.ptr(adj.offsets), L.ptr(adj.faces), adj.faces.numel(),
                                             L.ptr(adj.chi), n, flags, L.ptr(out), L.ptr(scratch), scratch.numel(),
                                             L.stream_ptr(pos.device)))
            return out
        L.check(lib.wsb_moments3d(self._h, KERNEL_ID[_kernel_name(kernel)], fid, L.ptr(adj.offsets), L.ptr(adj.faces),
                                  adj.faces.numel(), L.ptr(adj.chi), L.ptr(pos), L.ptr(sup), n, flags, L.ptr(out),
                                  L.ptr(scratch), scratch.numel(), L.stream_ptr(pos.device)))
        return out

    def hessian(self, adj: Adjacency3D, pos: torch.Tensor, sup: torch.Tensor, kernel="w2", form="fp64",
                physical: bool = True, out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
        """[N, 6] Hessian entries int_V d_i d_j W (i <= j: xx, xy, xz, yy, yz, zz), the requests baseline(3)[22:28]."""
        self._check_inputs(pos, sup)
        n = pos.shape[0]
        if out is None:
            out = torch.zeros((n, 6), dtype=torch.float64, device=pos.device)
        elif out.shape != (n, 6) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 [N, 6] tensor")
        flags = (L.OUT_PHYSICAL if physical else 0) | (L.OUT_ACCUMULATE if accumulate else 0)
        fid = L.form_id(form)
        if fid == 0:
            raise NotImplementedError("the 3-D path provides the fp64 and mixed forms")
        lib = L.load()
        scratch = self._buf("_mom_scratch", lib.wsb_moments3d_scratch_bytes(n), pos.device)
        if getattr(adj, "prow", None) is not None:
            raise NotImplementedError("the Hessian entries take the particle (rows) layout of the adjacency")
        L.check(lib.wsb_hessian3d(self._h, KERNEL_ID[_kernel_name(kernel)], fid, L.ptr(adj.offsets), L.ptr(adj.faces),
                                  adj.faces.numel(), L.ptr(pos), L.ptr(sup), n, flags, L.ptr(out), L.ptr(scratch),
                                  scratch.numel(), L.stream_ptr(pos.device)))
        return out

    def evaluate(self, pos, sup, kernel="w2", form="fp64", physical=True, hessian=False):
        """[N, 22] stored moments, [N, 28] with the Hessian entries (all of baseline(3))."""
        adj = self.adjacency(pos, sup)
        m = self.moments(adj, pos, sup, kernel, form, physical)
        if not hessian:
            return m
        return torch.cat([m, self.hessian(adj, pos, sup, kernel, form, physical)], 1)

    # ------------------------------------------------------------------------------------------ host evaluation
    def adjacency_host(self, pos, sup):
        """the same adjacency and winding arithmetic on the host (numpy in, numpy out)."""
        P = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        H = np.ascontiguousarray(np.broadcast_to(np.asarray(sup, dtype=np.float64), (len(P),)))
        off = np.zeros(len(P) + 1, dtype=np.int64)
        chi = np.zeros(len(P))
        lib = L.load()
        L.check(lib.wsb_adjacency3d_host(self._h, L.nptr(P), L.nptr(H), len(P), L.nptr(off), None, 0, L.nptr(chi)))
        faces = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
        L.check(lib.wsb_adjacency3d_host(self._h, L.nptr(P), L.nptr(H), len(P), L.nptr(off), L.nptr(faces), len(faces),
                                         L.nptr(chi)))
        return off, faces[: off[-1]], chi

    def winding_exact_host(self, pos):
        P = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        chi = np.zeros(len(P))
        L.check(L.load().wsb_winding3d_exact_host(self._h, L.nptr(P), len(P), L.nptr(chi)))
        return chi

    def face_terms_host(self, kernel, pair_face, pair_pos, pair_sup, form="fp64"):
        f = np.ascontiguousarray(np.asarray(pair_face, dtype=np.int32))
        P = np.ascontiguousarray(np.asarray(pair_pos, dtype=np.float64).reshape(-1, 3))
        H = np.ascontiguousarray(np.broadcast_to(np.asarray(pair_sup, dtype=np.float64), (len(f),)))
        out = np.zeros((len(f), 23))
        L.check(L.load().wsb_face_terms3d_host(self._h, KERNEL_ID[_kernel_name(kernel)], L.form_id(form), L.nptr(f),
                                               L.nptr(P), L.nptr(H), len(f), L.nptr(out)))
        return out

    def face_hessian_host(self, kernel, pair_face, pair_pos, pair_sup, form="fp64"):
        f = np.ascontiguousarray(np.asarray(pair_face, dtype=np.int32))
        P = np.ascontiguousarray(np.asarray(pair_pos, dtype=np.float64).reshape(-1, 3))
        H = np.ascontiguousarray(np.broadcast_to(np.asarray(pair_sup, dtype=np.float64), (len(f),)))
        out = np.zeros((len(f), 6))
        L.check(L.load().wsb_face_hessian3d_host(self._h, KERNEL_ID[_kernel_name(kernel)], L.form_id(form), L.nptr(f),
                                                 L.nptr(P), L.nptr(H), len(f), L.nptr(out)))
        return out

    def hessian_host(self, pos, sup, kernel="w2", physical=True, form="fp64"):
        """the device Hessian arithmetic run on the host: [N, 6]."""
        P = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        H = np.ascontiguousarray(np.broadcast_to(np.asarray(sup, dtype=np.float64), (len(P),)))
        off, faces, _ = self.adjacency_host(P, H)
        q = np.repeat(np.arange(len(P)), np.diff(off))
        acc = np.zeros((len(P), 6))
        if len(faces):
            np.add.at(acc, q, self.face_hessian_host(kernel, faces, P[q], H[q], form))
        acc /= np.pi
        return acc / (H * H)[:, None] if physical else acc

    def evaluate_host(self, pos, sup, kernel="w2", physical=True, form="fp64"):
        """the device arithmetic run on the host: [N, 22]."""
        P = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        H = np.ascontiguousarray(np.broadcast_to(np.asarray(sup, dtype=np.float64), (len(P),)))
        off, faces, chi = self.adjacency_host(P, H)
        q = np.repeat(np.arange(len(P)), np.diff(off))
        terms = self.face_terms_host(kernel, faces, P[q], H[q], form) if len(faces) else np.zeros((0, 23))
        acc = np.zeros((len(P), 23))
        np.add.at(acc, q, terms)
        out = np.zeros((len(P), NCH3))
        acc = np.ascontiguousarray(acc)
        L.check(L.load().wsb_finalize22_host(KERNEL_ID[_kernel_name(kernel)], L.nptr(acc), L.nptr(chi), L.nptr(H), len(P),
                                             L.OUT_PHYSICAL if physical else 0, L.nptr(out)))
        return out


# ---------------------------------------------------------------------------------------------- fluxbridge surface
@dataclass(frozen=True)
class Request:
    """one kernel-integral request (reference fluxbridge/contract.py:10-30)."""
    alpha: tuple
    derivative: tuple = ()

    def __post_init__(self):
        object.__setattr__(self, "alpha", tuple(self.alpha))
        object.__setattr__(self, "derivative", tuple(self.derivative))
        d = len(self.alpha)
        if d not in (2, 3) or any(not isinstance(a, Integral) or a < 0 for a in self.alpha):
            raise ValueError("alpha must contain two or three nonnegative integers")
        if sum(self.alpha) > 9:
            raise ValueError("polynomial degree exceeds nine")
        if len(self.derivative) > 2 or any(not isinstance(j, Integral) or j not in range(d) for j in self.derivative):
            raise ValueError("kernel derivatives must have order 0..2 and valid axes")

    @property
    def scale(self):
        return sum(self.alpha) - len(self.derivative)


def monomials(dimension, degree):
    """graded monomial order (reference fluxbridge/contract.py:29-38)."""
    if dimension not in (2, 3) or not isinstance(degree, Integral) or degree not in range(10):
        raise ValueError("dimension must be 2 or 3 and degree 0..9")

    def lattice(d, n):
        if d == 1:
            return [(n,)]
        return [(i,) + b for i in range(n, -1, -1) for b in lattice(d - 1, n - i)]
    return tuple(a for n in range(degree + 1) for a in lattice(dimension, n))


def baseline(dimension):
    """stored baseline moments, first gradient moments, symmetric Hessian (reference fluxbridge/contract.py:41-46)."""
    z = (0,) * dimension
    return tuple(Request(a) for a in monomials(dimension, 2)) + tuple(
        Request(a, (j,)) for a in monomials(dimension, 1) for j in range(dimension)) + tuple(
        Request(z, (i, j)) for i in range(dimension) for j in range(i, dimension))


_COLUMN = {Request(a, d): k for k, (a, d) in enumerate(REQUESTS22)}
# the Hessian requests of baseline(3), columns 22..27 of a 28-column evaluation
_HESSIAN = tuple((i, j) for i in range(3) for j in range(i, 3))
_COLUMN.update({Request((0, 0, 0), d): NCH3 + k for k, d in enumerate(_HESSIAN)})


def _tet_faces(V):
    faces = []
    for i in range(4):
        f = [j for j in range(4) if j != i]
        a, b, c = V[f]
        side = np.dot(np.cross(b - a, c - a), V[i] - a)
        if side == 0:
            raise ValueError("degenerate tetrahedron")
        if side > 0:
            f[1], f[2] = f[2], f[1]
        faces.append(f)
    return faces


class Geometry:
    """closed-surface geometry (reference fluxbridge/geometry.py:21-81): ``Geometry.closed_surface(vertices, faces)``,
    or ``Geometry(tetrahedra, 3)`` for a set of tetrahedra (each one a closed element; their moments add)."""

    def __init__(self, elements, dimension=3, measure="volume"):
        if dimension != 3 or measure != "volume":
            raise NotImplementedError("the device 3-D path evaluates closed 3-D volumes (tetrahedra or closed surfaces)")
        self.dimension, self.measure = 3, "volume"
        self.parts = []
        for v in elements:
            V = np.array(v, dtype=np.float64)
            if V.shape != (4, 3) or not np.isfinite(V).all():
                raise ValueError("volume elements must be tetrahedra; use closed_surface for a skin")
            self.parts.append((V, np.array(_tet_faces(V), dtype=np.int32)))

    @classmethod
    def closed_surface(cls, vertices, faces):
        obj = cls.__new__(cls)
        obj.dimension, obj.measure = 3, "volume"
        obj.parts = [(np.asarray(vertices, dtype=np.float64).reshape(-1, 3), np.asarray(faces, dtype=np.int32).reshape(-1, 3))]
        return obj


class Evaluator:
    """stored-moment evaluator (reference fluxbridge/evaluate.py:11-117 with compact=True): ``precision`` "float64"
    (the FP64 closed form) or "mixed" (FP32 hardware angles and square-root seeds)."""

    def __init__(self, geometry: Geometry, requests: Sequence[Request] = None, *, kernel="w2", precision="float64",
                 device="cuda:0", compact=True):
        requests = tuple(baseline(3) if requests is None else requests)
        for r in requests:
            if r not in _COLUMN:
                raise NotImplementedError(f"request {r} is outside baseline(3) (stored moments, first gradient "
                                          "moments, the Hessian at alpha = 0)")
        if precision not in _PRECISION:
            raise NotImplementedError(f"precision {precision!r}: expected float64 or mixed")
        self.geometry, self.requests = geometry, requests
        self.kernel = _kernel_name(kernel)
        self.form = _PRECISION[precision]
        self.device = torch.device(device)
        self.columns = [_COLUMN[r] for r in requests]
        self.hessian = any(c >= NCH3 for c in self.columns)
        self._meshes = {}

    def _mesh(self, part, smax):
        key = id(part)
        m = self._meshes.get(key)
        if m is None or m.support_max < smax:
            m = Mesh3D(part[0], part[1], smax, device=self.device)
            self._meshes[key] = m
        return m

    def evaluate_device(self, queries: torch.Tensor, supports: torch.Tensor) -> torch.Tensor:
        n = queries.shape[0]
        total = torch.zeros((n, NCH3 + (6 if self.hessian else 0)), dtype=torch.float64, device=queries.device)
        if n == 0:
            return total
        smax = float(supports.max())
        for part in self.geometry.parts:
            m = self._mesh(part, smax)
            adj = m.adjacency(queries, supports)
            total[:, :NCH3] += m.moments(adj, queries, supports, self.kernel, self.form, physical=True)
            if self.hessian:
                total[:, NCH3:] += m.hessian(adj, queries, supports, self.kernel, self.form, physical=True)
        return total

    def evaluate(self, queries, supports):
        q = np.asarray(queries, dtype=np.float64)
        if q.size == 0:
            return np.zeros((0, len(self.requests)))
        q = q.reshape(-1, 3)
        h = np.broadcast_to(np.asarray(supports, dtype=np.float64), (len(q),)).copy()
        if not np.isfinite(q).all() or not np.isfinite(h).all() or np.any(h <= 0):
            raise ValueError("finite queries and positive supports required")
        Q = torch.as_tensor(q, device=self.device)
        H = torch.as_tensor(h, device=self.device)
        out = self.evaluate_device(Q, H).cpu().numpy()
        return out[:, self.columns]


# ---------------------------------------------------------------------------------------------- mesh3d surface
def _alpha(alpha):
    alpha = tuple(alpha)
    if len(alpha) != 3 or any(not isinstance(i, Integral) or i < 0 for i in alpha):
        raise ValueError("alpha must be three nonnegative integers")
    return tuple(int(i) for i in alpha)


class ClosedSurface:
    """reference mesh3d/reference.py:219-325 signatures on the device: moment / moment_gradient / kernel_gradient for
    the stored-moment degrees (|alpha| <= 2 and |alpha| <= 1), kernel_hessian and laplacian at alpha = 0."""

    def __init__(self, vertices, faces, query=(0, 0, 0), device="cuda:0"):
        self.vertices = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
        self.indices = np.asarray(faces, dtype=np.int32).reshape(-1, 3)
        self.query = np.asarray(query, dtype=np.float64).reshape(3)
        self.device = device
        self._cache = {}

    def _row(self, name, h):
        key = (_kernel_name(name), float(h))
        v = self._cache.get(key)
        if v is None:
            m = Mesh3D(self.vertices, self.indices, float(h), device=self.device)
            q = torch.as_tensor(self.query[None], device=m.device)
            H = torch.full((1,), float(h), dtype=torch.float64, device=m.device)
            v = m.evaluate(q, H, key[0], "fp64", physical=True, hessian=True)[0].cpu().numpy()
            self._cache[key] = v
        return v

    def kernel_hessian(self, name="w2", alpha=(0, 0, 0), h=1):
        """int_V y^alpha Hessian_query W_h (mesh3d/reference.py:296-305) at alpha = 0, the 3 x 3 symmetric matrix."""
        if any(_alpha(alpha)):
            raise NotImplementedError("device kernel Hessians cover alpha = 0")
        row = self._row(name, h)
        H = [[0.0] * 3 for _ in range(3)]
        for k, (i, j) in enumerate(_HESSIAN):
            H[i][j] = H[j][i] = float(row[NCH3 + k])
        return tuple(tuple(r) for r in H)

    def laplacian(self, name="w2", alpha=(0, 0, 0), h=1):
        """int_V y^alpha Delta W_h (mesh3d/reference.py:288-294) at alpha = 0: the trace of the kernel Hessian."""
        if any(_alpha(alpha)):
            raise NotImplementedError("device Laplacians cover alpha = 0")
        H = self.kernel_hessian(name, alpha, h)
        return H[0][0] + H[1][1] + H[2][2]

    def moment(self, name="w2", alpha=(0, 0, 0), h=1):
        alpha = _alpha(alpha)
        k = _COLUMN.get(Request(alpha))
        if k is None:
            raise NotImplementedError("device moments cover |alpha| <= 2")
        return float(self._row(name, h)[k])

    def kernel_gradient(self, name="w2", alpha=(0, 0, 0), h=1):
        alpha = _alpha(alpha)
        cols = [_COLUMN.get(Request(alpha, (j,))) for j in range(3)]
        if None in cols:
            raise NotImplementedError("device kernel gradients cover |alpha| <= 1")
        row = self._row(name, h)
        return tuple(float(row[c]) for c in cols)

    def moment_gradient(self, name="w2", alpha=(0, 0, 0), h=1):
        """derivative of the complete translated moment (kernel gradient minus alpha_j M(alpha - e_j))."""
        alpha = _alpha(alpha)
        g = self.kernel_gradient(name, alpha, h)
        out = []
        for j in range(3):
            v = g[j]
            if alpha[j]:
                lower = tuple(a - int(i == j) for i, a in enumerate(alpha))
                v -= alpha[j] * self.moment(name, lower, h)
            out.append(v)
        return tuple(out)


class Tetrahedron(ClosedSurface):
    """single tetrahedron; either input orientation is accepted (mesh3d/reference.py:328-347)."""

    def __init__(self, vertices, query=(0, 0, 0), device="cuda:0"):
        V = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
        if len(V) != 4:
            raise ValueError("tetrahedron needs four vertices")
        super().__init__(V, _tet_faces(V), query, device)


class TetrahedralMesh:
    """additive tetrahedral volume (mesh3d/reference.py:350-412): the moments of its boundary surface."""

    def __init__(self, vertices, cells, query=(0, 0, 0), device="cuda:0"):
        V = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
        cells = [tuple(int(i) for i in c) for c in cells]
        if not cells:
            raise ValueError("volume mesh needs tetrahedra")
        face_map = {}
        for c in cells:
            if len(c) != 4 or len(set(c)) != 4 or any(i < 0 or i >= len(V) for i in c):
                raise ValueError("invalid tetrahedron indices")
            for f in _tet_faces(V[list(c)]):
                ids = tuple(c[i] for i in f)
                face_map.setdefault(tuple(sorted(ids)), []).append(ids)
        boundary = []
        for fs in face_map.values():
            if len(fs) == 1:
                boundary.append(fs[0])
            elif len(fs) == 2:
                a, b = fs
                if any(a == b[k:] + b[:k] for k in range(3)):
                    raise ValueError("shared tetrahedral face has inconsistent solid orientation")
            else:
                raise ValueError("nonmanifold tetrahedral face")
        self._boundary = ClosedSurface(V, boundary, query, device)

    def boundary(self):
        return self._boundary

    def moment(self, name="w2", alpha=(0, 0, 0), h=1):
        return self._boundary.moment(name, alpha, h)

    def kernel_gradient(self, name="w2", alpha=(0, 0, 0), h=1):
        return self._boundary.kernel_gradient(name, alpha, h)

    def moment_gradient(self, name="w2", alpha=(0, 0, 0), h=1):
        return self._boundary.moment_gradient(name, alpha, h)

    def kernel_hessian(self, name="w2", alpha=(0, 0, 0), h=1):
        return self._boundary.kernel_hessian(name, alpha, h)

    def laplacian(self, name="w2", alpha=(0, 0, 0), h=1):
        return self._boundary.laplacian(name, alpha, h)


# ---------------------------------------------------------------------------------------------- test geometries
def icosphere(level: int, radius=1.0, center=(0.0, 0.0, 0.0)):
    """(vertices, faces) of an icosahedron subdivided `level` times, projected to the sphere, outward faces."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    Fc = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
          (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7),
          (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(level):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                mid[key] = len(V) - 1
            return mid[key]
        nf = []
        for a, b, c in Fc:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        Fc = nf
    Vn = np.array(V) * radius + np.asarray(center, dtype=np.float64)
    return Vn, np.array(Fc, dtype=np.int32)


def box_surface(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """(vertices, faces) of an axis-aligned box, 12 outward triangles."""
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    V = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)])
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    Fc = []
    c = V.mean(axis=0)
    for a, b, cc, d in quads:
        for tri in ((a, b, cc), (a, cc, d)):
            p, q, r = V[list(tri)]
            n = np.cross(q - p, r - p)
            if np.dot(n, p - c) < 0:
                tri = (tri[0], tri[2], tri[1])
            Fc.append(tri)
    return V, np.array(Fc, dtype=np.int32)
