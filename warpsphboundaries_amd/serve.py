"""Continuous-batching server of boundary-moment queries from many concurrent simulations (no reference
counterpart: the reference evaluates one position set per call).

    srv = BoundaryServer(device="cuda:0", hbm_budget=8 << 30, streams=4, form="mixed")
    srv.add_scene("tank", [(loop, "outside"), (obstacle, "inside")], support_max=h)    # resident geometry
    req = srv.submit("tank", positions, supports, kernel="w4")                        # any time, from any client
    done = srv.poll()                                                                  # retire finished, admit, launch

* Geometry, cell lists and their coordinate records stay resident on the device (``add_scene``); every resident
  byte is charged to the HBM budget.
* Requests are admitted by HBM footprint: a batch is launched only when its buffers (inputs, slots, indicator,
  outputs, sort scratch) fit in the budget left after the resident scenes and the batches in flight.  A request
  larger than the whole budget is rejected with ``MemoryError``; otherwise it waits.
* Continuous batching: pending requests of one (scene, kernel) are merged into a batch (up to ``max_batch_rows``);
  each of ``streams`` HIP streams carries one batch at a time, so batches of different scenes overlap on the device;
  as soon as a stream's batch retires, the oldest pending work is admitted into it.
* Batched launches across scenes (``batched=True``, the default for the fp64 and mixed forms): every resident
  geometry has a device descriptor (grid parameters and array pointers) in one table; a batch gathers the pending
  requests of one kernel from any number of scenes, each row names its geometry, and one adjacency launch plus one
  moment launch (per surface representation index) serve the whole batch.  Host work per batch is a few copies.
  A scene with a conditioned triangle (a closed 3-vertex loop whose rows may take the small / sliver quadrature of
  ``triangle.py``) is launched per scene, where each representation's rows are finalised on their own.
* 3-D scenes (``add_scene(name, None, support_max, mesh=(vertices, faces))``): a resident ``Mesh3D`` (face records,
  cell grid, walk lists); their requests are batched per (scene, kernel) into one cell sort, one face adjacency +
  winding launch and one launch of the 22 stored moments.  The face list of a batch is sized from the scene's
  previous pair counts; the count is read back with the batch's results, and a batch whose count exceeded the list is
  launched again with the exact size before its requests complete (no host synchronisation on the launch path).  The
  3-D path has the fp64 and mixed forms; a server of the plain form runs its 3-D scenes in fp64.
* Captured launches (``graphs=True``): a cross-scene batch's whole sequence (segment-table copy, row gathers, per
  representation index the row-geometry fill, the counting cell sort, adjacency and moments, the result scatter) is
  captured once per (stream, row bucket, kernel, representation count, request kind, sort) over fixed buffers and
  replayed: the segment tables have a fixed width (``GRAPH_SEGMENTS`` requests, unused segments empty) and the batch's
  row count travels in a header in front of them, which the kernels read (launch shapes follow the bucket, the rows
  beyond the batch are never touched), so a batch costs the host one table write and one graph launch.  A new scene
  changes the descriptor table and retires the captured graphs.
  Per-scene batches (``batched=False``) capture their adjacency + moment launches per (scene, kernel, stream, bucket).
* Latency = arrival (``submit``) to retirement (the poll that sees the batch's event complete), host clock.
Results are bitwise independent of batching: every row's channels are a fixed-order sum over its own slots.
"""
from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib as L
from . import triangle
from .scene import SurfaceRep, kernelName

F64 = torch.float64
PARK = 1.0e30          # coordinate of padded rows: outside every grid
GRAPH_SEGMENTS = 256   # requests per captured cross-scene batch
_HEADER = 256          # bytes before a batch's segment tables: its row count (int32), read by the counted launches


@dataclass
class Request:
    rid: int
    scene: str
    kernel: str
    positions: object                         # [n, 2] float64: numpy (host request) or a device tensor
    supports: object                          # [n] float64, same residency
    arrival: float
    done: Optional[float] = None
    result: Optional[object] = None           # [n, 12] float64, physical units, device channel order (numpy or tensor)
    error: Optional[BaseException] = None
    cancelled: bool = False
    out: Optional[torch.Tensor] = None        # device output of a device-resident request
    ready: Optional[torch.cuda.Event] = None  # device request: its inputs are written (submitter's stream)

    @property
    def n(self):
        return len(self.positions)

    @property
    def on_device(self):
        return isinstance(self.positions, torch.Tensor)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


@dataclass
class ResidentScene:
    name: str
    reps: list            # SurfaceRep
    support_max: float
    geoms: list
    K: int
    bytes: int
    desc: list = field(default_factory=list)     # descriptor-table index of each geometry
    mesh: Optional[object] = None                # 3-D scene: its Mesh3D
    pairs_per_row: float = 8.0                   # 3-D: face-list sizing from the previous batches
    conditioned: bool = False                    # 2-D: a representation is a conditioned triangle (per-scene launches)

    @property
    def dim(self):
        return 2 if self.mesh is None else 3


@dataclass
class _Slot:
    """per-stream fixed buffers for one row capacity (reused by every batch that fits)."""
    cap: int
    Kcap: int
    lpos: torch.Tensor
    lsup: torch.Tensor
    slots: torch.Tensor
    counts: torch.Tensor
    ind: torch.Tensor
    order: torch.Tensor
    out: torch.Tensor
    host_pos: torch.Tensor
    host_sup: torch.Tensor
    host_out: torch.Tensor
    sort_scratch: torch.Tensor
    bytes: int
    graphs: dict = field(default_factory=dict)
    filled: int = 0           # rows of lpos / lsup holding request data (the rest is parked)
    ready: Optional[torch.cuda.Event] = None   # the initialising fills (current stream) are done


_SEG = np.dtype([("src", "<u8"), ("dst", "<u8"), ("first", "<i8"), ("count", "<i8"), ("value", "<i8")])


@dataclass
class _BSlot:
    """per-stream fixed buffers of the cross-scene batches for one row capacity."""
    cap: int
    Kstride: int
    lpos: torch.Tensor
    lsup: torch.Tensor
    rdesc: torch.Tensor
    slots: torch.Tensor
    counts: torch.Tensor
    ind: torch.Tensor
    out: torch.Tensor
    host_pos: torch.Tensor
    host_sup: torch.Tensor
    host_desc: torch.Tensor
    host_out: torch.Tensor
    bytes: int
    order: Optional[torch.Tensor] = None
    sort_scratch: Optional[torch.Tensor] = None
    host_seg: Optional[torch.Tensor] = None      # pinned segment tables (rewritten only after the slot's batch retired)
    dev_seg: Optional[torch.Tensor] = None
    ready: Optional[torch.cuda.Event] = None
    graphs: dict = field(default_factory=dict)   # captured batches by (kernel, nrep, on_dev, sort, table version)


@dataclass
class _MSlot:
    """per-stream fixed buffers of the 3-D batches for one row capacity (the face list grows with the pair count)."""
    cap: int
    pos: torch.Tensor
    sup: torch.Tensor
    order: torch.Tensor
    offsets: torch.Tensor
    chi: torch.Tensor
    out: torch.Tensor
    adj_scratch: torch.Tensor
    host_pos: torch.Tensor
    host_sup: torch.Tensor
    host_out: torch.Tensor
    host_count: torch.Tensor
    bytes: int
    faces: Optional[torch.Tensor] = None
    sort_scratch: Optional[torch.Tensor] = None
    host_seg: Optional[torch.Tensor] = None
    dev_seg: Optional[torch.Tensor] = None
    ready: Optional[torch.cuda.Event] = None
    mom_scratch: Optional[torch.Tensor] = None
    prow: Optional[torch.Tensor] = None           # [cap, 4] the sorted-position layout's (x, y, z, h) records


def _mark_ready(slot, device):
    """slot buffers are allocated (and filled) on the current stream; a batch stream waits for that once."""
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(device))
    slot.ready = ev
    return slot


def _wait_ready(slot, stream):
    if slot.ready is not None:
        stream.wait_event(slot.ready)
        slot.ready = None


@dataclass
class _Batch:
    stream_index: int
    scene: ResidentScene
    kernel: str
    requests: List[Request]
    rows: int
    slot: object
    event: torch.cuda.Event
    bytes: int
    keep: Optional[list] = None       # host buffers an asynchronous copy of this batch still reads
    pair_capacity: int = 0            # 3-D: the face list size this batch was launched with


class BoundaryServer:
    def __init__(self, device="cuda:0", hbm_budget: int = 8 << 30, streams: int = 4, form: str = "mixed",
                 max_batch_rows: int = 1 << 18, sort_threshold: int = 1 << 16, graphs: bool = False,
                 batched: Optional[bool] = None):
        self.device = torch.device(device)
        self.budget = int(hbm_budget)
        self.form = form
        self.max_batch_rows = int(max_batch_rows)
        self.sort_threshold = int(sort_threshold)
        self.use_graphs = bool(graphs)
        self.batched = True if batched is None else bool(batched)
        self._desc_host = bytearray()
        self._desc_dev = None
        self._desc_bytes = int(L.load().wsb_geom2d_desc_bytes())
        self._nbins = 0                     # bins of the batch cell sort over every geometry of the table
        self._table_version = 0             # bumped by every change of the descriptor table
        self.captured_batches = 0
        self.replayed_batches = 0
        self._Kmax = 1
        self.bslots: List[Dict[int, _BSlot]] = []
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(int(streams))]
        self.busy = [None] * len(self.streams)
        self.slots: List[Dict[int, _Slot]] = [dict() for _ in self.streams]
        self.bslots = [dict() for _ in self.streams]
        self.mslots: List[Dict[int, _MSlot]] = [dict() for _ in self.streams]
        self.relaunched_batches = 0
        self.min_pair_capacity = 4096       # face-list slack of a 3-D batch beyond the scene's running estimate
        self.scenes: Dict[str, ResidentScene] = {}
        self.pending: deque = deque()
        self.resident_bytes = 0
        self.slot_bytes = 0
        self._rid = 0
        self.completed: List[Request] = []
        self.launched_batches = 0
        self.launched_rows = 0
        self.host_seconds = {"submit": 0.0, "launch": 0.0, "poll": 0.0}     # host time spent in the server

    # ------------------------------------------------------------------------------------------------ scenes
    def add_scene(self, name: str, loops, support_max: float, mesh=None):
        """make a scene resident: ``loops`` = [(vertices [n,2], 'inside' | 'outside')] or SurfaceReps (2-D), or
        ``mesh`` = (vertices [V,3], outward faces [F,3]) or a ``Mesh3D`` built for at least ``support_max`` (3-D)."""
        if name in self.scenes:
            raise ValueError(f"scene {name!r} exists")
        if mesh is not None:
            return self._add_mesh_scene(name, mesh, float(support_max))
        reps = [l if isinstance(l, SurfaceRep) else SurfaceRep.polygon(l[0], l[1]) for l in loops]
        geoms = [r.geom(float(support_max), self.device) for r in reps]
        nbytes = sum(int(g.info["device_bytes"]) for g in geoms)
        if self.resident_bytes + nbytes > self.budget:
            raise MemoryError(f"scene {name!r} needs {nbytes} resident bytes; {self.budget - self.resident_bytes} left")
        self.resident_bytes += nbytes
        sc = ResidentScene(name, reps, float(support_max), geoms, max(g.K for g in geoms), nbytes,
                           conditioned=any(r._triangle is not None for r in reps))
        lib = L.load()
        for g in geoms:
            buf = ctypes.create_string_buffer(self._desc_bytes)
            L.check(lib.wsb_geom2d_desc(g.handle, buf))
            sc.desc.append(len(self._desc_host) // self._desc_bytes)
            self._desc_host += buf.raw
        table = (ctypes.c_char * len(self._desc_host)).from_buffer(self._desc_host)
        nbins = ctypes.c_int32()
        L.check(lib.wsb_desc2d_cell_bases(table, len(self._desc_host) // self._desc_bytes, ctypes.byref(nbins)))
        del table
        self._nbins = int(nbins.value)
        self._desc_dev = torch.frombuffer(bytearray(self._desc_host), dtype=torch.uint8).to(self.device)
        self._table_version += 1
        for pool in self.bslots:             # captured batches hold the previous table's address
            for bs in pool.values():
                bs.graphs.clear()
        self._Kmax = max(self._Kmax, sc.K)
        self.scenes[name] = sc
        return sc

    def _add_mesh_scene(self, name, mesh, support_max):
        from .mesh3d import Mesh3D
        if isinstance(mesh, Mesh3D):
            m = mesh
            if m.host_only or m.device != self.device or m.support_max < support_max:
                raise ValueError("the Mesh3D must be resident on the server's device with support_max >= the scene's")
        else:
            V, F = mesh
            m = Mesh3D(V, F, support_max, device=self.device)
        nbytes = int(m.bytes)
        if self.resident_bytes + nbytes > self.budget:
            raise MemoryError(f"scene {name!r} needs {nbytes} resident bytes; {self.budget - self.resident_bytes} left")
        self.resident_bytes += nbytes
        sc = ResidentScene(name, [], support_max, [], 1, nbytes, mesh=m)
        self.scenes[name] = sc
        return sc

    def remove_scene(self, name: str):
        """drop a scene (its descriptor-table entries stay, unreferenced, until the server ends)."""
        if any(r.scene == name for r in self.pending) or any(b is not None and any(r.scene == name for r in b.requests)
                                                              for b in self.busy):
            raise ValueError(f"scene {name!r} has requests pending or in flight")
        sc = self.scenes.pop(name)
        self.resident_bytes -= sc.bytes

    # ------------------------------------------------------------------------------------------------ requests
    def submit(self, scene: str, positions, supports, kernel="w4") -> Request:
        """queue one timestep of one simulation.  Host arrays give a numpy result; a device tensor of positions (a
        simulation resident on this GPU) gives a device tensor result and never crosses PCIe."""
        t0 = time.perf_counter()
        try:
            return self._submit(scene, positions, supports, kernel)
        finally:
            self.host_seconds["submit"] += time.perf_counter() - t0

    def _submit(self, scene: str, positions, supports, kernel) -> Request:
        sc = self.scenes.get(scene)
        if sc is None:
            raise KeyError(f"unknown scene {scene!r}")
        k = kernelName(kernel)
        dim = sc.dim
        if isinstance(positions, torch.Tensor):
            if dim == 2 and (not self.batched or sc.conditioned):
                raise NotImplementedError("device-resident requests use the cross-scene batches" + (
                    ", which a scene with a conditioned triangle does not take" if sc.conditioned else ""))
            pos = positions.to(device=self.device, dtype=F64).reshape(-1, dim).contiguous()
            if isinstance(supports, torch.Tensor):
                sup = supports.to(device=self.device, dtype=F64).reshape(-1).contiguous()
                bad = len(sup) != len(pos) or (len(sup) and bool((~((sup > 0) & (sup <= sc.support_max))).any()))
            else:
                h = float(supports)
                sup = torch.full((len(pos),), h, dtype=F64, device=self.device)
                bad = not (0 < h <= sc.support_max)
            if bad:
                raise ValueError("supports must lie in (0, support_max] of the scene (the resident cell list is built for it)")
        else:
            pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, dim))
            sup = np.ascontiguousarray(np.broadcast_to(np.asarray(supports, dtype=np.float64), (len(pos),)))
            if len(pos) and not (np.all(sup > 0) and np.all(sup <= sc.support_max)):
                raise ValueError("supports must lie in (0, support_max] of the scene (the resident cell list is built for it)")
        req = Request(self._rid, scene, k, pos, sup, time.perf_counter())
        if req.on_device:
            req.out = torch.empty((len(pos), 12 if dim == 2 else 22), dtype=F64, device=self.device)
            req.ready = torch.cuda.Event()
            req.ready.record(torch.cuda.current_stream(self.device))
        self._rid += 1
        if self._footprint(sc, max(len(pos), 1)) + self.resident_bytes > self.budget:
            req.error = MemoryError(f"request of {len(pos)} rows exceeds the HBM budget")
            req.done = time.perf_counter()
            self.completed.append(req)
            return req
        self.pending.append(req)
        return req

    def cancel(self, req: Request):
        req.cancelled = True
        if req in self.pending:
            self.pending.remove(req)

    # ------------------------------------------------------------------------------------------------ scheduling
    def _footprint(self, sc: ResidentScene, rows: int) -> int:
        cap = _bucket(rows)
        if sc.dim == 3:
            pairs = int(sc.pairs_per_row * cap * 1.25) + 4096
            return cap * (24 + 8 + 4 + 8 + 8 + 176 + 16 + 32) + 4 * pairs + int(L.load().wsb_adjacency3d_scratch_bytes(cap)) + (1 << 20)
        K = self._Kmax if self.batched else sc.K
        per_row = 16 + 8 + 4 * K + 4 + 8 + 4 + 96 + 16 + 16 + 96
        return cap * per_row + (1 << 20)

    def _in_flight_bytes(self):
        return sum(b.bytes for b in self.busy if b is not None)

    def poll(self) -> List[Request]:
        """retire finished batches, admit and launch pending work; returns the requests completed now."""
        t0 = time.perf_counter()
        l0 = self.host_seconds["launch"]
        try:
            return self._poll()
        finally:
            self.host_seconds["poll"] += (time.perf_counter() - t0) - (self.host_seconds["launch"] - l0)

    def _poll(self) -> List[Request]:
        now_done = []
        for i, b in enumerate(self.busy):
            if b is not None and b.event.query():
                if b.scene is not None and b.scene.dim == 3:
                    total = int(b.slot.host_count[0])
                    if total < 0:       # the adjacency's device check: a support beyond the mesh's support_max
                        now = time.perf_counter()
                        for r in b.requests:
                            r.error = ValueError("support exceeds the mesh's support_max")
                            r.done = now
                            now_done.append(r)
                        self.busy[i] = None
                        continue
                    b.scene.pairs_per_row = max(b.scene.pairs_per_row, total / max(b.rows, 1))
                    if total > b.pair_capacity:           # face list too short: launch again, exactly sized
                        self.relaunched_batches += 1
                        self._launch_mesh(i, b.scene, b.kernel, b.requests, b.rows, b.bytes, pairs=total)
                        continue
                now = time.perf_counter()
                res = None
                off = 0
                for r in b.requests:
                    if not r.cancelled:
                        if r.on_device:
                            r.result = r.out
                        else:
                            if res is None:
                                res = b.slot.host_out[:b.rows].numpy()
                            r.result = res[off:off + r.n].copy()
                    r.done = now
                    off += r.n
                    now_done.append(r)
                self.busy[i] = None
        while self.pending:
            free = [i for i, b in enumerate(self.busy) if b is None]
            if not free:
                break
            head = self.pending[0]
            sc = self.scenes[head.scene]
            group, rows = [], 0
            dim = sc.dim
            for r in list(self.pending):
                rs = self.scenes[r.scene]
                cross = self.batched and dim == 2 and not sc.conditioned and not rs.conditioned
                same = (r.kernel == head.kernel and r.on_device == head.on_device and rs.dim == dim
                        and (cross or r.scene == head.scene))
                if same and len(group) >= self._group_limit(dim):
                    break
                if same and (not group or rows + r.n <= self.max_batch_rows):
                    group.append(r)
                    rows += r.n
                    if rows >= self.max_batch_rows:
                        break
            fp = self._footprint(sc, max(rows, 1))
            if self.resident_bytes + self._in_flight_bytes() + fp > self.budget:
                if len(group) > 1:          # admit less: the head alone
                    group, rows = [head], head.n
                    fp = self._footprint(sc, max(rows, 1))
                if self.resident_bytes + self._in_flight_bytes() + fp > self.budget:
                    break                   # wait for memory to free up
            for r in group:
                self.pending.remove(r)
            si = free[0]
            if self.use_graphs:
                # stream affinity: a free stream whose buffers already hold this batch's captured graph
                cap = _bucket(rows)
                for i in free:
                    if dim == 2 and self.batched and not sc.conditioned:
                        sl = self.bslots[i].get(cap)
                        hit = sl is not None and any(k[0] == head.kernel and k[-1] == self._table_version for k in sl.graphs)
                    else:
                        sl = self.slots[i].get(cap)
                        hit = sl is not None and (sc.name, head.kernel) in sl.graphs
                    if hit:
                        si = i
                        break
            t1 = time.perf_counter()
            if dim == 3:
                self._launch_mesh(si, sc, head.kernel, group, rows, fp)
            elif self.batched and not sc.conditioned:
                self._launch_batched(si, head.kernel, group, rows, fp)
            else:
                self._launch(si, sc, head.kernel, group, rows, fp)
            self.host_seconds["launch"] += time.perf_counter() - t1
        self.completed.extend(now_done)
        return now_done

    def _group_limit(self, dim):
        return GRAPH_SEGMENTS if (self.use_graphs and self.batched and dim == 2) else 1 << 30

    def run(self, until_idle=True, sleep=0.0):
        while self.pending or any(b is not None for b in self.busy):
            self.poll()
            if sleep:
                time.sleep(sleep)

    # ------------------------------------------------------------------------------------------------ launches
    def _slot(self, si: int, sc: ResidentScene, rows: int) -> _Slot:
        cap = _bucket(rows)
        pool = self.slots[si]
        key = cap
        s = pool.get(key)
        if s is None or s.Kcap < sc.K:
            dev = self.device
            Kcap = max(sc.K, 1 if s is None else s.Kcap)
            sort_bytes = max(int(L.load().wsb_sort2d_scratch_bytes(g.handle, cap)) for g in sc.geoms)
            s = _Slot(cap, Kcap,
                      torch.full((cap, 2), PARK, dtype=F64, device=dev), torch.ones(cap, dtype=F64, device=dev),
                      torch.empty(cap * Kcap, dtype=torch.int32, device=dev), torch.empty(cap, dtype=torch.int32, device=dev),
                      torch.empty(cap, dtype=F64, device=dev), torch.empty(cap, dtype=torch.int32, device=dev),
                      torch.empty((cap, 12), dtype=F64, device=dev),
                      torch.empty((cap, 2), dtype=F64).pin_memory(), torch.empty(cap, dtype=F64).pin_memory(),
                      torch.empty((cap, 12), dtype=F64).pin_memory(),
                      torch.empty(max(sort_bytes, 1), dtype=torch.uint8, device=dev), cap * (16 + 8 + 4 * Kcap + 4 + 8 + 4 + 96))
            pool[key] = _mark_ready(s, dev)
        return s

    def _evaluate(self, sc: ResidentScene, kernel: str, s: _Slot, n: int, sort: bool):
        """the launches of one batch over the slot's buffers (current stream)."""
        for j, g in enumerate(sc.geoms):
            order = None
            if sort:
                order = g.sort(s.lpos[:n], order=s.order[:n], scratch=s.sort_scratch)
            g.adjacency(s.lpos[:n], s.lsup[:n], pad=False, slots=s.slots[:n * g.K], counts=s.counts[:n], ind=s.ind[:n],
                        order=order)
            rep = sc.reps[j]
            if rep._triangle is None:
                flags = L.OUT_PHYSICAL | (L.OUT_ACCUMULATE if j else 0)
                g.moments(kernel, s.slots[:n * g.K], s.counts[:n], s.ind[:n], s.lpos[:n], s.lsup[:n], form=self.form,
                          flags=flags, out=s.out[:n])
            else:
                # a conditioned triangle: its own rows first, the quadrature rows overwritten, then accumulated
                own = s.out[:n] if j == 0 else torch.empty((n, 12), dtype=F64, device=self.device)
                g.moments(kernel, s.slots[:n * g.K], s.counts[:n], s.ind[:n], s.lpos[:n], s.lsup[:n], form=self.form,
                          flags=L.OUT_PHYSICAL, out=own)
                triangle.override(rep.vertices, rep._triangle, kernel, rep.background, s.lpos[:n], s.lsup[:n],
                                  flags=L.OUT_PHYSICAL, out=own)
                if j:
                    s.out[:n] += own

    def _launch(self, si: int, sc: ResidentScene, kernel: str, group: List[Request], rows: int, fp: int):
        s = self._slot(si, sc, rows)
        st = self.streams[si]
        hp, hs = s.host_pos.numpy(), s.host_sup.numpy()
        off = 0
        for r in group:
            hp[off:off + r.n] = r.positions
            hs[off:off + r.n] = r.supports
            off += r.n
        use_graph = self.use_graphs
        with torch.cuda.stream(st):
            _wait_ready(s, st)
            s.lpos[:rows].copy_(s.host_pos[:rows], non_blocking=True)
            s.lsup[:rows].copy_(s.host_sup[:rows], non_blocking=True)
            if use_graph:
                if s.filled > rows:                                   # park the rows of a larger previous batch
                    s.lpos[rows:s.filled].fill_(PARK)
                    s.lsup[rows:s.filled].fill_(sc.support_max)
                s.filled = rows
                key = (sc.name, kernel)
                gr = s.graphs.get(key)
                if gr is None:
                    self._evaluate(sc, kernel, s, s.cap, False)      # warm-up outside the capture
                    gr = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(gr, stream=st):
                        self._evaluate(sc, kernel, s, s.cap, False)
                    s.graphs[key] = gr
                gr.replay()
            else:
                s.filled = max(s.filled, rows)
                self._evaluate(sc, kernel, s, rows, rows >= self.sort_threshold)
            s.host_out[:rows].copy_(s.out[:rows], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        self.busy[si] = _Batch(si, sc, kernel, group, rows, s, ev, fp)
        self.launched_batches += 1
        self.launched_rows += rows

    def _bslot(self, si: int, rows: int) -> _BSlot:
        cap = _bucket(rows)
        pool = self.bslots[si]
        s = pool.get(cap)
        if s is None or s.Kstride < self._Kmax:
            dev, K = self.device, self._Kmax
            s = _BSlot(cap, K, torch.empty((cap, 2), dtype=F64, device=dev), torch.empty(cap, dtype=F64, device=dev),
                       torch.empty(cap, dtype=torch.int32, device=dev), torch.empty(cap * K, dtype=torch.int32, device=dev),
                       torch.empty(cap, dtype=torch.int32, device=dev), torch.empty(cap, dtype=F64, device=dev),
                       torch.empty((cap, 12), dtype=F64, device=dev),
                       torch.empty((cap, 2), dtype=F64).pin_memory(), torch.empty(cap, dtype=F64).pin_memory(),
                       torch.empty(cap, dtype=torch.int32).pin_memory(), torch.empty((cap, 12), dtype=F64).pin_memory(),
                       cap * (16 + 8 + 4 + 4 * K + 4 + 8 + 96))
            pool[cap] = _mark_ready(s, dev)
        return s

    def _batch_tables(self, s: _BSlot, group: List[Request], nrep: int, on_dev: bool, width: int):
        """the segment tables of one batch: [pos gather | sup gather | out scatter | row geometry of pass 0 .. nrep-1],
        ``width`` segments each (the segments beyond the group are empty)."""
        seg = np.zeros((3 + nrep, width), dtype=_SEG)
        off = 0
        hp = hs = None
        if not on_dev:
            hp, hs = s.host_pos.numpy(), s.host_sup.numpy()
        for g, r in enumerate(group):
            seg["first"][:, g] = off
            seg["count"][:, g] = r.n
            d = self.scenes[r.scene].desc
            for j in range(nrep):
                seg["value"][3 + j, g] = d[j] if j < len(d) else -1
            if on_dev:
                seg["src"][0, g] = r.positions.data_ptr()
                seg["src"][1, g] = r.supports.data_ptr()
                seg["dst"][2, g] = r.out.data_ptr()
            else:
                hp[off:off + r.n] = r.positions
                hs[off:off + r.n] = r.supports
            off += r.n
        return seg

    def _batch_sequence(self, s: _BSlot, kernel: str, nrep: int, on_dev: bool, n: int, width: int, sort: bool, nbytes: int,
                        counted: bool):
        """the launches of one cross-scene batch over the slot's buffers, on the current stream: launch shapes for
        ``n`` rows and ``width`` segments per table; ``counted``: the kernels serve the row count the batch header
        (the first bytes of the copied tables) carries, so a captured sequence serves any batch up to n rows."""
        lib = L.load()
        from .kernels import KERNEL_ID
        from .device import surface_device_plan
        kid, fid = KERNEL_ID[kernel], L.form_id(self.form)
        plan = surface_device_plan(kernel, self.device.index or 0).handle if fid == 0 else None
        sb = _SEG.itemsize
        s.dev_seg[:nbytes].copy_(s.host_seg[:nbytes], non_blocking=True)
        sp = L.stream_ptr(self.device)
        base = s.dev_seg.data_ptr()
        nrows = ctypes.c_void_p(base) if counted else None
        base += _HEADER
        tab = [ctypes.c_void_p(base + t * width * sb) for t in range(3 + nrep)]
        if on_dev:
            L.check(lib.wsb_gather_rows(tab[0], width, 2, L.ptr(s.lpos), sp))
            L.check(lib.wsb_gather_rows(tab[1], width, 1, L.ptr(s.lsup), sp))
        else:
            s.lpos[:n].copy_(s.host_pos[:n], non_blocking=True)
            s.lsup[:n].copy_(s.host_sup[:n], non_blocking=True)
        for j in range(nrep):
            L.check(lib.wsb_fill_rows_i32(tab[3 + j], width, L.ptr(s.rdesc), sp))
            order = None
            if sort:
                L.check(lib.wsb_batch2d_cellsort(L.ptr(self._desc_dev), L.ptr(s.rdesc), L.ptr(s.lpos), n, self._nbins,
                                                 L.ptr(s.order), L.ptr(s.sort_scratch), s.sort_scratch.numel(), nrows, sp))
                order = s.order
            L.check(lib.wsb_batch2d_adjacency(L.ptr(self._desc_dev), L.ptr(s.rdesc), L.ptr(order), L.ptr(s.lpos), L.ptr(s.lsup),
                                              n, s.Kstride, L.ptr(s.slots), L.ptr(s.counts), L.ptr(s.ind), nrows, sp))
            flags = L.OUT_PHYSICAL | (L.OUT_ACCUMULATE if j else 0)
            if fid == 0:
                L.check(lib.wsb_batch2d_plain_moments(L.ptr(self._desc_dev), L.ptr(s.rdesc), plan, L.ptr(s.slots),
                                                      L.ptr(s.counts), s.Kstride, L.ptr(s.ind), L.ptr(s.lpos),
                                                      L.ptr(s.lsup), n, flags, L.ptr(s.out), nrows, sp))
            else:
                L.check(lib.wsb_batch2d_moments(L.ptr(self._desc_dev), L.ptr(s.rdesc), kid, fid, L.ptr(s.slots),
                                                L.ptr(s.counts), s.Kstride, L.ptr(s.ind), L.ptr(s.lpos), L.ptr(s.lsup), n,
                                                flags, L.ptr(s.out), nrows, sp))
        if on_dev:
            L.check(lib.wsb_scatter_rows(tab[2], width, 12, L.ptr(s.out), sp))
        else:
            s.host_out[:n].copy_(s.out[:n], non_blocking=True)

    def _launch_batched(self, si: int, kernel: str, group: List[Request], rows: int, fp: int):
        """one cross-scene batch: rows in (pinned copies for host requests, a gather kernel for device requests), per
        surface-representation index one fill of the row geometries, one cell sort (large batches), one adjacency and
        one moment launch, rows out; captured once and replayed with ``graphs=True``."""
        s = self._bslot(si, rows)
        st = self.streams[si]
        nrep = max(len(self.scenes[r.scene].geoms) for r in group)
        on_dev = group[0].on_device
        captured = self.use_graphs
        width = GRAPH_SEGMENTS if captured else len(group)
        n = s.cap if captured else rows
        sort = (s.cap if captured else rows) >= self.sort_threshold
        seg = self._batch_tables(s, group, nrep, on_dev, width)
        raw = seg.view(np.uint8).reshape(-1)
        need = _HEADER + (3 + nrep) * GRAPH_SEGMENTS * _SEG.itemsize if captured else _HEADER + raw.size
        if s.host_seg is None or s.host_seg.numel() < need:
            nb = max(need, 1 << 16)
            s.host_seg = torch.empty(nb, dtype=torch.uint8).pin_memory()
            s.dev_seg = torch.empty(nb, dtype=torch.uint8, device=self.device)
            s.graphs.clear()                 # they copy from the previous buffers
        hseg = s.host_seg.numpy()
        hseg[:4] = np.array([rows], dtype=np.int32).view(np.uint8)     # the batch header: its row count
        hseg[_HEADER:_HEADER + raw.size] = raw
        nbytes = _HEADER + raw.size
        if sort:
            lib = L.load()
            sbytes = max(int(lib.wsb_batch2d_cellsort_scratch_bytes(s.cap, self._nbins)), 1)
            if s.order is None or s.sort_scratch is None or s.sort_scratch.numel() < sbytes:
                s.order = torch.empty(s.cap, dtype=torch.int32, device=self.device)
                s.sort_scratch = torch.empty(sbytes, dtype=torch.uint8, device=self.device)
                s.graphs.clear()
        with torch.cuda.stream(st):
            _wait_ready(s, st)
            if on_dev:
                for r in group:
                    st.wait_event(r.ready)
            if captured:
                key = (kernel, nrep, on_dev, sort, self._table_version)
                gr = s.graphs.get(key)
                if gr is None:
                    # this batch runs eagerly (it also warms every launch up); the same sequence is captured for the next
                    self._batch_sequence(s, kernel, nrep, on_dev, n, width, sort, nbytes, True)
                    # captured on this stream alone (no device-wide synchronisation, unlike torch.cuda.graph's entry):
                    # the sequence allocates nothing and makes no other call while the capture is open
                    gr = torch.cuda.CUDAGraph()
                    gr.capture_begin(capture_error_mode="thread_local")
                    try:
                        self._batch_sequence(s, kernel, nrep, on_dev, n, width, sort, nbytes, True)
                    finally:
                        gr.capture_end()
                    s.graphs[key] = gr
                    self.captured_batches += 1
                else:
                    gr.replay()
                    self.replayed_batches += 1
            else:
                self._batch_sequence(s, kernel, nrep, on_dev, n, width, sort, nbytes, False)
            ev = torch.cuda.Event()
            ev.record(st)
        self.busy[si] = _Batch(si, None, kernel, group, rows, s, ev, fp)
        self.launched_batches += 1
        self.launched_rows += rows

    def _mslot(self, si: int, rows: int) -> _MSlot:
        cap = _bucket(rows)
        pool = self.mslots[si]
        s = pool.get(cap)
        if s is None:
            dev = self.device
            ab = max(int(L.load().wsb_adjacency3d_scratch_bytes(cap)), 1)
            mb = max(int(L.load().wsb_moments3d_scratch_bytes(cap)), 1)
            s = _MSlot(cap, torch.empty((cap, 3), dtype=F64, device=dev), torch.empty(cap, dtype=F64, device=dev),
                       torch.empty(cap, dtype=torch.int32, device=dev), torch.empty(cap + 1, dtype=torch.int64, device=dev),
                       torch.empty(cap, dtype=F64, device=dev), torch.empty((cap, 22), dtype=F64, device=dev),
                       torch.empty(ab, dtype=torch.uint8, device=dev),
                       torch.empty((cap, 3), dtype=F64).pin_memory(), torch.empty(cap, dtype=F64).pin_memory(),
                       torch.empty((cap, 22), dtype=F64).pin_memory(), torch.empty(1, dtype=torch.int64).pin_memory(),
                       cap * (24 + 8 + 4 + 8 + 8 + 176 + 32) + ab + mb)
            s.mom_scratch = torch.empty(mb, dtype=torch.uint8, device=dev)
            s.prow = torch.empty((cap, 4), dtype=F64, device=dev)
            pool[cap] = _mark_ready(s, dev)
        return s

    def _launch_mesh(self, si: int, sc: ResidentScene, kernel: str, group: List[Request], rows: int, fp: int,
                     pairs: Optional[int] = None):
        """one 3-D batch of one scene: rows in, cell sort, face adjacency + winding, 22 moments, rows out, the pair
        count back to the host (pinned) for the face-list check at retirement."""
        from .kernels import KERNEL_ID
        s = self._mslot(si, rows)
        st = self.streams[si]
        m = sc.mesh
        lib = L.load()
        fid = L.form_id(self.form)
        if fid == 0:
            fid = L.form_id("fp64")          # the 3-D path's all-FP64 form
        cap_pairs = int(pairs if pairs is not None else sc.pairs_per_row * rows * 1.25) + self.min_pair_capacity
        if s.faces is None or s.faces.numel() < cap_pairs:
            s.faces = torch.empty(cap_pairs, dtype=torch.int32, device=self.device)
        cap_pairs = s.faces.numel()
        sb = int(lib.wsb_sort3d_scratch_bytes(m.handle, s.cap))
        if s.sort_scratch is None or s.sort_scratch.numel() < sb:
            s.sort_scratch = torch.empty(max(sb, 1), dtype=torch.uint8, device=self.device)
        on_dev = group[0].on_device
        ng = len(group)
        seg = np.zeros((3, ng), dtype=_SEG)
        off = 0
        if not on_dev:
            hp, hs = s.host_pos.numpy(), s.host_sup.numpy()
        for g, r in enumerate(group):
            seg["first"][:, g] = off
            seg["count"][:, g] = r.n
            if on_dev:
                seg["src"][0, g] = r.positions.data_ptr()
                seg["src"][1, g] = r.supports.data_ptr()
                seg["dst"][2, g] = r.out.data_ptr()
            else:
                hp[off:off + r.n] = r.positions
                hs[off:off + r.n] = r.supports
            off += r.n
        raw = seg.view(np.uint8).reshape(-1)
        if on_dev:
            if s.host_seg is None or s.host_seg.numel() < raw.size:
                nb = max(raw.size, 1 << 12)
                s.host_seg = torch.empty(nb, dtype=torch.uint8).pin_memory()
                s.dev_seg = torch.empty(nb, dtype=torch.uint8, device=self.device)
            s.host_seg.numpy()[:raw.size] = raw
        with torch.cuda.stream(st):
            sp = L.stream_ptr(self.device)
            _wait_ready(s, st)
            if on_dev:
                for r in group:
                    st.wait_event(r.ready)
                s.dev_seg[:raw.size].copy_(s.host_seg[:raw.size], non_blocking=True)
                base = s.dev_seg.data_ptr()
                tab = [ctypes.c_void_p(base + t * ng * seg.dtype.itemsize) for t in range(3)]
                L.check(lib.wsb_gather_rows(tab[0], ng, 3, L.ptr(s.pos), sp))
                L.check(lib.wsb_gather_rows(tab[1], ng, 1, L.ptr(s.sup), sp))
            else:
                s.pos[:rows].copy_(s.host_pos[:rows], non_blocking=True)
                s.sup[:rows].copy_(s.host_sup[:rows], non_blocking=True)
            L.check(lib.wsb_sort3d(m.handle, L.ptr(s.pos), rows, L.ptr(s.order), L.ptr(s.sort_scratch),
                                   s.sort_scratch.numel(), sp))
            # the sorted-position layout: the adjacency at the cell-order positions, the moments read it contiguously
            # and store each row's 22 values at the row (the values of the particle layout, bitwise)
            L.check(lib.wsb_adjacency3d_sorted(m.handle, L.ptr(s.pos), L.ptr(s.sup), L.ptr(s.order), rows,
                                               L.ptr(s.offsets), L.ptr(s.faces), cap_pairs, L.ptr(s.chi), L.ptr(s.prow),
                                               L.ptr(s.adj_scratch), s.adj_scratch.numel(), sp))
            L.check(lib.wsb_moments3d_sorted(m.handle, KERNEL_ID[kernel], fid, L.ptr(s.order), L.ptr(s.prow),
                                             L.ptr(s.offsets), L.ptr(s.faces), cap_pairs, L.ptr(s.chi), rows,
                                             L.OUT_PHYSICAL, L.ptr(s.out), L.ptr(s.mom_scratch), s.mom_scratch.numel(),
                                             sp))
            s.host_count.copy_(s.offsets[rows:rows + 1], non_blocking=True)
            if on_dev:
                L.check(lib.wsb_scatter_rows(tab[2], ng, 22, L.ptr(s.out), sp))
            else:
                s.host_out[:rows].copy_(s.out[:rows], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        self.busy[si] = _Batch(si, sc, kernel, group, rows, s, ev, fp, pair_capacity=cap_pairs)
        if pairs is None:
            self.launched_batches += 1
            self.launched_rows += rows

    # ------------------------------------------------------------------------------------------------ stats
    def latency_stats(self, requests=None):
        reqs = [r for r in (requests or self.completed) if r.latency is not None and r.error is None]
        lat = np.array([r.latency for r in reqs])
        if not len(lat):
            return {}
        return {"requests": len(lat), "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}


def _bucket(rows: int) -> int:
    """row capacity bucket: the next power of two, at least 1024."""
    cap = 1024
    while cap < rows:
        cap *= 2
    return cap
