"""Continuous-batching service for hologram reconstruction and phase-retrieval requests.

A request is one hologram to reconstruct (xy min-projection over its slices), or a two-plane
Gerchberg-Saxton pair whose retrieved wavefront is then reconstructed (the reference's
workflow: cu_phase_retrieval_holo followed by cu_get_reconst_xyprojection, pr.md:95-109).

The serving loop (one host thread):

* requests arrive over time (their ``arrival`` offsets) into a FIFO;
* admission: the oldest waiting request and the requests of its size among the next 63 are
  admitted (one size per group, so its launches carry several holograms and GS pairs) while
  their device-memory footprint
  (inputs, outputs, retrieved wavefronts) fits the budget left by the in-flight work and the
  slots' reserved workspaces, up to ``max_batch`` per launch and ``max_group_work`` bytes of
  work per group (the reference-traffic measure of ``Request.work_bytes``; 32 GB, about 5 ms
  of work: a group finishes when its slowest
  request does, so small groups keep the tail short -- measured p99 211 -> ~120 ms at 75% load
  with unchanged capacity). The cap grows to an eighth of the waiting work per slot when the
  queue is that deep: latency is then set by queueing, and larger groups (more holograms and
  GS pairs per launch) drain it sooner -- capacity of a 1024^2 stream 235-244k -> 253k
  slices/s, p50 / p99 at 75% of it 6.7 / 21.9 -> 8.2 / 25.0 ms;
* an admitted group runs on a free slot -- a default-priority stream with its own libholo
  handle (own workspace, plans and tuning) -- as one batched phase retrieval (its GS requests)
  and one batched reconstruction, so groups admitted while others run overlap on the device;
* completion is detected by polling an event per group; the projection is copied back, the
  footprint released and the latency (completion - arrival) recorded.

A request's result does not depend on what it is batched with: the batched entry points
compute every hologram exactly as the single-hologram calls do.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from ._lib import check
from .batch import GSJob, ReconJob, phase_retrieval_batch, reconst_batch


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start of the stream
    holo: torch.Tensor             # Float32 (ny, nx), host (pinned) or device
    z_front: float
    z_dz: float
    slices: int
    holo2: Optional[torch.Tensor] = None   # GS requests: the second plane
    gs_z: float = 0.0
    gs_iters: int = 0
    wavlen: float = 0.6328
    dx: float = 10.0
    # filled by the server
    proj: Optional[torch.Tensor] = None
    t_admit: float = math.nan
    t_done: float = math.nan

    @property
    def is_gs(self) -> bool:
        return self.holo2 is not None

    def footprint(self) -> int:
        """Device bytes held while in flight: inputs, projection, GS wavefront."""
        ny, nx = self.holo.shape
        px = nx * ny
        b = 4 * px + 4 * px
        if self.is_gs:
            b += 4 * px + 8 * px
        return b

    def work_bytes(self) -> int:
        """The request's work measure for admission: the HBM bytes of the reference's per-slice
        sequence, 32 B/px per slice and 72 B/px per GS iteration ("reference traffic")."""
        ny, nx = self.holo.shape
        return nx * ny * (32 * self.slices + 72 * self.gs_iters)

    def fused_bytes(self, px_bytes_per_slice: float) -> float:
        """Algorithmic HBM bytes of the fused design: ``px_bytes_per_slice`` (16 + 8/spg_gen +
        8/spg_out for the size's tiles) per pixel and slice, 72 B/px per GS iteration."""
        ny, nx = self.holo.shape
        return nx * ny * (px_bytes_per_slice * self.slices + 72.0 * self.gs_iters)


@dataclass
class _Slot:
    stream: torch.cuda.Stream
    handle: _lib.Handle
    busy: bool = False


@dataclass
class _Group:
    slot: _Slot
    reqs: List[Request]
    event: torch.cuda.Event
    bytes: int
    outs: list = field(default_factory=list)


def select_batch(waiting: deque, used: int, budget: int, max_batch: int, others_in_flight: bool,
                 max_work: Optional[int] = None, lookahead: int = 0, key=None) -> list:
    """Admission: the oldest waiting requests, in order, while their footprints fit the budget
    left by the work in flight, up to ``max_batch`` requests and (when given) ``max_work``
    algorithmic bytes per group, so that one group's slowest request does not hold back many
    others. A request larger than the whole budget is still admitted alone when nothing else is
    in flight (it could never fit otherwise); the first request of a group always fits the work
    cap. With ``lookahead`` > 0 the group is the oldest request and, among the next
    ``lookahead`` - 1, those of the same ``key`` (the hologram size), in order: one size per
    group, so that its launches carry several holograms and GS pairs of that size; the others
    keep their places."""
    if lookahead > 0 and key is not None and waiting:
        return _select_same_key(waiting, used, budget, max_batch, others_in_flight, max_work, lookahead, key)
    batch, fp, work = [], 0, 0
    while waiting and len(batch) < max_batch:
        f = waiting[0].footprint()
        if used + fp + f > budget and (batch or others_in_flight):
            break
        w = waiting[0].work_bytes() if max_work is not None else 0
        if max_work is not None and batch and work + w > max_work:
            break
        batch.append(waiting.popleft())
        fp += f
        work += w
    return batch


def _select_same_key(waiting, used, budget, max_batch, others_in_flight, max_work, lookahead, key):
    head = waiting[0]
    f = head.footprint()
    if used + f > budget and others_in_flight:
        return []
    k0 = key(head)
    picked, fp, work = [0], f, head.work_bytes() if max_work is not None else 0
    for i in range(1, min(lookahead, len(waiting))):
        if len(picked) >= max_batch:
            break
        r = waiting[i]
        if key(r) != k0:
            continue
        f = r.footprint()
        if used + fp + f > budget:
            break
        w = r.work_bytes() if max_work is not None else 0
        if max_work is not None and work + w > max_work:
            break
        picked.append(i)
        fp += f
        work += w
    # only the look-ahead window is rewritten: the picked requests leave it, the others keep
    # their order at the front of the queue
    front = [waiting.popleft() for _ in range(picked[-1] + 1)]
    sel = set(picked)
    for r in reversed([r for i, r in enumerate(front) if i not in sel]):
        waiting.appendleft(r)
    return [front[i] for i in picked]


# Round sizes of the slot handles where they differ from a lone stream's (batch, spg_gen,
# spg_out, cols, rows). None at present: with the blocked intermediate slices the slots serve
# more with the lone tiles (capacity of the 1500-request stream on MI355X, 1024^2-equivalent
# slices/s: 214.9k against 208.2k with 16-slice rounds at 2048^2), and equal tiles keep a served
# result bitwise equal to the same request alone.
SLOT_TUNING: dict = {}


def group_work_cap(max_work: Optional[int], waiting_work: int, n_slots: int) -> Optional[int]:
    """The per-group work cap, grown with the queue: when the waiting work exceeds eight groups
    of the cap per slot, latency is set by queueing, and larger groups (more holograms and GS
    pairs per launch) drain the queue sooner. None: no cap."""
    if max_work is None:
        return None
    return max(max_work, waiting_work // (8 * n_slots))


def _size_key(r: Request):
    return tuple(r.holo.shape)


class Server:
    # Two slots: measured on the 1000-request stream (capacity, then p50/p99 at 75% of it),
    # four slots 165k eq. slices/s, 12.2 / 132 ms; three 171k, 13.8 / 116 ms; two 176k,
    # 15.2 / 113 ms. Fewer slots share the device among fewer groups, so the 2048^2 requests
    # that make the tail finish sooner; small requests wait a little longer.
    def __init__(self, device: int = 0, n_slots: int = 2, max_batch: int = 32,
                 hbm_budget: int = 32 << 30, workspace_per_slot: int = 4 << 30,
                 max_group_work: Optional[int] = 32 << 30, keep_results: bool = True,
                 lookahead: int = 64):
        # keep_results: each request keeps its host projection (Request.proj); False hands the
        # result over at completion and releases its pinned host buffer for reuse, so that a
        # long stream holds pinned memory only for the groups in flight
        self.keep_results = keep_results
        # groups of one hologram size from the oldest request and the next `lookahead` - 1
        # (0: strict arrival order, mixed sizes). Measured on the 3000-request mixed stream,
        # three runs each: capacity 220-226k -> 226-230k eq. slices/s with 64, p50 / p99 at 75%
        # load unchanged (16.4-16.7 / 85-87 ms against 16.5-17.1 / 86 ms)
        self.lookahead = lookahead
        self.device = device
        self.dev = torch.device(f"cuda:{device}")
        self.max_batch = max_batch
        self.max_group_work = max_group_work
        self.budget = hbm_budget - n_slots * workspace_per_slot
        if self.budget <= 0:
            raise ValueError("HBM budget smaller than the slots' workspaces")
        self.slots = [_Slot(torch.cuda.Stream(device=self.dev), _lib.Handle(device)) for _ in range(n_slots)]
        for s in self.slots:
            s.handle.set_workspace_limit(workspace_per_slot)
            # the whole workspace up front: the batched calls plan within the limit, so no call
            # grows it (growth synchronizes the device and would drain the other slots' work)
            s.handle.reserve_workspace(workspace_per_slot)
            for (nx, ny), t in SLOT_TUNING.items():
                check(_lib.load().holo_set_tuning_for_size(s.handle.ptr, nx, ny, *t))

    def workspace_sizes(self) -> list:
        return [s.handle.workspace_size() for s in self.slots]

    def _launch(self, slot: _Slot, reqs: List[Request], now: float) -> _Group:
        with torch.cuda.stream(slot.stream):
            dev_in = []
            for r in reqs:
                h1 = r.holo.to(self.dev, non_blocking=True)
                h2 = r.holo2.to(self.dev, non_blocking=True) if r.is_gs else None
                dev_in.append((h1, h2))
            gs = [(i, r) for i, r in enumerate(reqs) if r.is_gs]
            wfs = {}
            if gs:
                outs = phase_retrieval_batch([GSJob(dev_in[i][0], dev_in[i][1], r.gs_z, r.gs_iters, r.wavlen, r.dx)
                                              for i, r in gs], self.device, slot.handle)
                wfs = {i: o for (i, _), o in zip(gs, outs)}
            jobs = [ReconJob(wfs.get(i, dev_in[i][0]), r.z_front, r.z_dz, r.slices, r.wavlen, r.dx)
                    for i, r in enumerate(reqs)]
            projs = reconst_batch(jobs, self.device, slot.handle)
            host = [p.to("cpu", non_blocking=True) for p in projs]
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        for r in reqs:
            r.t_admit = now
        g = _Group(slot, reqs, ev, sum(r.footprint() for r in reqs))
        g.outs = [dev_in, projs, wfs, host]  # keep buffers alive until the group completes
        return g

    def run(self, requests: List[Request], spin: float = 50e-6) -> dict:
        """Serves the stream in real time (arrival offsets from now); returns statistics."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque = deque()
        waiting_work = 0  # reference-traffic bytes of the waiting requests
        inflight: List[_Group] = []
        used = 0
        t0 = time.perf_counter()
        groups = 0
        while pending or waiting or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting_work += pending[0].work_bytes()
                waiting.append(pending.popleft())
            for g in list(inflight):
                if g.event.query():
                    done = time.perf_counter() - t0
                    for r, hp in zip(g.reqs, g.outs[3]):
                        r.proj = hp if self.keep_results else None
                        r.t_done = done
                    used -= g.bytes
                    g.slot.busy = False
                    inflight.remove(g)
            free = [s for s in self.slots if not s.busy]
            if waiting and free:
                cap = group_work_cap(self.max_group_work, waiting_work, len(self.slots))
                batch = select_batch(waiting, used, self.budget, self.max_batch, bool(inflight), cap,
                                     self.lookahead, _size_key)
                waiting_work -= sum(r.work_bytes() for r in batch)
                fp = sum(r.footprint() for r in batch)
                if batch:
                    slot = free[0]
                    slot.busy = True
                    inflight.append(self._launch(slot, batch, now))
                    used += fp
                    groups += 1
                    continue
            if not (waiting and free):
                if pending and not inflight and not waiting:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.002)))
                else:
                    time.sleep(spin)
        makespan = time.perf_counter() - t0
        return summarize(requests, makespan, groups)


def fused_px_bytes(sizes, device: int = 0) -> dict:
    """Fused-design bytes per pixel and slice for each (ny, nx): the spectrum read once per
    pass-A slice group, the column-transformed slice written and read, the projection
    read-modify-written once per pass-B group (16 + 8/spg_gen + 8/spg_out)."""
    from .batch import get_tuning_for_size

    out = {}
    for (ny, nx) in sizes:
        t = get_tuning_for_size(nx, ny, 0, device)
        out[(ny, nx)] = 16.0 + 8.0 / t["spg_gen"] + 8.0 / t["spg_out"]
    return out


def _traffic(requests: List[Request], makespan: float) -> dict:
    """Reference-traffic and fused-design (algorithmic) GB/s of a served set of requests."""
    ref = sum(r.work_bytes() for r in requests)
    pxb = fused_px_bytes({tuple(r.holo.shape) for r in requests})
    alg = sum(r.fused_bytes(pxb[tuple(r.holo.shape)]) for r in requests)
    return {"algorithmic_GBs": alg / makespan / 1e9, "reference_traffic_GBs": ref / makespan / 1e9}


def summarize(requests: List[Request], makespan: float, groups: int = 0) -> dict:
    lat = np.array([r.t_done - r.arrival for r in requests])
    slices = sum(r.slices for r in requests)
    eq1024 = sum(r.slices * r.holo.shape[0] * r.holo.shape[1] for r in requests) / (1024 * 1024)
    per_size = {}
    for (ny, nx) in sorted({tuple(r.holo.shape) for r in requests}):
        ls = np.array([r.t_done - r.arrival for r in requests if tuple(r.holo.shape) == (ny, nx)])
        per_size[f"{nx}x{ny}"] = {"requests": int(ls.size), "p50_ms": float(np.percentile(ls, 50) * 1e3),
                                  "p99_ms": float(np.percentile(ls, 99) * 1e3)}
    return {"requests": len(requests), "groups": groups, "makespan_s": makespan,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "slices_per_s": slices / makespan,
            "slices_1024eq_per_s": eq1024 / makespan, **_traffic(requests, makespan),
            "per_size": per_size}


def plain_one_at_a_time(requests: List[Request], device: int = 0) -> dict:
    """The comparator: every request alone, in arrival order, with the single-hologram entry
    points on one stream (H2D, phase retrieval, reconstruction, D2H), no overlap."""
    dev = torch.device(f"cuda:{device}")
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for r in requests:
        h1 = r.holo.to(dev)
        if r.is_gs:
            src = _single_gs(h1, r.holo2.to(dev), r)
        else:
            src = h1
        proj = _single_recon(src, r)
        r.proj = proj.cpu()
        r.t_done = time.perf_counter() - t0
    makespan = time.perf_counter() - t0
    return summarize_closed(requests, makespan)


def _single_gs(h1, h2, r):
    import ctypes
    out = torch.empty(h1.shape, dtype=torch.complex64, device=h1.device)
    ny, nx = h1.shape
    f = _lib.holo_tf(None, r.gs_z, r.wavlen, r.dx, 1)
    b = _lib.holo_tf(None, -r.gs_z, r.wavlen, r.dx, 1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.load().holo_phase_retrieval(_lib.handle(0).ptr, st, nx, ny, ctypes.c_void_p(h1.data_ptr()),
                                                ctypes.c_void_p(h2.data_ptr()), ctypes.byref(f), ctypes.byref(b),
                                                r.gs_iters, ctypes.c_void_p(out.data_ptr()), 1))
    return out


def alone_gs(h1, h2, r):
    """The request's phase retrieval computed by itself through the batched entry point the
    server uses (the same tiles, so the served result equals it bit for bit)."""
    return phase_retrieval_batch([GSJob(h1, h2, r.gs_z, r.gs_iters, r.wavlen, r.dx)], h1.device.index or 0)[0]


def _single_recon(src, r):
    import ctypes
    ny, nx = src.shape
    proj = torch.empty((ny, nx), dtype=torch.float32, device=src.device)
    f = _lib.holo_tf(None, r.z_front, r.wavlen, r.dx, 1)
    d = _lib.holo_tf(None, r.z_dz, r.wavlen, r.dx, 1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _lib.load()
    fn = lib.holo_reconst_xyprojection_holo if src.dtype == torch.float32 else lib.holo_reconst_xyprojection
    _lib.check(fn(_lib.handle(0).ptr, st, nx, ny, ctypes.c_void_p(src.data_ptr()), ctypes.byref(f),
                  ctypes.byref(d), r.slices, ctypes.c_void_p(proj.data_ptr())))
    return proj


def summarize_closed(requests: List[Request], makespan: float) -> dict:
    slices = sum(r.slices for r in requests)
    return {"requests": len(requests), "makespan_s": makespan, "slices_per_s": slices / makespan,
            **_traffic(requests, makespan)}


SIZES = [(512, 512), (768, 768), (1024, 1024), (1280, 1024), (2048, 2048)]
GS_Z = 2000.0  # um between the two planes of a GS request


def make_stream(n: int, rate: float, seed: int = 105, pool: int = 3, gs_fraction: float = 0.2,
                slices=(128, 1024), gs_iters: int = 50, device: int = 0, pin: bool = True,
                sizes=SIZES) -> List[Request]:
    """A seeded Poisson stream of ``n`` requests at ``rate`` requests/s. Holograms come from a
    pool of ``pool`` seeded scenes per size (formed on the device, copied to pinned host
    memory, as a camera would deliver them); each request draws its own size, slice count
    (uniform in ``slices``), depth range and, for GS requests, 50 iterations on the scene's two
    recordings GS_Z apart."""
    from . import synth

    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = {}
    for si, (nx, ny) in enumerate(sizes):
        for k in range(pool):
            s = seed * 1000 + si * 10 + k
            # one scene recorded at the sensor and GS_Z further away: a genuine two-plane pair
            h1, h2 = synth.gs_pair_device(s, 300, nx, ny, separation=GS_Z, zmin=1e4, zmax=1e5, device=device)
            scenes[(si, k)] = tuple(t.cpu().pin_memory() if pin else t.cpu() for t in (h1, h2))
    t, reqs = 0.0, []
    for i in range(n):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        si = int(rng.integers(len(sizes)))
        k = int(rng.integers(pool))
        s = int(rng.integers(slices[0], slices[1] + 1))
        h1, h2 = scenes[(si, k)]
        gs = rng.random() < gs_fraction
        reqs.append(Request(i, t, h1, -1e4, -(2e5 - 1e4) / s, s, holo2=h2 if gs else None,
                            gs_z=GS_Z if gs else 0.0, gs_iters=gs_iters if gs else 0))
    return reqs
