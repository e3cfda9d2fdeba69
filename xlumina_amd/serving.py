"""Continuous-batching request loop for RS / VRS propagation and gradient requests.

A request is one optical set-up on a square grid of side n:
  * scalar:     Gaussian beam (waist `waist`) -> SLM with a phase mask ->
                RS to distance z (ScalarLight.RS_propagation,
                wave_optics.py:173-196, with the SLM of optical_elements.py:76-103
                fused into the row pass);
  * vectorial:  Gaussian beam with Jones vector (1, 1) -> sSLM with two phase
                masks (alpha on Ex, phi on Ey) -> VRS to z (Ex, Ey, Ez)
                (VectorizedLight.VRS_propagation, vectorized_optics.py:244-284,
                sSLM optical_elements.py:186-222).
The phase masks are U[-pi, pi) from a generator seeded per request, or the
client's own (`Request.phases`).  Every request returns the output intensity
|E|^2.  A gradient request also returns dL/d(first phase mask) of
L = sum_pixels w * |E|^2 with the fixed radial weight
w = (x^2 + y^2) / half_width^2, through the hand-written adjoints
(conjugate-kernel RS reusing the forward's kernel spectra, then the
sSLM/SLM and Ez-construction adjoint).

Execution model (one GPU):
  * A bucket is (n, scalar/vectorial, gradient).  Each bucket owns a pool of
    slots.  A slot holds the device buffers of up to `capacity` requests of
    its bucket (phase masks, distances, fields, intensities, gradients,
    workspace) and one HIP graph per batch size k = 1..capacity, captured
    over its first k members: a batch is one set of launches
    (xl_rs_propagate with batch = k, member b propagating over its own
    distance z_dev[b]).  The beam and the loss weight of each grid size are
    shared, read-only, by all slots of that size.
  * Continuous batching: whenever a slot of a bucket is free, the requests of
    that bucket that have arrived (up to `capacity`) are admitted together;
    the host loop keeps admitting while earlier batches run on their own
    streams, so the work of different batches overlaps on the device.
  * Admission by HBM footprint: slots are created while the pool total stays
    under `hbm_budget` bytes; when the budget is full, idle slots of buckets
    with no queued request are released (least recently used first) to make
    room; a bucket whose oldest request has waited longer than another
    bucket's may also take that bucket's idle slots, and buckets are served in
    the order of their oldest request, so no bucket starves under a tight
    budget.  A bucket whose one slot would exceed the whole budget is refused:
    its requests end with status "rejected" and run nothing.  Per-bucket
    queues: a waiting 2048^2 request never blocks a 512^2 one.
  * Quarantine: each batch also reports, per member, whether its outputs are
    all finite.  A member with a non-finite input (a NaN phase or distance)
    ends with status "nonfinite" and its result is dropped; members run in
    separate thread blocks (blockIdx.y) on separate workspace blocks, so its
    batch mates are unaffected.
  * Cancel: Server.cancel(rid) (from any host code, e.g. the on_launch hook)
    removes a queued request (status "cancelled", no GPU work); a request
    already on the GPU lets its batch finish and only its result is dropped.
The distance z is read by the kernels from device memory (z_dev), which is
what lets one captured graph serve every z.  Latency = completion seen by
the host - arrival time.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from . import ops
from .ops import MaskSpec
from .toolbox import space

OK, NONFINITE, REJECTED, CANCELLED = "ok", "nonfinite", "rejected", "cancelled"


@dataclass
class Request:
    rid: int
    n: int
    vectorial: bool
    z: float            # propagation distance, um
    grad: bool
    seed: int
    arrival: float = 0.0  # seconds after the start of the stream
    phases: Optional[np.ndarray] = field(default=None, compare=False)   # client phase masks (nin, n, n)


@dataclass(frozen=True)
class Bucket:
    n: int
    vectorial: bool
    grad: bool


def bucket_of(r: Request) -> Bucket:
    return Bucket(r.n, r.vectorial, r.grad)


def make_requests(count: int, rate: Optional[float] = None, seed: int = 7,
                  sizes: Sequence[int] = (512, 1024, 2048), size_p: Sequence[float] = (0.5, 0.3, 0.2),
                  p_vectorial: float = 0.3, z_range=(5e3, 1e5), p_grad: float = 0.2) -> List[Request]:
    """Seeded request stream: Poisson arrivals at `rate` req/s (None: all at
    t = 0), grid side from `sizes` with `size_p`, vectorial with
    p_vectorial, z ~ U[z_range] um, gradient with p_grad."""
    rng = np.random.default_rng(seed)
    n = rng.choice(np.asarray(sizes), size=count, p=np.asarray(size_p, dtype=np.float64))
    vec = rng.random(count) < p_vectorial
    z = rng.uniform(z_range[0], z_range[1], count)
    grad = rng.random(count) < p_grad
    seeds = rng.integers(0, 2 ** 62, count)
    if rate is None or math.isinf(rate):
        t = np.zeros(count)
    else:
        t = np.cumsum(rng.exponential(1.0 / rate, count))
        t -= t[0]
    return [Request(i, int(n[i]), bool(vec[i]), float(z[i]), bool(grad[i]), int(seeds[i]), float(t[i]))
            for i in range(count)]


def _client_phases(r: Request) -> np.ndarray:
    nin = 2 if r.vectorial else 1
    ph = np.asarray(r.phases, dtype=np.float32)
    if ph.shape != (nin, r.n, r.n):
        raise ValueError(f"request {r.rid}: phases must be {(nin, r.n, r.n)}, got {ph.shape}")
    return ph


def request_phases(r: Request, device, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """The request's phase masks: the client's (`r.phases`), else U[-pi, pi)
    generated on the device from r.seed (ops.uniform_phases, the generator the
    slots run inside their graphs); (1, n, n) for a scalar request (SLM),
    (2, n, n) for a vectorial one (sSLM alpha, phi).  Written into `out` when
    given."""
    nin = 2 if r.vectorial else 1
    if out is None:
        out = torch.empty((nin, r.n, r.n), dtype=torch.float32, device=device)
    if r.phases is not None:
        return out.copy_(torch.from_numpy(_client_phases(r)))
    seeds = torch.tensor([r.seed], dtype=torch.int64, device=device)
    ops.uniform_phases(out.view(1, -1), seeds)
    return out


@dataclass
class _Shared:
    """Read-only per-grid-size state: grid, beams, loss weight."""
    x: np.ndarray
    y: np.ndarray
    dx: float
    beam_s: torch.Tensor     # (1, n, n) complex64
    beam_v: torch.Tensor     # (2, n, n) complex64, Jones (1, 1)
    w2: torch.Tensor         # (n, n) float32, 2 * w (cotangent of |E|^2 -> E)


def slot_bytes(b: Bucket, capacity: int) -> int:
    """Device bytes of one slot of bucket b holding `capacity` requests."""
    n2 = b.n * b.n
    nin, nc = (2, 3) if b.vectorial else (1, 1)
    per = nin * n2 * 4 + 8 + nc * n2 * 8 + n2 * 4
    if b.grad:
        per += 2 * nc * n2 * 8 + n2 * 4 + ops.rs_hhat_bytes(b.n)
    return capacity * per + ops.rs_batch_workspace_bytes(b.n, nc, capacity)


class _Slot:
    """Device buffers and per-batch-size graphs of one bucket (GPU)."""

    def __init__(self, server: "Server", b: Bucket, capacity: int):
        dev = server.device
        n = b.n
        sh = server.shared(n)
        self.bucket = b
        self.capacity = capacity
        self.nin = 2 if b.vectorial else 1
        self.ncomp = 3 if b.vectorial else 1
        self.src = sh.beam_v if b.vectorial else sh.beam_s
        B = capacity
        self.params = torch.zeros((B, self.nin, n, n), dtype=torch.float32, device=dev)
        # per-member control words, one host-to-device copy per batch: distance (float64
        # bits), phase seed, and whether the graph generates the member's phases
        self.ctl_host = torch.zeros((3, B), dtype=torch.int64).pin_memory()
        self.ctl_host[0].view(torch.float64).fill_(1e4)
        self.ctl_np = self.ctl_host.numpy()                 # host writes through numpy (no per-element torch calls)
        self.ctl = self.ctl_host.to(dev)
        self.zd = self.ctl[0].view(torch.float64)
        self.seeds = self.ctl[1]
        self.gen_flags = self.ctl[2].view(torch.int32)[:B]   # int32 flags in the first half of row 2
        self.ok = torch.ones((B,), dtype=torch.bool, device=dev)
        self.ok_host = torch.ones((B,), dtype=torch.bool).pin_memory()
        self.out = torch.empty((B, self.ncomp, n, n), dtype=torch.complex64, device=dev)
        self.inten = torch.empty((B, n, n), dtype=torch.float32, device=dev)
        self.ws = torch.empty(ops.rs_batch_workspace_bytes(n, self.ncomp, B), dtype=torch.uint8, device=dev)
        if b.grad:
            self.gout = torch.empty_like(self.out)
            self.gpre = torch.empty_like(self.out)
            self.gmask = torch.empty((B, n, n), dtype=torch.float32, device=dev)
            self.hhat = torch.empty(B * ops.rs_hhat_bytes(n), dtype=torch.uint8, device=dev)
        else:
            self.gout = self.gpre = self.gmask = self.hhat = None
        self.sh = sh
        self.wavelength = server.wavelength
        self.stream = torch.cuda.Stream(device=dev)   # default priority
        self.event = torch.cuda.Event()
        self.graphs: Dict[int, torch.cuda.CUDAGraph] = {}
        self.requests: List[Request] = []
        self.last_used = 0.0
        self.bytes = sum(t.numel() * t.element_size() for t in
                         (self.params, self.zd, self.out, self.inten, self.ws, self.gout, self.gpre, self.gmask,
                          self.hhat) if t is not None)
        self.stream.wait_stream(torch.cuda.current_stream(dev))   # buffers and shared state were made there
        with torch.cuda.stream(self.stream):
            self.program(B)                        # eager run: fills caches
        self.stream.synchronize()
        if server.use_graphs:
            for k in range(1, B + 1):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.stream):
                    self.program(k)
                self.graphs[k] = g
            self.stream.synchronize()

    def _masks(self, k):
        p = self.params[:k]
        if self.bucket.vectorial:
            return [MaskSpec(N.XL_MASK_SSLM, p[:, 0], p[:, 1])], N.XL_EZ_VRS
        return [MaskSpec(N.XL_MASK_SLM, p[:, 0])], N.XL_EZ_NONE

    def program(self, k: int):
        """The bucket's kernel sequence over the first k members of this slot:
        seeded phase masks, propagation (+ adjoint and mask gradient)."""
        sh = self.sh
        masks, ez = self._masks(k)
        grad = self.bucket.grad
        ops.uniform_phases(self.params[:k], self.seeds[:k], self.gen_flags[:k])
        ops.rs_propagate_batched(self.src, sh.x, sh.y, self.wavelength, self.zd[:k], masks=masks, ez_mode=ez,
                                 nin=self.nin, intensity=True, batch=k, out=self.out[:k], inten=self.inten[:k],
                                 workspace=self.ws, hhat=self.hhat, hhat_mode=1 if grad else 0)
        if grad:
            torch.mul(self.out[:k], sh.w2, out=self.gout[:k])
            ops.rs_propagate_batched(self.gout[:k], sh.x, sh.y, self.wavelength, self.zd[:k], adjoint=True,
                                     out=self.gpre[:k], workspace=self.ws, hhat=self.hhat, hhat_mode=2)
            ops.pre_backward_batched(self.src, self.gpre[:k], sh.x, sh.y, k, masks=masks, ez_mode=ez, nin=self.nin,
                                     need_src=False, need_params=((True, False), (False, False)),
                                     gparams=[[self.gmask[:k], None], [None, None]], z_dev=self.zd[:k])

    def launch(self, reqs: List[Request], keep: bool):
        """Writes the requests' inputs into the first len(reqs) members and
        runs the batch on the slot stream; returns the kept results (or None)."""
        k = len(reqs)
        self.requests = list(reqs)
        ctl = self.ctl_np
        zh, flags = ctl[0].view(np.float64), ctl[2].view(np.int32)
        with torch.cuda.stream(self.stream):
            for i, r in enumerate(reqs):
                zh[i] = r.z
                ctl[1, i] = r.seed
                flags[i] = 0 if r.phases is not None else 1
                if r.phases is not None:
                    self.params[i].copy_(torch.from_numpy(_client_phases(r)))
            self.ctl.copy_(self.ctl_host, non_blocking=True)
            if k in self.graphs:
                self.graphs[k].replay()
            else:
                self.program(k)
            # a non-finite input (phase or distance) reaches the member's intensity, and
            # max propagates NaN.  Two stages (k x 256 partial maxima, then k): a
            # one-stage reduction runs one block per member over n^2 values
            fin = torch.isfinite(self.inten[:k].view(k, 256, -1).amax(2).amax(1))
            self.ok_host[:k].copy_(fin, non_blocking=True)
            results = None
            if keep:
                results = [(self.inten[i].clone(), self.gmask[i].clone() if self.gmask is not None else None)
                           for i in range(k)]
            self.event.record(self.stream)
        return results

    def done(self) -> bool:
        return self.event.query()

    def finite(self) -> List[bool]:
        """Per member of the last batch: all outputs finite (after done())."""
        return [bool(v) for v in self.ok_host[:len(self.requests)].tolist()]


@dataclass
class Report:
    completed: int                  # requests that ran (ok + nonfinite)
    seconds: float
    throughput: float               # completed requests / s
    p50_ms: float
    p99_ms: float
    mean_ms: float
    offered_rate: Optional[float]
    per_bucket: Dict[str, dict] = field(default_factory=dict)
    latencies_ms: Optional[np.ndarray] = None
    status: Dict[int, str] = field(default_factory=dict)
    batches: int = 0

    def counts(self) -> Dict[str, int]:
        out = {OK: 0, NONFINITE: 0, REJECTED: 0, CANCELLED: 0}
        for s in self.status.values():
            out[s] += 1
        return out

    def summary(self) -> dict:
        return {"completed": self.completed, "seconds": round(self.seconds, 4),
                "throughput_req_s": round(self.throughput, 2), "p50_ms": round(self.p50_ms, 4),
                "p99_ms": round(self.p99_ms, 4), "mean_ms": round(self.mean_ms, 4),
                "offered_rate_req_s": None if self.offered_rate is None else round(self.offered_rate, 2),
                "batches": self.batches, "mean_batch": round(self.completed / max(1, self.batches), 3),
                "status": self.counts()}


DEFAULT_CAPACITY = {512: 8, 1024: 4, 2048: 2}


def _head(q) -> tuple:
    """(arrival, rid) of the oldest request of a queue; empty queues sort last."""
    return (q[0].arrival, q[0].rid) if q else (math.inf, 0)


class Server:
    """Slot pools per bucket under an HBM budget; see the module docstring.

    slot_factory(server, bucket, capacity) -> slot and slot_size(bucket,
    capacity) -> bytes replace the GPU slots (the admission logic is the same
    for any slot with the _Slot interface: capacity, bytes, requests,
    last_used, launch, done, finite)."""

    def __init__(self, device=None, wavelength: float = 0.65, half_width: float = 1200.0, waist: float = 400.0,
                 hbm_budget: Optional[int] = None, slots_per_bucket: int = 2, use_graphs: bool = True,
                 capacity: Optional[Dict[int, int]] = None,
                 slot_factory: Optional[Callable] = None, slot_size: Optional[Callable] = None):
        self.fake = slot_factory is not None
        if self.fake:
            self.device = torch.device("cpu") if device is None else torch.device(device)
        else:
            self.device = torch.device(device) if device is not None else torch.device("cuda",
                                                                                       torch.cuda.current_device())
            N.ensure_device(self.device)
        self.wavelength = wavelength
        self.half_width = half_width
        self.waist = waist
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = min(int(free * 0.5), 64 << 30)
        self.hbm_budget = int(hbm_budget)
        self.slots_per_bucket = slots_per_bucket
        self.use_graphs = use_graphs
        self.capacity = dict(DEFAULT_CAPACITY if capacity is None else capacity)
        self._factory = slot_factory or (lambda srv, b, cap: _Slot(srv, b, cap))
        self._size = slot_size or slot_bytes
        self._shared: Dict[int, _Shared] = {}
        self._shared_bytes = 0
        self.pools: Dict[Bucket, list] = {}
        self.used_bytes = 0
        self.evictions = 0
        self._cancelled: set = set()
        self.results = None

    # -- resources -------------------------------------------------------
    def shared(self, n: int) -> _Shared:
        s = self._shared.get(n)
        if s is None:
            x, y = space(self.half_width, n)
            X, Y = np.meshgrid(x, y)
            g = np.exp(-(X ** 2 + Y ** 2) / self.waist ** 2)
            beam_s = torch.from_numpy(g[None].astype(np.complex64)).to(self.device)
            beam_v = torch.from_numpy((np.stack([g, g]) / math.sqrt(2.0)).astype(np.complex64)).to(self.device)
            w2 = torch.from_numpy((2.0 * (X ** 2 + Y ** 2) / self.half_width ** 2).astype(np.float32)).to(self.device)
            s = _Shared(x, y, float(x[1] - x[0]), beam_s, beam_v, w2)
            self._shared[n] = s
            nb = sum(t.numel() * t.element_size() for t in (beam_s, beam_v, w2))
            self._shared_bytes += nb
            self.used_bytes += nb
        return s

    def cap(self, b: Bucket) -> int:
        return max(1, int(self.capacity.get(b.n, 1)))

    def admissible(self, b: Bucket) -> bool:
        """A slot of b fits the budget at all (next to the shared state)."""
        return self._shared_bytes + self._size(b, self.cap(b)) <= self.hbm_budget

    def _release(self, slot) -> None:
        pool = self.pools[slot.bucket]
        pool.remove(slot)
        self.used_bytes -= slot.bytes
        self.evictions += 1
        slot.graphs = {}
        del slot
        if not self.fake:
            torch.cuda.empty_cache()

    def _grow(self, b: Bucket, queues, evict: bool = True) -> Optional[object]:
        pool = self.pools.setdefault(b, [])
        if len(pool) >= self.slots_per_bucket:
            return None
        need = self._size(b, self.cap(b))
        if self.used_bytes + need > self.hbm_budget and evict:
            # release idle slots of buckets that nobody has waited for longer (no queued request,
            # or a younger oldest request than b's), least recently used first
            head = _head(queues.get(b))
            idle = [s for bb, ps in self.pools.items()
                    if bb != b and (not queues.get(bb) or _head(queues[bb]) > head)
                    for s in ps if not s.requests]
            idle.sort(key=lambda s: s.last_used)
            while idle and self.used_bytes + need > self.hbm_budget:
                self._release(idle.pop(0))
        if self.used_bytes + need > self.hbm_budget:
            return None
        if not self.fake:
            self.shared(b.n)
        s = self._factory(self, b, self.cap(b))
        pool.append(s)
        self.used_bytes += s.bytes
        return s

    def prewarm(self, buckets) -> None:
        """Creates (and captures) the slots of the admissible buckets while
        the budget allows -- first one per bucket, then the rest -- so that no
        request pays for a capture unless the budget forces slots to be
        released and rebuilt."""
        order = sorted((b for b in buckets if self.admissible(b)), key=lambda b: (b.n, b.vectorial, b.grad))
        for b in order:
            if not self.pools.get(b):
                self._grow(b, {}, evict=False)
        for b in order:
            while self._grow(b, {}, evict=False) is not None:
                pass
        if not self.fake:
            torch.cuda.synchronize(self.device)

    def _free_slot(self, b: Bucket, queues):
        for s in self.pools.get(b, ()):
            if not s.requests:
                return s
        return self._grow(b, queues)

    def cancel(self, rid: int) -> None:
        """Cancels request rid (see the module docstring)."""
        self._cancelled.add(rid)

    # -- the loop ----------------------------------------------------------
    def run(self, requests: Sequence[Request], keep_results: bool = False, serial: bool = False,
            offered_rate: Optional[float] = None, on_launch: Optional[Callable] = None) -> Report:
        """Serves `requests` (sorted by arrival) in real time.  serial=True
        runs one request at a time, alone on the device.  on_launch(reqs) is
        called after each batch is put on the GPU."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        self._cancelled = set()
        self.prewarm({bucket_of(r) for r in reqs})
        queues: Dict[Bucket, deque] = {}
        busy: list = []
        lat: List[float] = []
        status: Dict[int, str] = {}
        results = {} if keep_results else None
        pending: Dict[int, object] = {}
        bucket_lat: Dict[Bucket, list] = {}
        batches = 0
        nxt = 0
        done = 0
        if not self.fake:
            torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        last = t0
        while done < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                r = reqs[nxt]
                nxt += 1
                b = bucket_of(r)
                if not self.admissible(b):
                    status[r.rid] = REJECTED
                    done += 1
                    continue
                queues.setdefault(b, deque()).append(r)
            if self._cancelled:
                for q in queues.values():
                    for r in [r for r in q if r.rid in self._cancelled]:
                        q.remove(r)
                        status[r.rid] = CANCELLED
                        done += 1
            # oldest waiting request first, across buckets
            for b, q in sorted(queues.items(), key=lambda kv: _head(kv[1])):
                while q and not (serial and busy):
                    s = self._free_slot(b, queues)
                    if s is None:
                        break
                    k = 1 if serial else min(len(q), s.capacity)
                    take = [q.popleft() for _ in range(k)]
                    res = s.launch(take, keep_results)
                    s.last_used = time.perf_counter()
                    batches += 1
                    if keep_results:
                        for r, v in zip(take, res):
                            pending[r.rid] = v
                    busy.append(s)
                    if on_launch is not None:
                        on_launch(take)
            still = []
            for s in busy:
                if s.done():
                    tdone = time.perf_counter()
                    for r, fin in zip(s.requests, s.finite()):
                        if r.rid in self._cancelled:
                            status[r.rid] = CANCELLED
                            pending.pop(r.rid, None)
                        else:
                            status[r.rid] = OK if fin else NONFINITE
                            ms = (tdone - t0 - r.arrival) * 1e3
                            lat.append(ms)
                            bucket_lat.setdefault(s.bucket, []).append(ms)
                            if keep_results:
                                v = pending.pop(r.rid)
                                if fin:
                                    results[r.rid] = v
                        done += 1
                    last = tdone
                    s.requests = []
                else:
                    still.append(s)
            busy = still
            if serial and busy and not self.fake:
                busy[0].event.synchronize()
        secs = last - t0
        a = np.asarray(lat) if lat else np.zeros(1)
        rep = Report(len(lat), secs, len(lat) / secs if secs > 0 else float("inf"),
                     float(np.percentile(a, 50)), float(np.percentile(a, 99)), float(a.mean()), offered_rate,
                     status=status, batches=batches)
        for b, v in bucket_lat.items():
            av = np.asarray(v)
            rep.per_bucket[f"{b.n}{'v' if b.vectorial else 's'}{'g' if b.grad else ''}"] = {
                "count": len(av), "p50_ms": round(float(np.percentile(av, 50)), 4),
                "p99_ms": round(float(np.percentile(av, 99)), 4)}
        rep.latencies_ms = np.asarray(lat)
        self.results = results
        return rep


def rate_sweep(server: Server, count: int = 10000, fractions=(0.25, 0.5, 0.75, 1.0, 2.0), seed: int = 7) -> dict:
    """Measures capacity (all requests at t = 0), then serves Poisson streams
    at each fraction of it; returns the summaries."""
    cap = server.run(make_requests(count, None, seed))
    out = {"capacity": cap.summary(), "rates": []}
    for f in fractions:
        rate = f * cap.throughput
        rep = server.run(make_requests(count, rate, seed), offered_rate=rate)
        d = rep.summary()
        d["fraction_of_capacity"] = f
        out["rates"].append(d)
    return out
