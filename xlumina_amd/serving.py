"""Continuous-batching request loop for RS / VRS propagation and gradient requests.

A request is one seeded random optical set-up on a square grid of side n:
  * scalar:     Gaussian beam (waist `waist`) -> SLM with a random phase
                (U[-pi, pi], seeded per request) -> RS to distance z
                (ScalarLight.RS_propagation, wave_optics.py:173-196, with the
                SLM of optical_elements.py:76-103 fused into the row pass);
  * vectorial:  Gaussian beam with Jones vector (1, 1) -> sSLM with two random
                phases (alpha on Ex, phi on Ey) -> VRS to z (Ex, Ey, Ez)
                (VectorizedLight.VRS_propagation, vectorized_optics.py:244-284,
                sSLM optical_elements.py:186-222).
Every request returns the output intensity |E|^2.  A gradient request also
returns dL/d(first phase mask) of L = sum_pixels w * |E|^2 with the fixed
radial weight w = (x^2 + y^2) / half_width^2, through the hand-written
adjoints (conjugate-kernel RS, then the sSLM/SLM and Ez-construction adjoint).

Execution model (one GPU):
  * A bucket is (n, scalar/vectorial, gradient).  Each bucket owns a pool of
    slots; a slot holds every device buffer one request needs (phase masks,
    distance, field, intensity, gradient, workspace) and a HIP graph of the
    bucket's whole kernel sequence captured over those buffers.  The beam
    and the loss weight of each grid size are shared, read-only, by all
    slots of that size.
  * Admission is by HBM footprint: slots are created while the pool total
    stays under `hbm_budget` bytes; a request waits in its bucket's queue
    until one of that bucket's slots is free.  No other request waits on it
    (per-bucket queues, no head-of-line blocking across sizes).
  * Each slot has its own stream.  Admitting a request writes its phase masks
    (seeded generator) and z into the slot on that stream, replays the graph
    and records an event; the host loop keeps admitting arrivals while
    earlier requests run, so work of different requests overlaps on the
    device.  Latency = completion seen by the host - arrival time.
The distance z is read by the kernels from device memory (z_dev), which is
what lets one captured graph serve every z.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from . import ops
from .ops import MaskSpec
from .toolbox import space

TWO_PI = 2.0 * math.pi


@dataclass
class Request:
    rid: int
    n: int
    vectorial: bool
    z: float            # propagation distance, um
    grad: bool
    seed: int
    arrival: float = 0.0  # seconds after the start of the stream


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


@dataclass
class _Shared:
    """Read-only per-grid-size state: grid, beams, loss weight."""
    x: np.ndarray
    y: np.ndarray
    dx: float
    beam_s: torch.Tensor     # (1, n, n) complex64
    beam_v: torch.Tensor     # (2, n, n) complex64, Jones (1, 1)
    w2: torch.Tensor         # (n, n) float32, 2 * w (cotangent of |E|^2 -> E)


class _Slot:
    def __init__(self, server: "Server", b: Bucket):
        dev = server.device
        n = b.n
        sh = server.shared(n)
        self.bucket = b
        self.nin = 2 if b.vectorial else 1
        self.ncomp = 3 if b.vectorial else 1
        self.src = sh.beam_v if b.vectorial else sh.beam_s
        self.params = torch.zeros((self.nin, n, n), dtype=torch.float32, device=dev)
        self.zd = torch.full((1,), 1e4, dtype=torch.float64, device=dev)
        self.out = torch.empty((self.ncomp, n, n), dtype=torch.complex64, device=dev)
        self.inten = torch.empty((n, n), dtype=torch.float32, device=dev)
        self.ws = torch.empty(ops.rs_workspace_bytes(n, self.ncomp), dtype=torch.uint8, device=dev)
        if b.grad:
            self.gout = torch.empty_like(self.out)
            self.gpre = torch.empty_like(self.out)
            self.gmask = torch.empty((n, n), dtype=torch.float32, device=dev)
        else:
            self.gout = self.gpre = self.gmask = None
        if b.vectorial:
            self.masks = [MaskSpec(N.XL_MASK_SSLM, self.params[0], self.params[1])]
            self.ez = N.XL_EZ_VRS
        else:
            self.masks = [MaskSpec(N.XL_MASK_SLM, self.params[0])]
            self.ez = N.XL_EZ_NONE
        self.sh = sh
        self.wavelength = server.wavelength
        self.stream = torch.cuda.Stream(device=dev)   # default priority
        self.event = torch.cuda.Event()
        self.graph = None
        self.request: Optional[Request] = None
        self.bytes = sum(t.numel() * t.element_size() for t in
                         (self.params, self.zd, self.out, self.inten, self.ws, self.gout, self.gpre, self.gmask)
                         if t is not None)
        self.stream.wait_stream(torch.cuda.current_stream(dev))   # buffers and shared state were made there
        with torch.cuda.stream(self.stream):
            self.program()                         # eager run: sets kernel attributes, fills caches
        self.stream.synchronize()
        if server.use_graphs:
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph, stream=self.stream):
                self.program()
            self.stream.synchronize()

    def program(self):
        """The bucket's kernel sequence over this slot's buffers."""
        sh = self.sh
        ops.rs_propagate(self.src, sh.x, sh.y, 1e4, self.wavelength, masks=self.masks, ez_mode=self.ez,
                         nin=self.nin, intensity=True, out=self.out, workspace=self.ws, z_dev=self.zd,
                         inten=self.inten)
        if self.bucket.grad:
            torch.mul(self.out, sh.w2, out=self.gout)
            ops.rs_propagate(self.gout, sh.x, sh.y, 1e4, self.wavelength, adjoint=True, out=self.gpre,
                             workspace=self.ws, z_dev=self.zd)
            ops.pre_backward(self.src, self.gpre, sh.x, sh.y, 1e4, masks=self.masks, ez_mode=self.ez, nin=self.nin,
                             need_src=False, need_params=((True, False), (False, False)), z_dev=self.zd,
                             gparams=[[self.gmask, None], [None, None]])

    def launch(self, r: Request, keep: bool):
        """Writes the request's inputs into the slot and runs it (on the slot stream)."""
        self.request = r
        with torch.cuda.stream(self.stream):
            gen = torch.Generator(device=self.params.device)
            gen.manual_seed(r.seed)
            torch.rand(self.params.shape, generator=gen, device=self.params.device, out=self.params)
            self.params.mul_(TWO_PI).sub_(math.pi)
            self.zd.fill_(r.z)
            if self.graph is not None:
                self.graph.replay()
            else:
                self.program()
            result = None
            if keep:
                result = (self.inten.clone(), self.gmask.clone() if self.gmask is not None else None)
            self.event.record(self.stream)
        return result


@dataclass
class Report:
    completed: int
    seconds: float
    throughput: float               # requests / s
    p50_ms: float
    p99_ms: float
    mean_ms: float
    offered_rate: Optional[float]
    per_bucket: Dict[str, dict] = field(default_factory=dict)
    latencies_ms: Optional[np.ndarray] = None

    def summary(self) -> dict:
        return {"completed": self.completed, "seconds": round(self.seconds, 4),
                "throughput_req_s": round(self.throughput, 2), "p50_ms": round(self.p50_ms, 4),
                "p99_ms": round(self.p99_ms, 4), "mean_ms": round(self.mean_ms, 4),
                "offered_rate_req_s": None if self.offered_rate is None else round(self.offered_rate, 2)}


class Server:
    """Slot pools per bucket under an HBM budget; see the module docstring."""

    def __init__(self, device=None, wavelength: float = 0.65, half_width: float = 1200.0, waist: float = 400.0,
                 hbm_budget: Optional[int] = None, slots_per_bucket: int = 2, use_graphs: bool = True):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
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
        self._shared: Dict[int, _Shared] = {}
        self.pools: Dict[Bucket, List[_Slot]] = {}
        self.used_bytes = 0

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
            self.used_bytes += sum(t.numel() * t.element_size() for t in (beam_s, beam_v, w2))
        return s

    def slot_bytes(self, b: Bucket) -> int:
        n2 = b.n * b.n
        nin, nc = (2, 3) if b.vectorial else (1, 1)
        by = nin * n2 * 4 + 8 + nc * n2 * 8 + n2 * 4 + ops.rs_workspace_bytes(b.n, nc)
        if b.grad:
            by += 2 * nc * n2 * 8 + n2 * 4
        return by

    def _grow(self, b: Bucket) -> Optional[_Slot]:
        pool = self.pools.setdefault(b, [])
        if len(pool) >= self.slots_per_bucket:
            return None
        if self.used_bytes + self.slot_bytes(b) > self.hbm_budget:
            return None
        s = _Slot(self, b)
        pool.append(s)
        self.used_bytes += s.bytes
        return s

    def prewarm(self, buckets) -> None:
        """Creates (and captures) the slot pools of `buckets` up front, so
        no request pays for a capture."""
        for b in buckets:
            while self._grow(b) is not None:
                pass
            if not self.pools.get(b):
                raise MemoryError(f"HBM budget {self.hbm_budget} B admits no slot of {b}")
        torch.cuda.synchronize(self.device)

    def _free_slot(self, b: Bucket) -> Optional[_Slot]:
        for s in self.pools.get(b, ()):
            if s.request is None:
                return s
        return self._grow(b)

    # -- the loop ----------------------------------------------------------
    def run(self, requests: Sequence[Request], keep_results: bool = False, serial: bool = False,
            offered_rate: Optional[float] = None) -> Report:
        """Serves `requests` (sorted by arrival) in real time.  serial=True
        admits one request at a time (each alone on the device)."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        self.prewarm({bucket_of(r) for r in reqs})
        queues: Dict[Bucket, deque] = {}
        busy: List[_Slot] = []
        lat = np.zeros(len(reqs))
        results = {} if keep_results else None
        bucket_lat: Dict[Bucket, list] = {}
        nxt = 0
        done = 0
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        last = t0
        while done < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                r = reqs[nxt]
                queues.setdefault(bucket_of(r), deque()).append(r)
                nxt += 1
            for b, q in queues.items():
                while q and not (serial and busy):
                    s = self._free_slot(b)
                    if s is None:
                        break
                    r = q.popleft()
                    res = s.launch(r, keep_results)
                    if keep_results:
                        results[r.rid] = res
                    busy.append(s)
            still = []
            for s in busy:
                if s.event.query():
                    r = s.request
                    tdone = time.perf_counter()
                    lat[done] = (tdone - t0 - r.arrival) * 1e3
                    bucket_lat.setdefault(s.bucket, []).append(lat[done])
                    done += 1
                    last = tdone
                    s.request = None
                else:
                    still.append(s)
            busy = still
            if serial and busy:
                busy[0].event.synchronize()
        secs = last - t0
        rep = Report(len(reqs), secs, len(reqs) / secs if secs > 0 else float("inf"),
                     float(np.percentile(lat, 50)), float(np.percentile(lat, 99)), float(lat.mean()), offered_rate)
        for b, v in bucket_lat.items():
            a = np.asarray(v)
            rep.per_bucket[f"{b.n}{'v' if b.vectorial else 's'}{'g' if b.grad else ''}"] = {
                "count": len(a), "p50_ms": round(float(np.percentile(a, 50)), 4),
                "p99_ms": round(float(np.percentile(a, 99)), 4)}
        rep.latencies_ms = lat
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
