#!/usr/bin/env python
"""Benchmark of the light-propagation hot path on one MI355X.

Headline workload (one step): 2048^2 radially polarised beam through a
per-pixel LCD retarder, VRS to z = 30 mm (Ex, Ey, Ez complex64), then VCZT
zoom to a 512^2 ROI of 0.2 mm -- one full VRS + VCZT chain, inputs resident
in HBM.  `value` = chains per second on the one GPU.

Also reported (`secondary`): VCZT alone in both forms (FP32 Bluestein and
split-BF16 MFMA), scalar RS 1024^2 (SLM fused, field + |E|^2)
propagations/s, high-NA 2048^2 -> 256^2 forward+backward steps/s, and the
serving loop on the mixed request stream (capacity, then Poisson arrivals at
0.25-2x capacity with p50/p99 latency; summarised in `latency`), the chain
with the split-BF16 GEMM CZT form, and (with the CPU baseline) the relative
L2 of both forms' chain outputs against the float64 oracle chain.

`roofline` covers the kernel with the largest share of the timed region,
timed live with HIP events recorded on its launch stream; `cpu_baseline` is
the float64 NumPy/SciPy restatement of the reference algorithm (the oracle)
on the host cores for a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
METRIC = "2048² VRS+CZT propagations/s & fwd+bwd steps/s vs plain FP32; p50/p99 latency"


def kernel_bytes(name, n, P, ncomp, nin, mask_params, m_out, czt_rows_in, czt_cols_in, czt_nin):
    """Algorithmic HBM bytes of one launch (streaming model: complex64 8 B,
    float32 mask 4 B, each array read or written once per pass)."""
    c8 = 8
    if name == "rs_hrow":
        return (P // 2) * P * c8
    if name == "rs_urow":
        return nin * n * n * c8 + mask_params * n * n * 4 + ncomp * n * P * c8
    if name == "rs_hcol":
        return 2 * (P // 2) * P * c8
    if name == "rs_col":
        return (P // 2) * P * c8 + 2 * ncomp * n * P * c8
    if name == "rs_final":
        return ncomp * n * P * c8 + ncomp * n * n * c8
    if name == "czt_rows":
        return czt_nin * czt_rows_in * czt_cols_in * c8 + 3 * czt_rows_in * m_out * c8
    if name == "czt_cols":
        return 3 * czt_rows_in * m_out * c8 + 3 * m_out * m_out * c8
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--fourf-launches", type=int, default=2,
                    help="timed 100-step launches of the batched 4f optimisation (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests per serving run (0 skips the serving sweep)")
    ap.add_argument("--reference-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (scale reference for HBM byte counters)")
    ap.add_argument("--workload", choices=("chain", "high_na", "fourf", "rs1024"), default="chain",
                    help="chain (default): the full benchmark; the others time one secondary workload alone "
                         "(for profiler runs) and print its JSON entry")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); several GPUs would each run "
              "whole requests / candidates with no exchange between them", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from xlumina_amd import _native as N
    from xlumina_amd import workloads

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N.ensure_device(dev)
    if args.workload != "chain":
        print(json.dumps(run_single(dev, workloads, args)), flush=True)
        return

    w = workloads.vrs_vczt_2048(dev)
    geom = w.geometry()
    n = w.field.shape[-1]
    P = 1 << (2 * n - 1).bit_length()
    for _ in range(args.warmup):
        w.run(geom)
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(args.steps):
        w.run(geom)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / args.steps
    value = 1000.0 / ms
    # per-kernel durations: HIP events recorded on the launch stream around every
    # launch, over a second timed run of the same K steps (the events put gaps
    # between launches, so the throughput above is timed without them)
    N.timing_enable(capacity=16 * (args.steps + 2))
    torch.cuda.synchronize()
    for _ in range(args.steps):
        w.run(geom)
    torch.cuda.synchronize()
    kt = N.timing_collect()
    N.timing_disable()

    # roofline of the kernel with the largest share of the timed region
    dom = max(kt.items(), key=lambda kv: kv[1][1])
    name, (launches, total_ms) = dom
    per_launch_ms = total_ms / launches
    nbytes = kernel_bytes(name, n, P, 3, 2, 2, len(w.xout), n, n, 2)
    achieved = nbytes / (per_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
                "algorithmic_bytes": nbytes, "avg_launch_us": round(per_launch_ms * 1e3, 2),
                "timing": "HIP events around each launch on its stream, second timed run of the same steps",
                "kernels_ms_per_step": {k: round(v[1] / args.steps, 4) for k, v in kt.items()}}

    if args.reference_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    _progress(f"headline: {value:.1f} chains/s ({ms:.4f} ms/step)")
    secondary = {}
    checks = {}
    if not args.no_secondary:
        secondary = run_secondary(dev, workloads, args, checks)

    cpu = None
    accuracy = None
    if not args.no_cpu_baseline:
        _progress("CPU baseline (float64 oracle chains, cold + warm, complex128 and complex64)")
        cpu, ref = cpu_baseline(w)
        accuracy = chain_accuracy(w, geom, ref)
        _progress(f"accuracy vs float64: {accuracy}")
        secondary_accuracy(secondary, checks)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "VRS+VCZT chains/s (2048^2 -> 512^2)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "vrs_vczt_2048: 2048^2 radial beam, lambda 635 nm, window 2.4 mm, per-pixel LCD "
                               "(seeded eta, theta U[0,2pi]), VRS z=30 mm, VCZT z=30 mm to 512^2 ROI of 0.2 mm",
                   "form": "plain FP32 (LDS radix-16 FFT RS + Bluestein CZT)", "grid": n, "fft_len": P},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "latency": _latency(secondary.get("serving_mix")),
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_vrs_vczt_2048.txt")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (copy-kernel scaled), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == kernel:
                    return int(parts[2])
    except OSError:
        pass
    return None


def _progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _latency(sweep):
    """p50/p99 of the mixed request stream at 0.5, 0.75 and 1.0 of the measured capacity."""
    if not sweep:
        return None
    out = {"capacity_req_s": sweep["capacity"]["throughput_req_s"]}
    for r in sweep["rates"]:
        if r["fraction_of_capacity"] in (0.5, 0.75, 1.0):
            out[f"at_{r['fraction_of_capacity']}x"] = {"p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
                                                      "throughput_req_s": r["throughput_req_s"]}
    return out


def _kernel_split(fn, steps):
    """Per-kernel ms per call of fn (HIP events around each launch)."""
    import torch
    from xlumina_amd import _native as N
    torch.cuda.synchronize()
    N.timing_enable(capacity=64 * steps)
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    kt = N.timing_collect()
    N.timing_disable()
    return {k: round(v[1] / steps, 4) for k, v in kt.items()}


def _time(fn, steps, warmup):
    """Mean ms per call over `steps` calls after `warmup` calls.  The objects
    that exist before the timed calls are collected once and then frozen out
    of Python's cyclic garbage collector, which keeps running for the objects
    the calls create: a full pass over the whole heap (several ms here) inside
    an eager, host-paced loop would otherwise leave the GPU idle and be
    charged to the kernels."""
    import gc
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    gc.collect()
    gc.freeze()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    try:
        s.record()
        for _ in range(steps):
            fn()
        e.record()
        torch.cuda.synchronize()
    finally:
        gc.unfreeze()
    return s.elapsed_time(e) / steps


def run_secondary(dev, workloads, args, checks):
    """Timings of the other workloads; the GPU outputs the float64 checks
    compare are kept in `checks` (host arrays, taken outside the timed regions)."""
    import numpy as np
    import torch
    from xlumina_amd import _native as N
    from xlumina_amd import ops
    out = {}
    # VCZT 2048^2 -> 512^2 alone in both forms (split-BF16 MFMA: 193 GFLOP of bf16 MFMA work)
    w = workloads.vrs_vczt_2048(dev)
    g = w.geometry()
    vrs, _ = w.run(g)
    for form in ("bluestein", "gemm"):
        ms = _time(lambda: ops.czt(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2, form=form), max(10, args.steps), 3)
        ent = {"value": round(1000.0 / ms, 2), "unit": "VCZT/s", "ms": round(ms, 4)}
        ent["kernels_ms"] = _kernel_split(lambda: ops.czt(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2, form=form), 10)
        if form == "gemm":
            flops = 3 * 8 * (512 * 2048 * 2048 + 512 * 512 * 2048) * 3
            ent["mfma_tflops"] = round(flops / (ms * 1e-3) / 1e12, 1)
            ent["mfma_frac_of_2500"] = round(flops / (ms * 1e-3) / 2.5e15, 4)
        out[f"vczt_2048_to_512_{form}"] = ent
        _progress(f"vczt {form}: {ms:.4f} ms")
    # the whole chain with the chosen CZT form (split-BF16 MFMA zoom-DFT)
    ms = _time(lambda: w.run(g, czt_fn=lambda v, gg, **k: ops.czt(v, gg, form="gemm", **k)), max(10, args.steps), 3)
    out["chain_2048_to_512_gemm_form"] = {"value": round(1000.0 / ms, 2), "unit": "VRS+VCZT chains/s",
                                          "ms": round(ms, 4)}
    _progress(f"chain with GEMM CZT: {ms:.4f} ms")
    # CZT form crossover: VCZT 2048^2 -> M^2 for small output windows
    cross = {}
    for m_out in (64, 128, 192, 256):
        xo = np.linspace(-100.0, 100.0, m_out)
        gm = ops.czt_geometry(w.x, w.y, xo, xo, w.z_czt, w.wavelength)
        row = {}
        for form in ("bluestein", "gemm"):
            row[form + "_ms"] = round(_time(lambda: ops.czt(vrs, gm, ez_mode=N.XL_EZ_VCZT, nin=2, form=form),
                                            max(10, args.steps // 2), 3), 4)
        cross[str(m_out)] = row
    out["vczt_2048_form_crossover"] = cross
    _progress(f"CZT form crossover: {cross}")
    rs = workloads.scalar_rs_1024(dev)
    ms = _time(lambda: rs.run(), max(10, args.steps), 3)
    out["rs_1024_slm"] = {"value": round(1000.0 / ms, 2), "unit": "propagations/s", "ms": round(ms, 4)}
    got, _ = rs.run()
    checks["rs_1024_slm"] = {"field": got[0].cpu().numpy(), "src": rs.field[0].cpu().numpy(),
                             "phase": rs.phase.cpu().numpy(), "x": rs.x, "y": rs.y, "wavelength": rs.wavelength,
                             "z": rs.z}
    # the same propagation replayed from a captured HIP graph over static buffers
    # (what the serving slots do): no per-call host work
    o_buf = torch.empty((1, 1024, 1024), dtype=torch.complex64, device=dev)
    ws = torch.empty(ops.rs_workspace_bytes(1024, 1), dtype=torch.uint8, device=dev)
    rs.run(out=o_buf, workspace=ws)
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        rs.run(out=o_buf, workspace=ws)
    torch.cuda.current_stream(dev).wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        rs.run(out=o_buf, workspace=ws)
    ms = _time(graph.replay, max(10, args.steps), 3)
    out["rs_1024_slm_graph"] = {"value": round(1000.0 / ms, 2), "unit": "propagations/s", "ms": round(ms, 4)}
    del graph, o_buf, ws
    h = workloads.high_na_2048(dev)
    geom = h.geometry()
    checks["high_na"] = {"h": h, "fields": {
        form: ops.czt(h.field, geom, masks=h.masks(), nin=2, form=form).cpu().numpy()
        for form in ("bluestein", "gemm")}}
    alpha = h.alpha.clone().requires_grad_(True)
    phi = h.phi.clone().requires_grad_(True)
    for form in ("bluestein", "gemm"):
        prev = ops.get_czt_form()
        ops.set_czt_form(form)

        def step():
            L = h.loss(alpha, phi, geom)
            L.backward()

        # warm-up covers the first calls' one-time work after the form switch
        # (workspace and split-K buffers of the caching allocator)
        ms = _time(step, max(10, args.steps // 2), 4)
        out[f"high_na_fwd_bwd_2048_to_256_{form}"] = {"value": round(1000.0 / ms, 2), "unit": "fwd+bwd steps/s",
                                                      "ms": round(ms, 4)}
        _progress(f"high-NA fwd+bwd {form}: {ms:.4f} ms")
        # the same step captured in a HIP graph (no launch or Python overhead per step)
        a_g = h.alpha.clone().requires_grad_(True)
        p_g = h.phi.clone().requires_grad_(True)
        cap = workloads.CapturedStep(lambda: h.loss(a_g, p_g, geom), [a_g, p_g])
        ms = _time(cap.replay, max(5, args.steps // 4), 2)
        out[f"high_na_fwd_bwd_2048_to_256_{form}_graph"] = {"value": round(1000.0 / ms, 2),
                                                            "unit": "fwd+bwd steps/s", "ms": round(ms, 4)}
        del cap, a_g, p_g
        ops.set_czt_form(prev)
        _progress(f"high-NA fwd+bwd {form}, graph-captured: {ms:.4f} ms")
    # hybrid optical table of experiments/hybrid_sharp_optical_table.py: 1024^2, 635 nm, 3 building
    # blocks, 9 splitters, 6 high-NA detectors of 400^2 (NA 0.9); loss softmin(small-area losses)
    table = workloads.SharpFocusTable(dev, n=1024, m=400)
    tp = table.params
    # a few warm-up steps first: the first ones also grow the allocator's pools
    ms = _time(table.step, max(10, args.steps // 10), 3)
    out["hybrid_sharp_focus_table_1024_fwd_bwd"] = {"value": round(1000.0 / ms, 2), "unit": "fwd+bwd steps/s",
                                                    "ms": round(ms, 3), "propagations_per_step": "16 VRS + 6 high-NA VCZT forward, their adjoints backward"}
    _progress(f"hybrid table fwd+bwd: {ms:.3f} ms")
    table_loss = table.loss
    cap = workloads.CapturedStep(table_loss, tp)       # the same step in one HIP graph
    ms = _time(cap.replay, max(10, args.steps // 10), 3)
    out["hybrid_sharp_focus_table_1024_fwd_bwd_graph"] = {"value": round(1000.0 / ms, 2),
                                                          "unit": "fwd+bwd steps/s", "ms": round(ms, 3)}
    del cap, tp, table
    torch.cuda.empty_cache()
    _progress(f"hybrid table fwd+bwd, graph-captured: {ms:.3f} ms")
    if args.fourf_launches > 0:
        from xlumina_amd import fourf
        opt = fourf.FourFOptimizer(dev, batch=64, n=1024, steps_per_launch=100)
        checks["fourf"] = {"opt": opt, "intensity0": opt.forward_intensity()[0].cpu().numpy(),
                           "phases0": opt.phi[0].cpu().numpy(), "einit": opt.e_in.cpu().numpy()}
        first = opt.launch().clone()          # captures the 100-step graph, then runs it
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fourf_launches):
            last = opt.launch()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.fourf_launches
        split = _kernel_split(lambda: opt.step(opt.losses[0]), 3)
        out["fourf_batched_opt_64x1024"] = {"kernels_ms_per_step": split,
            "value": round(opt.candidate_steps_per_launch() / dt, 1), "unit": "candidate-steps/s",
            "ms_per_launch": round(dt * 1e3, 2), "steps_per_launch": 100, "candidates": 64,
            "mean_loss_step0": float(first[0].mean()), "mean_loss_last": float(last[-1].mean())}
        checks["fourf"]["opt"] = None
        del opt
        torch.cuda.empty_cache()
        _progress(f"4f batched optimisation: {dt * 1e3:.1f} ms per 100-step launch")
    if args.serve_requests > 0:
        from xlumina_amd import serving as S
        srv = S.Server(dev)
        sweep = S.rate_sweep(srv, count=args.serve_requests, fractions=(0.25, 0.5, 0.75, 1.0, 2.0))
        sweep["workload"] = (f"{args.serve_requests} requests, n in 512/1024/2048 (p .5/.3/.2), scalar/vectorial "
                             "70/30, z ~ U[5, 100] mm, 20% gradient w.r.t. one phase mask, Poisson arrivals at "
                             "fractions of the measured capacity")
        sweep["hbm_pool_bytes"] = srv.used_bytes
        out["serving_mix"] = sweep
        # one scalar and one vectorial 512^2 request, kept for the float64 check
        probe = [S.Request(0, 512, False, 37000.0, False, 123456789), S.Request(1, 512, True, 52000.0, False, 987654321)]
        srv.run(probe, keep_results=True)
        checks["serving"] = {"requests": probe, "intensities": [srv.results[r.rid][0].cpu().numpy() for r in probe],
                             "phases": [S.request_phases(r, dev).cpu().numpy() for r in probe],
                             "shared": srv.shared(512), "wavelength": srv.wavelength}
        _progress(f"serving: capacity {sweep['capacity']['throughput_req_s']} req/s")
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _stats(samples):
    import numpy as np
    a = np.asarray(samples, dtype=np.float64)
    return {"n": int(a.size), "mean_s": round(float(a.mean()), 4), "p50_s": round(float(np.percentile(a, 50)), 4),
            "p99_s": round(float(np.percentile(a, 99)), 4)}


def run_single(dev, workloads, args):
    """One secondary workload alone, `args.steps` timed steps after
    `args.warmup` (the profiler runs of the committed summaries use this)."""
    from xlumina_amd import ops
    if args.workload == "high_na":
        h = workloads.high_na_2048(dev)
        geom = h.geometry()
        alpha = h.alpha.clone().requires_grad_(True)
        phi = h.phi.clone().requires_grad_(True)
        out = {"workload": "high_na_2048_to_256 forward + backward (second-moment loss)"}
        for form in ("bluestein", "gemm"):
            ops.set_czt_form(form)

            def step():
                L = h.loss(alpha, phi, geom)
                L.backward()

            ms = _time(step, args.steps, args.warmup)
            out[form] = {"ms": round(ms, 4), "steps_s": round(1000.0 / ms, 2),
                         "kernels_ms": _kernel_split(step, max(1, min(args.steps, 10)))}
        ops.set_czt_form("auto")
        return out
    if args.workload == "rs1024":
        rs = workloads.scalar_rs_1024(dev)
        ms = _time(lambda: rs.run(), args.steps, args.warmup)
        return {"workload": "scalar RS 1024^2 with SLM, field + |E|^2", "ms": round(ms, 4),
                "kernels_ms": _kernel_split(lambda: rs.run(), max(1, min(args.steps, 10)))}
    from xlumina_amd import fourf
    opt = fourf.FourFOptimizer(dev, batch=64, n=1024, steps_per_launch=100, use_graph=False)
    for _ in range(args.warmup):
        opt.step(opt.losses[0])
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        opt.step(opt.losses[0])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {"workload": "4f optimisation 64 x 1024^2, eager steps", "ms_per_step": round(dt * 1e3, 3),
            "kernels_ms": _kernel_split(lambda: opt.step(opt.losses[0]), max(1, min(args.steps, 3)))}


def cpu_baseline(w, budget_s=24.0, max_warm=4):
    """VRS + VCZT chains of the same workload through the oracle (the
    reference algorithm restated in NumPy/SciPy: 2N-1 padding, per-call
    kernel FFT, two-pass Bluestein) on the host cores, following the method of
    examples/scalar_xlumina.py:31-55 within a time budget: one cold chain,
    then warm chains alternating complex128 (the reference's precision) and
    complex64 until the budget or max_warm of each is reached; mean / p50 /
    p99 per precision.  `value` is 1 / mean of the warm complex128 chains."""
    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import xlumina_ref as R
    cores = len(os.sched_getaffinity(0))
    R.set_workers(cores)
    Ex = w.field[0].cpu().numpy().astype(np.complex128)
    Ey = w.field[1].cpu().numpy().astype(np.complex128)
    eta = w.eta.cpu().numpy().astype(np.float64)
    th = w.theta.cpu().numpy().astype(np.float64)

    def chain(dtype):
        mx, my, _ = R.LCD(Ex, Ey, None, eta, th)
        (vx, vy, _), _ = R.vrs_propagation(mx, my, w.x, w.y, w.wavelength, w.z_vrs, dtype=dtype)
        return R.vczt(vx, vy, w.x, w.y, w.wavelength, w.z_czt, w.xout, w.yout, dtype=dtype)

    t0 = time.perf_counter()
    ref = chain(np.complex128)
    cold = time.perf_counter() - t0
    warm = {"complex128": [], "complex64": []}
    start = time.perf_counter()
    while min(len(v) for v in warm.values()) < max_warm and time.perf_counter() - start < budget_s:
        for name, dt in (("complex128", np.complex128), ("complex64", np.complex64)):
            t1 = time.perf_counter()
            chain(dt)
            warm[name].append(time.perf_counter() - t1)
    mean128 = float(np.mean(warm["complex128"]))
    return ({"value": round(1.0 / mean128, 5), "unit": "VRS+VCZT chains/s", "cores": cores, "kind": "port",
             "cpu_model": _cpu_model(), "cold_s": round(cold, 3),
             "warm_complex128": _stats(warm["complex128"]), "warm_complex64": _stats(warm["complex64"]),
             "complex64_value": round(1.0 / float(np.mean(warm["complex64"])), 5),
             "sample": (f"2048^2 -> 512^2 chains: 1 cold complex128, then {len(warm['complex128'])} warm complex128 + "
                        f"{len(warm['complex64'])} warm complex64 (about {budget_s:.0f} s budget), "
                        f"scipy.fft workers={cores}")},
            np.stack(ref))


def _rel(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def secondary_accuracy(secondary, checks):
    """Relative L2 of every secondary workload's output against the float64
    oracle on the same inputs (the checker side of the CPU leg): the 1024^2 RS
    field, the high-NA focal field in both CZT forms, the 4f detector
    intensity of candidate 0 before its first update, and the intensities of
    one scalar and one vectorial 512^2 served request."""
    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import xlumina_ref as R
    c = checks.get("rs_1024_slm")
    if c and "rs_1024_slm" in secondary:
        u, _ = R.slm(c["src"].astype(np.complex128), c["phase"].astype(np.float64))
        ref, _ = R.rs_propagation(u, c["x"], c["y"], c["wavelength"], c["z"])
        e = _rel(c["field"], ref)
        secondary["rs_1024_slm"]["rel_l2_vs_float64"] = e
        if "rs_1024_slm_graph" in secondary:
            secondary["rs_1024_slm_graph"]["rel_l2_vs_float64"] = e
    c = checks.get("high_na")
    if c:
        h = c["h"]
        Ex, Ey = (h.field[i].cpu().numpy().astype(np.complex128) for i in range(2))
        a64 = h.alpha.cpu().numpy().astype(np.float64)
        p64 = h.phi.cpu().numpy().astype(np.float64)
        ex, ey, _ = R.sSLM(Ex, Ey, None, a64, p64)
        ex, ey, _ = R.LCD(ex, ey, None, h.eta, h.theta)
        ref = R.vczt_objective_lens(ex, ey, h.x, h.y, h.wavelength, h.radius, h.f, h.xout, h.yout)
        for form, got in c["fields"].items():
            e = _rel(got, ref)
            for key in (f"high_na_fwd_bwd_2048_to_256_{form}", f"high_na_fwd_bwd_2048_to_256_{form}_graph"):
                if key in secondary:
                    secondary[key]["rel_l2_vs_float64"] = e
                    secondary[key]["rel_l2_of"] = "forward focal field (3 x 256^2)"
    c = checks.get("fourf")
    if c and "fourf_batched_opt_64x1024" in secondary:
        from xlumina_amd import fourf
        x, y = fourf.space(1200.0, 1024)
        ph = c["phases0"].astype(np.float64)
        e = c["einit"].astype(np.complex128)
        a1 = R.czt(e, x, y, 0.65, 25e3)
        a2 = R.czt(R.slm(a1, ph[0])[0], x, y, 0.65, 50e3)
        a3 = R.czt(R.slm(a2, ph[1])[0], x, y, 0.65, 25e3)
        secondary["fourf_batched_opt_64x1024"]["rel_l2_vs_float64"] = _rel(c["intensity0"], np.abs(a3) ** 2)
        secondary["fourf_batched_opt_64x1024"]["rel_l2_of"] = "detector intensity of candidate 0 before step 1"
    c = checks.get("serving")
    if c and "serving_mix" in secondary:
        sh, worst = c["shared"], 0.0
        for r, got, ph in zip(c["requests"], c["intensities"], c["phases"]):
            ph = ph.astype(np.float64)
            if r.vectorial:
                bx, by = (sh.beam_v[i].cpu().numpy().astype(np.complex128) for i in range(2))
                ex, ey, _ = R.sSLM(bx, by, None, ph[0], ph[1])
                (vx, vy, vz), _ = R.vrs_propagation(ex, ey, sh.x, sh.y, c["wavelength"], r.z)
                ref = np.abs(vx) ** 2 + np.abs(vy) ** 2 + np.abs(vz) ** 2
            else:
                u, _ = R.slm(sh.beam_s[0].cpu().numpy().astype(np.complex128), ph[0])
                E, _ = R.rs_propagation(u, sh.x, sh.y, c["wavelength"], r.z)
                ref = np.abs(E) ** 2
            worst = max(worst, _rel(got, ref))
        secondary["serving_mix"]["rel_l2_vs_float64"] = worst
        secondary["serving_mix"]["rel_l2_of"] = "|E|^2 of one scalar and one vectorial 512^2 request (worst)"
    _progress("secondary accuracy vs float64: " + ", ".join(
        f"{k} {v['rel_l2_vs_float64']:.2e}" for k, v in secondary.items()
        if isinstance(v, dict) and "rel_l2_vs_float64" in v))


def chain_accuracy(w, geom, ref):
    """Relative L2 of the GPU chain output (both CZT forms) against the
    float64 oracle chain of the same inputs."""
    import numpy as np
    from xlumina_amd import ops
    out = {}
    for form in ("bluestein", "gemm"):
        _, got = w.run(geom, czt_fn=lambda v, g, **k: ops.czt(v, g, form=form, **k))
        got = got.cpu().numpy().astype(np.complex128)
        out[f"{form}_rel_l2_vs_float64"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    return out


if __name__ == "__main__":
    main()
