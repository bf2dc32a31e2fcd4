"""Benchmark: dynamics-gradient knot points/s on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

A step is one pass of the hot path over one batch of synthetic knot points
already resident in HBM.  The headline workload (`value`) is the plain per-call
kernel on one trajectory of N = 128 knots, one launch per call; the large batch
(2^20 knots, 10% stretched-arm poses) and the 64-trajectory mixed-length batch
are reported beside it under "workloads".  Knot points/s is also given as a
fraction of the HBM roofline (504 algorithmic bytes per knot: 84 in, 420 out)
and of the FP32 roofline (31,340 algorithmic FLOP per knot, the op count of
the oracle's restatement of the reference method).  The same per-call and
large-batch workloads also run on a robot read from a URDF (the synthetic arm
of tests/golden/synthetic_arm7.urdf, librbdgrad_arm7.so).

The CPU leg comes last: the CPU baseline is the oracle's FP32 restatement of
the reference's multithreaded path (time_CPU.cpp ReusableThreads loop) timed
on this host, and the FP64 oracle then checks a sample of every GPU output
recorded above ("accuracy" entries; skipped with --no-cpu).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fast-rbd-gradients_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import rbdgrad  # noqa: E402

METRIC = "dynamics-gradient knot points/s @<=1e-4 rel err vs FP64; served p50/p99 latency"
BYTES_PER_KNOT = 504          # 21 FP32 in + 105 FP32 out
FLOPS_PER_KNOT = 31340        # algorithmic count of the reference method (oracle op count)
FLOPS_PER_KNOT_ARM7 = 31464   # the same for the synthetic arm of tests/golden/synthetic_arm7.urdf
HBM_PEAK_GBS = 8000.0         # MI355X HBM3E
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (x2.000 and x1.000, the factors a device-to-device copy of known size gives):
# profiles/fdgrad_per_call_n128_and_large_batch_2p20.md.  Keyed by (kernel,
# knots per launch); other configurations report null.
MEASURED_TRAFFIC_BYTES = {
    ("fdgrad_two_wave_kernel<PrecFP32>", 128): int((181.0 * 2.000 + 53.0 * 1.000) * 1024),
    ("fdgrad_soa_kernel<PrecFP32>", 1 << 20): int((41344.7 * 2.000 + 430080.0) * 1024),
    ("fdgrad_aos_kernel<PrecFP32>", 1 << 20): int((60402.1 * 2.000 + 881842.3) * 1024),
}
FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 (packed vector / matrix) peak


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def knots(n, seed, near_singular_frac=0.0):
    rng = np.random.default_rng(seed)
    q = rng.uniform(-np.pi, np.pi, (n, 7))
    qd = rng.uniform(-2.0, 2.0, (n, 7))
    u = rng.uniform(-50.0, 50.0, (n, 7))
    if near_singular_frac > 0:
        m = rng.random(n) < near_singular_frac
        for j in (1, 3, 5):
            q[m, j] = rng.uniform(-1e-3, 1e-3, int(m.sum()))
    return np.concatenate([q, qd, u], axis=1).astype(np.float32)


def time_launches(launch, steps, warmup, stream):
    """Warm up, then time exactly `steps` launches on `stream` with HIP events,
    synchronising the device on both sides.  Returns total milliseconds."""
    for _ in range(warmup):
        launch()
    rbdgrad.hip().hipDeviceSynchronize()
    t0, t1 = rbdgrad.Event(), rbdgrad.Event()
    t0.record(stream)
    for _ in range(steps):
        launch()
    t1.record(stream)
    t1.synchronize()
    rbdgrad.hip().hipDeviceSynchronize()
    ms = t0.elapsed_ms(t1)
    t0.destroy()
    t1.destroy()
    return ms


class OracleChecks:
    """GPU outputs kept for the CPU leg, which checks them against the FP64
    oracle on the same FP32 inputs: per block and knot max|d| / max(max|ref|,
    1e-6) (max and p99.9 over knots), and the elementwise max of
    |d| / (|ref| + 1e-3 max|ref|).  A sample of
    rows is taken when the output is recorded, so large outputs are not held;
    the result is written into target[key] when the CPU leg runs."""

    def __init__(self):
        self.items = []

    def add(self, target, key, out, x, sample=None, model=None):
        idx = np.arange(len(x)) if sample is None or sample >= len(x) else \
            np.random.default_rng(5).choice(len(x), sample, replace=False)
        target[key] = {"pending": "checked against the FP64 oracle in the CPU leg"}
        self.items.append((target, key, out[idx].astype(np.float64), x[idx].astype(np.float64), model))

    def run(self):
        from oracle import pyoracle
        threads = min(16, os.cpu_count() or 1)
        for target, key, o, x, model in self.items:
            if x.shape[1] == rbdgrad.VOL_PER_KNOT:          # reference-compatible mode: 98 out, no qdd
                ref = np.concatenate([x[:, 21:28], pyoracle.fdgrad_refmode(x)], axis=1)
                o = np.concatenate([x[:, 21:28], o], axis=1)
            elif model is not None:
                with pyoracle.using_model(*model):
                    ref = pyoracle.fdgrad(x, threads=threads)
            else:
                ref = pyoracle.fdgrad(x, threads=threads)
            res = {}
            for name, sl in (("qdd", slice(0, 7)), ("dqdd_dq", slice(7, 56)), ("dqdd_dqd", slice(56, 105))):
                d, r = np.abs(o[:, sl] - ref[:, sl]), np.abs(ref[:, sl])
                rmax = np.max(r, axis=1)
                e = np.max(d, axis=1) / np.maximum(rmax, 1e-6)               # normwise per block and knot
                el = d / (r + 1e-3 * rmax[:, None] + 1e-30)                  # elementwise
                res[name] = {"max": float(e.max()), "p99.9": float(np.percentile(e, 99.9)),
                             "elementwise_max": float(el.max())}
            res["knots_checked"] = int(len(x))
            target[key] = res

    def skip(self, why):
        for target, key, *_ in self.items:
            target[key] = {"skipped": why}


CHECKS = OracleChecks()


def roofline(knots_per_launch, ms_per_launch):
    sec = ms_per_launch * 1e-3
    gbs = knots_per_launch * BYTES_PER_KNOT / sec / 1e9
    tfl = knots_per_launch * FLOPS_PER_KNOT / sec / 1e12
    return gbs, tfl


def run_flat(ctx, x, steps, warmup, stream, layout="aos"):
    n = len(x)
    if layout == "soa":
        xs = np.ascontiguousarray(x.T)
        d_in = rbdgrad.DeviceBuffer.from_array(xs)
        d_out = rbdgrad.DeviceBuffer(n * rbdgrad.OUT_PER_KNOT * 4)
        launch = lambda: ctx.fdgrad_soa(n, d_in, n, d_out, n, stream)  # noqa: E731
    else:
        d_in = rbdgrad.DeviceBuffer.from_array(x)
        d_out = rbdgrad.DeviceBuffer(n * rbdgrad.OUT_PER_KNOT * 4)
        launch = lambda: ctx.fdgrad(n, d_in, d_out, stream)  # noqa: E731
    ms = time_launches(launch, steps, warmup, stream)
    if layout == "soa":
        out = d_out.download((rbdgrad.OUT_PER_KNOT, n)).T.copy()
    else:
        out = d_out.download((n, rbdgrad.OUT_PER_KNOT))
    d_in.free()
    d_out.free()
    return ms, out


def cpu_baseline(x, seconds, vol=None):
    """The oracle's FP32 restatement of the reference's ReusableThreads loop on
    this host: 4 threads (the reference's CPU_THREADS_GLOBAL) and all usable
    cores (capped at 16, the GPU host's CPU share); the faster is reported.
    With vol (77 floats per knot) the reference's own timed configuration
    (QDD_MINV_PASSED_IN: qdd and Minv given) is timed too, at the same thread
    counts, and reported under "reference_mode"."""
    from oracle import pyoracle
    n = len(x)
    cores = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())

    def leg(inp, mode, threads, secs):
        times, _ = pyoracle.cpu_bench(inp, threads, iters=10 ** 7, mode=mode, time_cap_s=secs)
        times = times[len(times) // 10:]          # drop the first 10% (pool warm-up)
        mean_us = float(np.mean(times))
        what = "full path (q, qd, u in)" if mode == 0 else "qdd and Minv given (QDD_MINV_PASSED_IN)"
        r = {"value": n / (mean_us * 1e-6), "unit": "knots/s", "cores": threads, "kind": "port",
             "sample": f"{len(times)} calls x {n} knots (config 1 inputs), {what}, FP32, ReusableThreads-style "
                       f"pool, {threads} threads, ~{secs:.0f} s",
             "mean_us": mean_us, "std_us": float(np.std(times)),
             "p50_us": float(np.percentile(times, 50)), "p99_us": float(np.percentile(times, 99))}
        log(f"cpu baseline mode {mode} {threads} threads: {r['value']:.3e} knots/s")
        return r

    results = [leg(x, 0, t, seconds) for t in sorted({4, cores})]
    best = dict(max(results, key=lambda r: r["value"]))
    best["other"] = [r for r in results if r["cores"] != best["cores"]]
    if vol is not None:
        ref = [leg(vol, 1, t, seconds / 2) for t in sorted({4, cores})]
        best["reference_mode"] = ref
    # the reference harness's trajectory lengths, 4 threads, about 1 s each
    sweep = {}
    for m in N_SET:
        times, _ = pyoracle.cpu_bench(knots(m, seed=110 + m), 4, iters=10 ** 7, mode=0, time_cap_s=1.0)
        times = times[len(times) // 10:]
        sweep[str(m)] = {"mean_us": float(np.mean(times)), "std_us": float(np.std(times)),
                         "p50_us": float(np.percentile(times, 50)), "p99_us": float(np.percentile(times, 99)),
                         "knots_per_s": m / (float(np.mean(times)) * 1e-6), "calls": int(len(times))}
    best["over_n_4_threads"] = sweep
    log(f"cpu baseline over N (4 threads, us per call): { {k: round(v['mean_us'], 1) for k, v in sweep.items()} }")
    return best


def mixed_trajectories(seed=7, count=64):
    rng = np.random.default_rng(seed)
    return rng.integers(8, 257, size=count)


def mixed_workload(ctx, stream, steps=500):
    """Config 2: 64 trajectories with lengths U{8..256}, each in its own device
    buffers.  (a) the plain per-call kernel, one launch per trajectory; (b) the
    batched form: one persistent-kernel launch over the device work queue per
    batch, replayed from a hipGraph."""
    lens = mixed_trajectories()
    n2 = int(lens.sum())
    x2 = knots(n2, seed=102)
    offs = np.concatenate([[0], np.cumsum(lens)])
    bufs = [(rbdgrad.DeviceBuffer.from_array(x2[a:b]), rbdgrad.DeviceBuffer((b - a) * rbdgrad.OUT_PER_KNOT * 4),
             int(b - a)) for a, b in zip(offs[:-1], offs[1:])]
    res = {}

    def per_call():
        for d_in, d_out, n in bufs:
            ctx.fdgrad(n, d_in, d_out, stream)
    log(f"[c2] mixed 64 trajectories ({n2} knots): per-call kernel, 64 launches per batch")
    ms = time_launches(per_call, steps // 5, 5, stream) / (steps // 5)
    res["mixed_64_traj_per_call"] = {"knots_per_batch": n2, "launches_per_batch": 64, "ms_per_batch": ms,
                                     "knots_per_s": n2 / (ms * 1e-3)}
    log(f"[c2] {res['mixed_64_traj_per_call']}")
    batch = rbdgrad.Batch(ctx, 64)
    tl = rbdgrad.TrajList(bufs)      # the descriptor array, built once (a C caller fills 64 structs per batch)
    for use_graph in (False, True):
        log(f"[c2] batched persistent kernel, graph={use_graph}: {steps} timed batches")
        ms = time_launches(lambda: batch.run(tl, use_graph=use_graph, stream=stream), steps, 20, stream) / steps
        gbs, tfl = roofline(n2, ms)
        out = np.concatenate([d_out.download((n, rbdgrad.OUT_PER_KNOT)) for _, d_out, n in bufs])
        res["mixed_64_traj_batched" + ("_graph" if use_graph else "")] = {
            "knots_per_batch": n2, "ms_per_batch": ms, "knots_per_s": n2 / (ms * 1e-3),
            "lanes_per_knot": batch.last_lanes, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
            "fp32_tflops": tfl, "fp32_frac": tfl / FP32_PEAK_TFLOPS}
        CHECKS.add(res["mixed_64_traj_batched" + ("_graph" if use_graph else "")], "accuracy", out, x2, sample=4096)
        log(f"[c2] {res['mixed_64_traj_batched' + ('_graph' if use_graph else '')]}")
    batch.close()
    for d_in, d_out, _ in bufs:
        d_in.free()
        d_out.free()
    return res


def served_workload(ctx, rates=(1000, 3000, 10000, 30000, 100000, 300000), duration_s=1.0):
    """Config 3: Poisson request arrivals at each rate; each request one
    trajectory with length uniform over {16, 32, 64, 128, 256}; device-resident
    constants, admission by a fixed memory budget, continuous batching over
    4 in-flight slots, a polling dispatcher.  Latency = submit -> result in the
    caller's host buffer.  Three serving modes: requests' buffers in device
    memory (copies), zero-copy, and zero-copy for batches up to 16384 knots with
    copies above ("auto"); the two that copy are also run at 1M req/s, past
    what zero-copy batches sustain.  The automatic mode is run again with
    two-wave batches."""
    pool = knots(1 << 16, seed=103)
    res = {}
    # the first three with lane-group batches (16 lanes per knot); the last the
    # automatic mode with two-wave batches (32), lower latency while batches
    # are small
    for mode, zero_copy, lanes in (("device_buffers", 0, 16), ("zero_copy", 1, 16), ("auto", 2, 16),
                                   ("auto_two_wave", 2, 32)):
        srv = rbdgrad.Server(ctx, slots=4, lanes_per_knot=lanes, memory_budget_bytes=256 << 20, zero_copy=zero_copy,
                             poll_dispatch=1)
        for rate in rates + ((1000000,) if zero_copy != 1 and lanes == 16 else ()):
            log(f"[c3] served ({mode}): Poisson {rate} req/s for {duration_s} s")
            st = srv.run_poisson(rate, duration_s, [16, 32, 64, 128, 256], pool, seed=rate)
            st["batch_phases_us"] = batch_phases(srv.trace())
            st["hbm_frac"] = st["knots_per_s"] * BYTES_PER_KNOT / 1e9 / HBM_PEAK_GBS
            st["fp32_frac"] = st["knots_per_s"] * FLOPS_PER_KNOT / 1e12 / FP32_PEAK_TFLOPS
            res[f"served_{mode}_{rate}_req_per_s"] = st
            log(f"[c3] {st}")
        srv.close()
    return res


def reference_mode_workload(ctx, stream, steps, warmup):
    """The reference's own timed configuration (dynamicsGradientKernel,
    time_GPU.cu:279-316): q, qd, u, qdd and the dense Minv in (77 floats per
    knot, qdd and Minv from forward dynamics, outside the timed region), the
    98 derivative values out; one launch of 128 knots per call."""
    n = 128
    x = knots(n, seed=101)
    qdd, minv = ctx.forward_dynamics(x)
    vol = np.ascontiguousarray(np.concatenate([x, qdd, minv.transpose(0, 2, 1).reshape(n, 49)], axis=1),
                               dtype=np.float32)                  # Minv column-major, as the reference packs it
    d_vol = rbdgrad.DeviceBuffer.from_array(vol)
    d_out = rbdgrad.DeviceBuffer(n * rbdgrad.DQDD_PER_KNOT * 4)
    ms = time_launches(lambda: ctx.dynamics_gradient_kernel(n, d_out, d_vol, stream=stream), steps, warmup, stream)
    out = d_out.download((n, rbdgrad.DQDD_PER_KNOT))
    per = ms / steps
    r = {"kernel": ctx.kernel_name(n, reference_mode=True), "us_per_launch": per * 1e3,
         "knots_per_s": n / (per * 1e-3), "bytes_per_knot": 700,
         "hbm_gbs": n * 700 / (per * 1e-3) / 1e9}
    CHECKS.add(r, "accuracy", out, vol)
    d_vol.free()
    d_out.free()
    log(f"[c1] reference-compatible mode: {r}")
    return {"per_call_n128_reference_mode": r}, vol


N_SET = (1, 10, 16, 32, 64, 128, 256)   # trajectory lengths of the reference's harness (time_CPU.cpp:149-156)


def n_sweep(ctx, stream, steps, warmup):
    """Per-call latency over the reference harness's trajectory lengths (GPU;
    the CPU leg adds the same lengths on the CPU)."""
    res = {}
    for n in N_SET:
        x = knots(n, seed=110 + n)
        ms, _ = run_flat(ctx, x, steps, warmup, stream)
        res[str(n)] = {"kernel": ctx.kernel_name(n), "us_per_launch": ms / steps * 1e3,
                       "knots_per_s": n * steps / (ms * 1e-3)}
    log(f"[sweep] per call over N: { {k: round(v['us_per_launch'], 2) for k, v in res.items()} }")
    return res


def copy_bandwidth(stream, nbytes=(1 << 20) * 420, reps=20):
    """Achievable HBM bandwidth: a device-to-device copy of nbytes (read + write counted)."""
    a, b = rbdgrad.DeviceBuffer(nbytes), rbdgrad.DeviceBuffer(nbytes)
    a.fill_zero()
    H = rbdgrad.hip()

    def launch():
        rbdgrad._hipcheck(H.hipMemcpyAsync(b.handle, a.handle, nbytes, rbdgrad.D2D, stream.handle), "hipMemcpyAsync")
    ms = time_launches(launch, reps, 3, stream) / reps
    a.free()
    b.free()
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    log(f"[copy] {nbytes} bytes device to device: {ms * 1e3:.1f} us, {gbs:.0f} GB/s (read + write)")
    return {"bytes": nbytes, "us": ms * 1e3, "gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}


def urdf_model_workload(stream, steps, warmup, name="arm7",
                        urdf=os.path.join(ROOT, "tests", "golden", "synthetic_arm7.urdf")):
    """A robot read from a URDF: librbdgrad_arm7.so, compiled on the tables that
    urdf_model.py generates from the synthetic 7-joint arm (joint frames at
    general angles, so the transforms have no structural zeros).  The same
    per-call and 2^20-knot SoA workloads as for the iiwa."""
    if not os.path.exists(rbdgrad.lib_path(name)):
        log(f"[model] {rbdgrad.lib_path(name)} not built: skipped")
        return {}
    import urdf_model
    tab = urdf_model.tables(urdf_model.read_urdf(urdf))
    tables = (np.array(tab["links"]),
              np.concatenate([np.array(tab["rot"]).reshape(7, 9), np.array(tab["offset"])], axis=1))
    ctx = rbdgrad.Context(device=0, model=name)
    ctx.autotune(stream)
    res = {}
    x1 = knots(128, seed=101)
    ms, out = run_flat(ctx, x1, steps, warmup, stream)
    r = res[f"urdf_model_{name}_per_call_n128"] = {
        "kernel": ctx.kernel_name(128), "us_per_launch": ms / steps * 1e3, "knots_per_s": 128 * steps / (ms * 1e-3)}
    CHECKS.add(r, "accuracy", out, x1, model=tables)
    log(f"[model] {name} per call: {r}")
    n4 = 1 << 20
    x4 = knots(n4, seed=104)
    ms, out = run_flat(ctx, x4, 20, 3, stream, "soa")
    per = ms / 20
    gbs, _ = roofline(n4, per)
    tfl = n4 * FLOPS_PER_KNOT_ARM7 / (per * 1e-3) / 1e12
    r = res[f"urdf_model_{name}_2p20_fp32_soa"] = {
        "knots_per_launch": n4, "ms_per_launch": per, "knots_per_s": n4 / (per * 1e-3), "hbm_gbs": gbs,
        "hbm_frac": gbs / HBM_PEAK_GBS, "fp32_tflops": tfl, "fp32_frac": tfl / FP32_PEAK_TFLOPS,
        "flops_per_knot": FLOPS_PER_KNOT_ARM7}
    CHECKS.add(r, "accuracy", out, x4, sample=8192, model=tables)
    log(f"[model] {name} 2^20 SoA: {r}")
    ctx.close()
    return res


def batch_phases(tr):
    """Where a served batch's time goes, from the server's per-batch trace:
    queued (oldest request's submit -> admission), enqueue (admission -> all
    device work enqueued), device (enqueued -> event seen), deliver (event ->
    results in the callers' buffers); p50/p99/max of each, and the phases of
    the batch with the longest submit-to-delivery time."""
    if len(tr["admit_us"]) == 0:
        return {}
    phases = {"queued": tr["admit_us"] - tr["oldest_submit_us"],
              "enqueue": tr["enqueued_us"] - tr["admit_us"],
              "device": tr["done_us"] - tr["enqueued_us"],
              "deliver": tr["delivered_us"] - tr["done_us"]}
    out = {k: {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)), "max": float(v.max())}
           for k, v in phases.items()}
    worst = int(np.argmax(tr["delivered_us"] - tr["oldest_submit_us"]))
    out["worst_batch"] = {k: float(v[worst]) for k, v in phases.items()}
    out["worst_batch"].update(requests=int(tr["requests"][worst]), knots=int(tr["knots"][worst]),
                              at_us=float(tr["oldest_submit_us"][worst] - tr["oldest_submit_us"][0]))
    return out


def profile_one(args, stream):
    """One workload, for rocprofv3 runs (kernel trace / counters)."""
    if args.only == "copy":
        # a plain device-to-device copy of known size, to calibrate FETCH_SIZE/WRITE_SIZE
        nbytes = (1 << 20) * rbdgrad.OUT_PER_KNOT * 4
        a, b = rbdgrad.DeviceBuffer(nbytes), rbdgrad.DeviceBuffer(nbytes)
        a.fill_zero()
        for _ in range(args.steps):
            rbdgrad._hipcheck(rbdgrad.hip().hipMemcpy(b.handle, a.handle, nbytes, rbdgrad.D2D), "hipMemcpy D2D")
        rbdgrad.hip().hipDeviceSynchronize()
        log(f"copy: {args.steps} x {nbytes} bytes device to device")
        return
    if args.only == "c2":
        ctx = rbdgrad.Context(device=0, precision=args.precision, model=args.model)
        res = mixed_workload(ctx, stream, steps=args.steps)
        log(json.dumps(res))
        return
    if args.only == "c3":
        ctx = rbdgrad.Context(device=0, precision=args.precision, model=args.model)
        if args.rate:
            # one automatic-mode run at one rate and mapping (profiler runs)
            srv = rbdgrad.Server(ctx, slots=4, lanes_per_knot=args.server_lanes, memory_budget_bytes=256 << 20,
                                 zero_copy=2, poll_dispatch=1)
            st = srv.run_poisson(args.rate, 1.0, [16, 32, 64, 128, 256], knots(1 << 16, seed=103), seed=args.rate)
            st["batch_phases_us"] = batch_phases(srv.trace())
            srv.close()
            log(json.dumps(st))
            return
        res = served_workload(ctx)
        log(json.dumps(res))
        return
    lanes = 1 if args.only in ("c1-one-lane", "c4-aos", "c4-soa") else (16 if args.only == "c1-lane-group" else 0)
    ctx = rbdgrad.Context(device=0, precision=args.precision, lanes_per_knot=lanes, model=args.model)
    if args.only.startswith("c1"):
        n, x, layout = 128, knots(128, seed=101), "aos"
    else:
        n, layout = 1 << 20, args.only[3:]
        x = knots(n, seed=104, near_singular_frac=0.1)
    ms, _ = run_flat(ctx, x, args.steps, min(args.warmup, 10), stream, layout)
    log(f"{args.only} ({ctx.model}, {layout}): {ctx.kernel_name(n)} mapping, n={n}: "
        f"{ms / args.steps * 1e3:.3f} us per launch over {args.steps}")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed launches of the per-call workload")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--precision", default="fp32", choices=list(rbdgrad.PRECISIONS))
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="CPU baseline time per thread count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip the large-batch and mixed workloads")
    ap.add_argument("--only", choices=["c1", "c1-one-lane", "c1-lane-group", "c4-aos", "c4-soa", "c2", "c3", "copy"],
                    help="run one workload only and print its timing (for profiler runs)")
    ap.add_argument("--rate", type=int, default=0,
                    help="with --only c3: one automatic-mode served run at this many requests/s")
    ap.add_argument("--server-lanes", type=int, default=16, choices=[16, 32],
                    help="with --only c3 --rate: the server's mapping (16: lane groups, 32: two waves)")
    ap.add_argument("--model", default=None,
                    help="with --only: the robot library librbdgrad_MODEL.so (default: the iiwa 14)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "knot points are independent, so N GPUs would each run their own share of knots", file=sys.stderr)
        sys.exit(2)
    if rbdgrad.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    stream = rbdgrad.Stream()
    if args.only:
        profile_one(args, stream)
        return
    ctx = rbdgrad.Context(device=0, precision=args.precision)
    threshold, tune_table = ctx.autotune(stream)
    log(f"autotune: two-wave mapping up to {ctx.two_wave_threshold} knots, lane groups up to {threshold}; "
        f"table (n: us one-lane, us lane-group, us two-wave) "
        f"{ {k: tuple(round(t, 2) for t in v) for k, v in tune_table.items()} }")
    bthreshold, btune_table = ctx.autotune_batch(stream)
    log(f"autotune (batched): lane-group mapping up to {bthreshold} knots; table (total knots: us one-lane, "
        f"us lane-group) { {k: (round(a, 2), round(b, 2)) for k, (a, b) in btune_table.items()} }")

    # ---- headline: per-call kernel, one trajectory of 128 knots per launch ----
    n1 = 128
    x1 = knots(n1, seed=101)
    log(f"[c1] per-call N={n1}: {args.warmup} warmup + {args.steps} timed launches")
    ms1, out1 = run_flat(ctx, x1, args.steps, args.warmup, stream)
    ms_step = ms1 / args.steps
    value = n1 * args.steps / (ms1 * 1e-3)
    gbs1, tfl1 = roofline(n1, ms_step)
    checked = {}
    CHECKS.add(checked, "accuracy", out1, x1)
    kname = ctx.kernel_name(n1)
    log(f"[c1] {kname}: {value:.3e} knots/s, {ms_step * 1e3:.2f} us/launch")

    workloads = {"autotune": {"group_threshold_knots": threshold,
                              "two_wave_threshold_knots": ctx.two_wave_threshold,
                              "us_per_launch": {str(k): {"one_lane": a, "lane_group": b, "two_wave": c}
                                                for k, (a, b, c) in tune_table.items()},
                              "batch_group_threshold_knots": bthreshold,
                              "batch_us_per_batch": {str(k): {"one_lane": a, "lane_group": b}
                                                     for k, (a, b) in btune_table.items()}}}
    # the same per-call workload with the mapping forced to one lane per knot
    one = rbdgrad.Context(device=0, precision=args.precision, lanes_per_knot=1)
    ms1b, _ = run_flat(one, x1, args.steps // 4, args.warmup, stream)
    workloads["per_call_n128_one_lane_per_knot"] = {
        "kernel": one.kernel_name(n1), "us_per_launch": ms1b / (args.steps // 4) * 1e3,
        "knots_per_s": n1 * (args.steps // 4) / (ms1b * 1e-3)}
    log(f"[c1] one lane per knot: {workloads['per_call_n128_one_lane_per_knot']}")
    one.close()
    # ... and to one 16-lane group per knot in one wave
    grp = rbdgrad.Context(device=0, precision=args.precision, lanes_per_knot=16)
    ms1c, _ = run_flat(grp, x1, args.steps // 4, args.warmup, stream)
    workloads["per_call_n128_lane_group"] = {
        "kernel": grp.kernel_name(n1), "us_per_launch": ms1c / (args.steps // 4) * 1e3,
        "knots_per_s": n1 * (args.steps // 4) / (ms1c * 1e-3)}
    log(f"[c1] one-wave lane group: {workloads['per_call_n128_lane_group']}")
    grp.close()
    # ... and with the hardware sin/cos precision mode
    hw = rbdgrad.Context(device=0, precision="fp32_hw_trig")
    hw.autotune(stream)
    ms1d, out1d = run_flat(hw, x1, args.steps // 4, args.warmup, stream)
    r = workloads["per_call_n128_fp32_hw_trig"] = {
        "kernel": hw.kernel_name(n1), "us_per_launch": ms1d / (args.steps // 4) * 1e3,
        "knots_per_s": n1 * (args.steps // 4) / (ms1d * 1e-3)}
    CHECKS.add(r, "accuracy", out1d, x1)
    log(f"[c1] hardware sin/cos: {r}")
    hw.close()
    ref_mode, vol1 = reference_mode_workload(ctx, stream, args.steps // 4, args.warmup)
    workloads.update(ref_mode)
    workloads["per_call_over_n"] = n_sweep(ctx, stream, args.steps // 4, args.warmup)
    if not args.quick:
        # ---- large batch: 2^20 knots, 10% stretched-arm poses ----
        n4 = 1 << 20
        x4 = knots(n4, seed=104, near_singular_frac=0.1)
        steps4 = 20
        for prec in ("fp32", "fp64_columns", "fp32_hw_trig"):
            c4 = ctx if prec == args.precision else rbdgrad.Context(device=0, precision=prec)
            for layout in (("aos", "soa") if prec != "fp32_hw_trig" else ("soa",)):
                log(f"[c4] 2^20 knots prec={prec} layout={layout}: 3 warmup + {steps4} timed")
                ms4, out4 = run_flat(c4, x4, steps4, 3, stream, layout)
                per = ms4 / steps4
                gbs, tfl = roofline(n4, per)
                workloads[f"large_batch_2p20_{prec}_{layout}"] = {
                    "knots_per_launch": n4, "ms_per_launch": per, "knots_per_s": n4 / (per * 1e-3),
                    "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS, "fp32_tflops": tfl,
                    "fp32_frac": tfl / FP32_PEAK_TFLOPS,
                    "traffic_bytes": MEASURED_TRAFFIC_BYTES.get(
                        (f"fdgrad_{layout}_kernel<PrecFP32>", n4)) if prec == "fp32" else None}
                CHECKS.add(workloads[f"large_batch_2p20_{prec}_{layout}"], "accuracy", out4, x4, sample=8192)
                log(f"[c4] {workloads[f'large_batch_2p20_{prec}_{layout}']}")
            if c4 is not ctx:
                c4.close()
        workloads.update(mixed_workload(ctx, stream))
        workloads.update(served_workload(ctx))
        workloads.update(urdf_model_workload(stream, args.steps // 4, args.warmup))
        workloads["hbm_copy"] = copy_bandwidth(stream)

    # ---- CPU leg: the oracle as the CPU baseline, and as the checker of the GPU outputs ----
    if args.no_cpu:
        cpu = None
        CHECKS.skip("--no-cpu")
    else:
        cpu = cpu_baseline(x1, args.cpu_seconds, vol1)
        log("checking the GPU outputs against the FP64 oracle")
        CHECKS.run()
    acc1 = checked["accuracy"]

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "knots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic: seeded uniform q~U(-pi,pi), qd~U(-2,2), u~U(-50,50); iiwa 14 model constants",
        "config": {"workload": "per-call kernel, one trajectory of 128 knot points per launch (config 1)",
                   "knots_per_call": n1, "layout": "aos", "precision": args.precision, "kernel": kname,
                   "lane_group_threshold": threshold, "two_wave_threshold": ctx.two_wave_threshold},
        "roofline": {"bound": "hbm", "achieved": gbs1, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs1 / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC_BYTES.get((kname, n1)),
                     "bytes_per_knot": BYTES_PER_KNOT, "knots_per_launch": n1},
        "roofline_fp32": {"bound": "fp32-valu", "achieved": tfl1, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": tfl1 / FP32_PEAK_TFLOPS, "flops_per_knot": FLOPS_PER_KNOT},
        "accuracy": acc1,
        "workloads": workloads,
        "cpu_baseline": cpu,
    }
    ctx.close()
    stream.destroy()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
