"""Benchmark: dynamics-gradient knot points/s on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--full]

A step is one pass of the hot path over one batch of synthetic knot points
already resident in HBM.  The headline workload (`value`) is the large batch:
2^20 knot points, 10% of them at the stretched-arm pose (joints 2, 4, 6 within
1e-3 rad of 0), one launch per step, FP32 arithmetic, structure-of-arrays
layout.  Every one of its 2^20 outputs is checked against the FP64 oracle on
the same FP32 inputs; should the FP32 mode break the 1e-4 bar the headline
moves to the fastest precision mode that holds it, and says so.  Beside it the
same knots run in the other precision modes (FP64 pivots, FP64 derivative
columns, hardware sin/cos) and in the array-of-structures layout, and through
the library's other entry points (the reference-compatible mode, the full path
with Minv, the WAFR layout, external forces, forward dynamics alone).

Knot points/s is given as a fraction of the FP32 roofline (31,340 algorithmic
FLOP per knot, the op count of the oracle's restatement of the reference
method; 157.3 TFLOP/s) and of the HBM roofline (504 algorithmic bytes per
knot: 84 in, 420 out; 8 TB/s).  Secondary workloads, reported in a few short
fields: the per-call kernel on one trajectory of 128 knots, the 64-trajectory
mixed-length batch (direct launch and graph replay), served Poisson request
streams, the reference harness's per-call latency statistics over its
trajectory lengths (when fast-rbd-gradients_amd/harness/time_gpu is built).
The full record of every workload goes to a JSON file (--detail); stdout gets
one JSON line.

The CPU leg comes last: the CPU baseline is the oracle's FP32 restatement of
the reference's multithreaded path (time_CPU.cpp ReusableThreads loop) timed
on this host on a sample of the same large-batch knots, and the FP64 oracle
then checks the GPU outputs recorded above (skipped with --no-cpu).
"""
import argparse
import json
import os
import subprocess
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
FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 (packed vector / matrix) peak
REL_TOL = 1e-4                # the accuracy bar: max relative error per output block vs the FP64 oracle
LARGE = 1 << 20               # knots of the large batch
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (x2.000 and x1.000, the factors a device-to-device copy of known size gives;
# KiB per dispatch): profiles/fdgrad_large_batch_2p20_soa_aos.md and
# profiles/fdgrad_per_call_n128.md.  Keyed by (kernel, knots per launch);
# other configurations report null.
MEASURED_TRAFFIC_BYTES = {
    ("fdgrad_two_wave_kernel<PrecFP32>", 128): int((176.0 * 2.000 + 53.0 * 1.000) * 1024),
    ("fdgrad_soa_kernel<PrecFP32>", LARGE): int((41326.1 * 2.000 + 430108.3) * 1024),
    ("fdgrad_aos_kernel<PrecFP32>", LARGE): int((43356.5 * 2.000 + 430334.7) * 1024),
}
# Executed work of the one-lane FP32 kernel per 64-knot block (one knot per
# lane), from rocprofv3 --pmc SQ_INSTS_VALU and SQ_INSTS_VALU_FLOPS_FP32 at 2^20
# knots (profiles/fdgrad_large_batch_2p20_soa_aos.md): VALU instructions, and
# FP32 FLOPs per knot (an FMA counts 2).  The algorithmic count above is the
# reference method's, whose dense 6x6 inertia products include structural
# zeros the kernel never computes, so its roofline fraction can pass 1.
EXECUTED_VALU_PER_KNOT = 6599
EXECUTED_FLOPS_PER_KNOT = 10900
N_SET = (1, 10, 16, 32, 64, 128, 256)   # trajectory lengths of the reference's harness (time_GPU.cu:344-351)
HARNESS = os.path.join(PKG, "harness", "time_gpu")


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


# Untimed work before every timed region: from idle, the GPU takes tens of
# milliseconds to reach its steady clocks (a first 11 ms of 2^20-knot launches
# measured 227 us per launch against 190-195 us after), and each workload here
# starts after host work (uploads, downloads, oracle checks) that idles it.
SETTLE_MS = 100.0


def time_launches(launch, steps, warmup, stream):
    """Warm up (`warmup` launches, then more until SETTLE_MS of wall time has
    passed), then time exactly `steps` launches on `stream` with HIP events
    recorded on that stream, synchronising the device on both sides.  Returns
    total milliseconds."""
    t_warm = time.perf_counter()
    for _ in range(warmup):
        launch()
    rbdgrad.hip().hipDeviceSynchronize()
    while (time.perf_counter() - t_warm) * 1e3 < SETTLE_MS:
        for _ in range(8):
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


GRADIENT_BLOCKS = (("qdd", slice(0, 7)), ("dqdd_dq", slice(7, 56)), ("dqdd_dqd", slice(56, 105)))
ORACLE_BLOCKS = {"fdgrad": GRADIENT_BLOCKS,
                 "fdgrad_minv": GRADIENT_BLOCKS + (("minv", slice(105, 154)),),   # + dqdd/du, column-major
                 "fd": (("qdd", slice(0, 7)), ("minv", slice(7, 56))),
                 "cdu": (("dtau_dq", slice(0, 49)), ("dtau_dqd", slice(49, 98))),
                 "dqdd": (("dqdd_dq", slice(0, 49)), ("dqdd_dqd", slice(49, 98))),
                 "vaf": (("v", slice(28, 70)), ("a", slice(70, 112)), ("f", slice(112, 154)))}


def error_stats(o, ref, blocks=GRADIENT_BLOCKS):
    """Per output block: the normwise error per knot max|d| / max(max|ref|, 1e-6),
    its max and p99.9 over knots, and the elementwise max of |d| / (|ref| +
    1e-3 max|ref|)."""
    res = {}
    for name, sl in blocks:
        if sl.stop > ref.shape[1]:
            continue
        d = np.abs(o[:, sl].astype(np.float64) - ref[:, sl])
        r = np.abs(ref[:, sl])
        rmax = np.max(r, axis=1)
        e = np.max(d, axis=1) / np.maximum(rmax, 1e-6)
        el = d / (r + 1e-3 * rmax[:, None] + 1e-30)
        res[name] = {"max": float(e.max()), "p99.9": float(np.percentile(e, 99.9)), "elementwise_max": float(el.max())}
    return res


class OracleChecks:
    """GPU outputs kept for the CPU leg, which checks them against the FP64
    oracle on the same FP32 inputs (error_stats).  `sample` rows are taken when
    the output is recorded, so large outputs need not be held; sample=None keeps
    every row.  Outputs of one input array share one oracle run.  The result is
    written into target[key] when the CPU leg runs."""

    def __init__(self):
        self.items = []

    def add(self, target, key, out, x, sample=None, model=None, kind="fdgrad", fext=None):
        """kind: "fdgrad" (105 per knot: qdd, dqdd/dq, dqdd/dqd), "fdgrad_minv"
        (those and Minv, 154), "fd" (qdd and Minv, 56), "dqdd" (98, no qdd),
        "cdu" (the inverse-dynamics gradient, 98, of x = q, qd, u, qdd), "vaf"
        (the split layout, 154, of the same x); Minv column-major.
        fext: the per-knot external forces (n, 42) the outputs were computed with."""
        if sample is None or sample >= len(x):
            o, xi, fi = out, x, fext
        else:
            idx = np.random.default_rng(5).choice(len(x), sample, replace=False)
            o, xi, fi = out[idx], x[idx], None if fext is None else fext[idx]
        target[key] = {"pending": "checked against the FP64 oracle in the CPU leg"}
        self.items.append((target, key, o, xi, model, kind, fi))

    def run(self, threads):
        from oracle import pyoracle
        refs = {}
        for target, key, o, x, model, kind, fext in self.items:
            rk = (id(x), None if model is None else id(model[0]), kind, None if fext is None else id(fext))
            if rk not in refs:
                if x.shape[1] == rbdgrad.VOL_PER_KNOT:          # reference-compatible mode: 98 out, no qdd
                    refs[rk] = np.concatenate([x[:, 21:28].astype(np.float64),
                                               pyoracle.fdgrad_refmode(x.astype(np.float64))], axis=1)
                elif kind == "cdu":                                 # x: q, qd, u, qdd (28 per knot)
                    refs[rk] = pyoracle.id_gradient(x.astype(np.float64))
                elif kind == "vaf":
                    refs[rk] = pyoracle.vaf(np.concatenate([x.astype(np.float64), np.zeros((len(x), 126))], axis=1))
                elif kind == "dqdd":                                # x: q, qd, u; the gradient without qdd
                    refs[rk] = pyoracle.fdgrad(x.astype(np.float64), threads=threads)[:, 7:]
                elif kind == "fd":
                    fd = [pyoracle.fd(k[:7], k[7:14], k[14:21]) for k in x.astype(np.float64)]
                    refs[rk] = np.array([np.concatenate([qdd, m.T.reshape(49)]) for qdd, m in fd])
                elif fext is not None:
                    refs[rk] = pyoracle.fdgrad(x.astype(np.float64), fext=fext.astype(np.float64))
                elif model is not None:
                    with pyoracle.using_model(*model):
                        refs[rk] = pyoracle.fdgrad(x.astype(np.float64), threads=threads)
                else:
                    refs[rk] = pyoracle.fdgrad(x.astype(np.float64), threads=threads)
                if kind == "fdgrad_minv":
                    minv = [pyoracle.fd(k[:7], k[7:14], k[14:21])[1].T.reshape(49) for k in x.astype(np.float64)]
                    refs[rk] = np.concatenate([refs[rk], np.array(minv)], axis=1)
            ref = refs[rk]
            if x.shape[1] == rbdgrad.VOL_PER_KNOT:
                o = np.concatenate([x[:, 21:28], o], axis=1)
            res = error_stats(o, ref, ORACLE_BLOCKS[kind])
            res["knots_checked"] = int(len(x))
            target[key] = res
        self.items = []

    def skip(self, why):
        for target, key, *_ in self.items:
            target[key] = {"skipped": why}


CHECKS = OracleChecks()


def roofline(knots_per_launch, ms_per_launch, flops_per_knot=FLOPS_PER_KNOT):
    sec = ms_per_launch * 1e-3
    gbs = knots_per_launch * BYTES_PER_KNOT / sec / 1e9
    tfl = knots_per_launch * flops_per_knot / sec / 1e12
    return gbs, tfl


def run_flat(ctx, x, steps, warmup, stream, layout="aos"):
    n = len(x)
    if layout == "soa":
        d_in = rbdgrad.DeviceBuffer.from_array(np.ascontiguousarray(x.T))
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


# ---------------------------------------------------------------------------
# Headline: the large batch
# ---------------------------------------------------------------------------
LARGE_LEGS = (("fp32", "soa"), ("fp32", "aos"), ("fp64_pivot", "soa"), ("fp64_columns", "soa"),
              ("fp32_hw_trig", "soa"))


def large_batch(stream, steps, warmup):
    """2^20 knots, 10% stretched-arm poses, one launch per step, in every
    precision mode (structure-of-arrays) and FP32 array-of-structures.  Every
    output is kept for the full oracle check of the CPU leg."""
    x4 = knots(LARGE, seed=104, near_singular_frac=0.1)
    legs = {}
    for prec, layout in LARGE_LEGS:
        ctx = rbdgrad.Context(device=0, precision=prec, lanes_per_knot=1)
        ms, out = run_flat(ctx, x4, steps, warmup, stream, layout)
        per = ms / steps
        gbs, tfl = roofline(LARGE, per)
        kname = ctx.kernel_name(LARGE).replace("_aos_", f"_{layout}_")   # one lane per knot, this layout
        r = legs[f"{prec}_{layout}"] = {
            "precision": prec, "layout": layout, "kernel": kname, "steps": steps, "ms_per_launch": per,
            "knots_per_s": LARGE / (per * 1e-3), "fp32_tflops": tfl, "fp32_frac": tfl / FP32_PEAK_TFLOPS,
            "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
            "traffic_bytes": MEASURED_TRAFFIC_BYTES.get((kname, LARGE))}
        CHECKS.add(r, "accuracy", out, x4)
        log(f"[large] {prec} {layout}: {per * 1e3:.1f} us/launch, {r['knots_per_s']:.3e} knots/s, "
            f"FP32 frac {r['fp32_frac']:.3f}")
        ctx.close()
    return legs, x4


# The library's other entry points, each a one-lane kernel over the same knots:
# (name, floats read per knot, floats written per knot).  External forces are
# one set of 42 per knot.
OTHER_ENTRIES = (("reference_mode", 77, 98), ("with_minv", 21, 105 + 49), ("wafr_layout", 21, 7 + 147),
                 ("external_forces", 21 + 42, 105), ("forward_dynamics", 21, 7 + 49),
                 # the reference's staged variants: inverse-dynamics gradient from the fused
                 # (q, qd, u, qdd) and split (+ v, a, f) layouts, v, a, f written into the
                 # split layout, and the finish stage -Minv cdu
                 ("fused_cdu", 28, 98), ("split_cdu", 154, 98), ("setup_split", 28, 126),
                 ("finish", 49 + 98, 98))
ORACLE_SAMPLE = 4096   # knots of each of these outputs checked against the FP64 oracle


def other_entries(stream, steps, warmup, x):
    """The other entry points on the large batch's knots, on device-resident
    buffers, one launch per step: the reference-compatible mode (q, qd, u,
    qdd and Minv in, 98 out), the full path with dqdd/du = Minv, the WAFR
    output layout, external forces, and forward dynamics alone.  Each is
    timed as the headline is and priced against HBM by the floats it reads
    and writes; a sample of its outputs goes to the oracle check."""
    n = len(x)
    ctx = rbdgrad.Context(device=0, precision="fp32", lanes_per_knot=1)
    d_in = rbdgrad.DeviceBuffer.from_array(x)
    d_qdd = rbdgrad.DeviceBuffer(n * 7 * 4)
    d_minv = rbdgrad.DeviceBuffer(n * 49 * 4)
    ctx.forward_dynamics_dev(n, d_in, d_qdd, d_minv, stream)
    rbdgrad.hip().hipDeviceSynchronize()
    qdd, minv = d_qdd.download((n, 7)), d_minv.download((n, 49))
    vol = np.ascontiguousarray(np.concatenate([x, qdd, minv], axis=1), dtype=np.float32)   # Minv column-major
    fext = np.random.default_rng(105).uniform(-20.0, 20.0, (n, 42)).astype(np.float32)
    d_vol = rbdgrad.DeviceBuffer.from_array(vol)
    d_fext = rbdgrad.DeviceBuffer.from_array(fext)
    d_out = rbdgrad.DeviceBuffer(n * 147 * 4)
    vol28 = np.ascontiguousarray(vol[:, :28])
    d_vol28 = rbdgrad.DeviceBuffer.from_array(vol28)
    split = np.zeros((n, 154), np.float32)
    split[:, :28] = vol28
    d_split = rbdgrad.DeviceBuffer.from_array(split)
    d_cdu = rbdgrad.DeviceBuffer(n * 98 * 4)
    launches = {
        "reference_mode": lambda: ctx.dynamics_gradient_kernel(n, d_out, d_vol, stream=stream),
        "with_minv": lambda: ctx.fdgrad_minv(n, d_in, d_out, d_minv, stream),
        "wafr_layout": lambda: ctx.fdgrad_wafr(n, d_in, d_qdd, d_out, stream),
        "external_forces": lambda: ctx.fdgrad_fext(n, d_in, d_fext, 42, d_out, stream),
        "forward_dynamics": lambda: ctx.forward_dynamics_dev(n, d_in, d_qdd, d_minv, stream),
        "fused_cdu": lambda: ctx.dynamics_gradient_kernel_fused(n, d_cdu, d_vol28, stream),
        "setup_split": lambda: ctx.gradient_setup_split_dev(n, d_split, stream),
        "split_cdu": lambda: ctx.dynamics_gradient_kernel_split(n, d_cdu, d_split, stream),
        "finish": lambda: ctx.gradient_finish_dev(n, d_out, d_cdu, d_minv, stream),
    }
    order = ("reference_mode", "with_minv", "wafr_layout", "external_forces", "forward_dynamics", "fused_cdu",
             "setup_split", "split_cdu", "finish")     # setup before split; forward dynamics leaves Minv for finish
    sizes = {name: (fin, fout) for name, fin, fout in OTHER_ENTRIES}
    res = {"knots": n, "oracle_sample": ORACLE_SAMPLE}
    for name in order:
        fin, fout = sizes[name]
        ms = time_launches(launches[name], steps, warmup, stream)
        per = ms / steps
        gbs = n * (fin + fout) * 4 / (per * 1e-3) / 1e9
        r = res[name] = {"us_per_launch": per * 1e3, "knots_per_s": n / (per * 1e-3),
                         "bytes_per_knot": (fin + fout) * 4, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        # outputs of the last launch, in the oracle's layouts
        if name == "reference_mode":
            CHECKS.add(r, "accuracy", d_out.download((n, rbdgrad.DQDD_PER_KNOT)), vol, sample=ORACLE_SAMPLE)
        elif name == "with_minv":
            o = np.concatenate([d_out.download((n, rbdgrad.OUT_PER_KNOT)), d_minv.download((n, 49))], axis=1)
            CHECKS.add(r, "accuracy", o, x, sample=ORACLE_SAMPLE, kind="fdgrad_minv")
        elif name == "wafr_layout":
            o = np.concatenate([d_qdd.download((n, 7)), d_out.download((n, 147))], axis=1)
            CHECKS.add(r, "accuracy", o, x, sample=ORACLE_SAMPLE, kind="fdgrad_minv")
        elif name == "external_forces":
            CHECKS.add(r, "accuracy", d_out.download((n, rbdgrad.OUT_PER_KNOT)), x, sample=ORACLE_SAMPLE,
                       fext=fext)
        elif name == "forward_dynamics":
            o = np.concatenate([d_qdd.download((n, 7)), d_minv.download((n, 49))], axis=1)
            CHECKS.add(r, "accuracy", o, x, sample=ORACLE_SAMPLE, kind="fd")
        elif name in ("fused_cdu", "split_cdu"):
            CHECKS.add(r, "accuracy", d_cdu.download((n, 98)), vol28, sample=ORACLE_SAMPLE, kind="cdu")
        elif name == "setup_split":
            CHECKS.add(r, "accuracy", d_split.download((n, 154)), vol28, sample=ORACLE_SAMPLE, kind="vaf")
        else:
            CHECKS.add(r, "accuracy", d_out.download((n, 98)), x, sample=ORACLE_SAMPLE, kind="dqdd")
        log(f"[other entries] {name}: {per * 1e3:.1f} us/launch, {r['knots_per_s']:.3e} knots/s, "
            f"HBM frac {r['hbm_frac']:.3f}")
    for b in (d_in, d_qdd, d_minv, d_vol, d_fext, d_out, d_vol28, d_split, d_cdu):
        b.free()
    ctx.close()
    return res


def choose_headline(legs):
    """The FP32 structure-of-arrays leg if it holds the bar on every checked
    knot, else the fastest leg that does (None when none was checked)."""
    def ok(r):
        acc = r.get("accuracy", {})
        return "knots_checked" in acc and max(acc[b]["max"] for b in ("qdd", "dqdd_dq", "dqdd_dqd")) <= REL_TOL
    if ok(legs["fp32_soa"]):
        return "fp32_soa"
    passing = [k for k, r in legs.items() if ok(r)]
    return max(passing, key=lambda k: legs[k]["knots_per_s"]) if passing else None


# ---------------------------------------------------------------------------
# Secondary workloads
# ---------------------------------------------------------------------------
def per_call(ctx, stream, steps, warmup):
    """One trajectory of 128 knots per launch: the mapping the context picks
    (two waves per knot group), one lane per knot, one-wave lane groups, the
    hardware sin/cos mode and the reference's own timed configuration."""
    n1 = 128
    x1 = knots(n1, seed=101)
    res = {}
    ms, out = run_flat(ctx, x1, steps, warmup, stream)
    per = ms / steps
    kname = ctx.kernel_name(n1)
    gbs, tfl = roofline(n1, per)
    res["auto"] = {"kernel": kname, "us_per_launch": per * 1e3, "knots_per_s": n1 / (per * 1e-3),
                   "hbm_frac": gbs / HBM_PEAK_GBS, "fp32_frac": tfl / FP32_PEAK_TFLOPS,
                   "traffic_bytes": MEASURED_TRAFFIC_BYTES.get((kname, n1))}
    CHECKS.add(res["auto"], "accuracy", out, x1)
    for name, kw in (("one_lane", {"lanes_per_knot": 1}), ("lane_group", {"lanes_per_knot": 16}),
                     ("fp32_hw_trig", {"precision": "fp32_hw_trig"})):
        c = rbdgrad.Context(device=0, **kw)
        if name == "fp32_hw_trig":
            c.autotune(stream)
        ms, out = run_flat(c, x1, steps // 4, warmup, stream)
        res[name] = {"kernel": c.kernel_name(n1), "us_per_launch": ms / (steps // 4) * 1e3,
                     "knots_per_s": n1 * (steps // 4) / (ms * 1e-3)}
        CHECKS.add(res[name], "accuracy", out, x1)
        c.close()
    # the reference's timed configuration (dynamicsGradientKernel, time_GPU.cu:279-316):
    # q, qd, u, qdd and the dense Minv in (77 floats per knot), 98 out
    qdd, minv = ctx.forward_dynamics(x1)
    vol = np.ascontiguousarray(np.concatenate([x1, qdd, minv.transpose(0, 2, 1).reshape(n1, 49)], axis=1),
                               dtype=np.float32)
    d_vol = rbdgrad.DeviceBuffer.from_array(vol)
    d_out = rbdgrad.DeviceBuffer(n1 * rbdgrad.DQDD_PER_KNOT * 4)
    ms = time_launches(lambda: ctx.dynamics_gradient_kernel(n1, d_out, d_vol, stream=stream), steps // 4, warmup,
                       stream)
    per = ms / (steps // 4)
    res["reference_mode"] = {"kernel": ctx.kernel_name(n1, reference_mode=True), "us_per_launch": per * 1e3,
                             "knots_per_s": n1 / (per * 1e-3), "bytes_per_knot": 700}
    CHECKS.add(res["reference_mode"], "accuracy", d_out.download((n1, rbdgrad.DQDD_PER_KNOT)), vol)
    d_vol.free()
    d_out.free()
    # back-to-back throughput over the reference harness's trajectory lengths
    res["over_n"] = {}
    for n in N_SET:
        ms, _ = run_flat(ctx, knots(n, seed=110 + n), steps // 4, warmup, stream)
        res["over_n"][str(n)] = {"kernel": ctx.kernel_name(n), "us_per_launch": ms / (steps // 4) * 1e3}
    log(f"[per-call] 128 knots: {res['auto']['us_per_launch']:.2f} us ({kname}); one lane "
        f"{res['one_lane']['us_per_launch']:.2f}, lane groups {res['lane_group']['us_per_launch']:.2f}, "
        f"reference mode {res['reference_mode']['us_per_launch']:.2f}")
    return res, vol


def mixed_trajectories(seed=7, count=64):
    rng = np.random.default_rng(seed)
    return rng.integers(8, 257, size=count)


def mixed_batch(ctx, stream, steps=500):
    """64 trajectories with lengths U{8..256}, each in its own device buffers:
    the per-call kernel (one launch per trajectory) against the batched form
    (one persistent-kernel launch over the device work queue), launched
    directly and replayed from a hipGraph."""
    lens = mixed_trajectories()
    n2 = int(lens.sum())
    x2 = knots(n2, seed=102)
    offs = np.concatenate([[0], np.cumsum(lens)])
    bufs = [(rbdgrad.DeviceBuffer.from_array(x2[a:b]), rbdgrad.DeviceBuffer((b - a) * rbdgrad.OUT_PER_KNOT * 4),
             int(b - a)) for a, b in zip(offs[:-1], offs[1:])]
    res = {"knots": n2, "trajectories": 64}

    def per_call_launches():
        for d_in, d_out, n in bufs:
            ctx.fdgrad(n, d_in, d_out, stream)
    ms = time_launches(per_call_launches, steps // 5, 5, stream) / (steps // 5)
    res["per_call_kernel"] = {"us_per_batch": ms * 1e3, "knots_per_s": n2 / (ms * 1e-3), "launches": 64}
    batch = rbdgrad.Batch(ctx, 64)
    tl = rbdgrad.TrajList(bufs)      # the descriptor array, built once (a C caller fills 64 structs per batch)
    for use_graph in (False, True):
        ms = time_launches(lambda: batch.run(tl, use_graph=use_graph, stream=stream), steps, 20, stream) / steps
        gbs, tfl = roofline(n2, ms)
        out = np.concatenate([d_out.download((n, rbdgrad.OUT_PER_KNOT)) for _, d_out, n in bufs])
        r = res["graph" if use_graph else "direct"] = {
            "us_per_batch": ms * 1e3, "knots_per_s": n2 / (ms * 1e-3), "lanes_per_knot": batch.last_lanes,
            "waves_per_group": next((w for n, l, w in ctx.mapping_table(batched=True)
                                     if n2 <= n and l == batch.last_lanes), 1),
            "hbm_frac": gbs / HBM_PEAK_GBS, "fp32_frac": tfl / FP32_PEAK_TFLOPS}
        CHECKS.add(r, "accuracy", out, x2, sample=4096)
    batch.close()
    for d_in, d_out, _ in bufs:
        d_in.free()
        d_out.free()
    log(f"[mixed] {n2} knots: per-call kernel {res['per_call_kernel']['us_per_batch']:.1f} us, batched "
        f"{res['direct']['us_per_batch']:.1f} us (lanes {res['direct']['lanes_per_knot']}, "
        f"{res['direct']['waves_per_group']} waves per workgroup), graph "
        f"{res['graph']['us_per_batch']:.1f} us")
    return res


SERVED_MODES = {
    # name: (zero_copy, lanes per knot)
    "auto_two_wave": (2, 32),
    "auto": (2, 16),
    "zero_copy": (1, 16),
    "device_buffers": (0, 16),
}


def served(ctx, modes, rates, duration_s=1.0):
    """Poisson request arrivals at each rate; each request one trajectory with
    length uniform over {16, 32, 64, 128, 256}; device-resident constants,
    admission by a fixed memory budget, continuous batching over 4 in-flight
    slots, a polling dispatcher.  Latency = submit -> result in the caller's
    host buffer."""
    pool = knots(1 << 16, seed=103)
    res = {}
    for mode in modes:
        zero_copy, lanes = SERVED_MODES[mode]
        srv = rbdgrad.Server(ctx, slots=4, lanes_per_knot=lanes, memory_budget_bytes=256 << 20, zero_copy=zero_copy,
                             poll_dispatch=1)
        res[mode] = {}
        for rate in rates:
            st = srv.run_poisson(rate, duration_s, [16, 32, 64, 128, 256], pool, seed=rate)
            st["batch_phases_us"] = batch_phases(srv.trace())
            st["hbm_frac"] = st["knots_per_s"] * BYTES_PER_KNOT / 1e9 / HBM_PEAK_GBS
            st["fp32_frac"] = st["knots_per_s"] * FLOPS_PER_KNOT / 1e12 / FP32_PEAK_TFLOPS
            res[mode][str(rate)] = st
            log(f"[served] {mode} {rate} req/s: p50 {st['p50_latency_us']:.1f} us, p99 {st['p99_latency_us']:.1f} us, "
                f"{st['knots_per_s']:.3e} knots/s")
        srv.close()
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


def harness_latency(seconds_per_leg):
    """The reference harness's per-call latency statistics (printStats:
    mean, std, min, max, the σ histogram and the 50..99th percentiles) over its
    trajectory lengths, from the native harness time_gpu: each call followed by
    a device synchronisation ("GPU Compute"), and with the host copies ("GPU
    Compute + I/O"), for the reference's timed configuration and the full
    path; plus the single-knot in-kernel latency at N = 1."""
    if not os.path.exists(HARNESS):
        log(f"[latency] {HARNESS} not built: skipped")
        return None
    r = subprocess.run([HARNESS, "--seconds", str(seconds_per_leg)], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"[latency] harness failed ({r.returncode}): {r.stderr[-500:]}")
        return {"failed": r.returncode}
    return json.loads(r.stdout)


def copy_bandwidth(stream, nbytes=LARGE * 420, reps=20):
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
    return {"bytes": nbytes, "us": ms * 1e3, "gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}


def urdf_model(stream, steps, warmup, name="arm7",
               urdf=os.path.join(ROOT, "tests", "golden", "synthetic_arm7.urdf")):
    """A robot read from a URDF: librbdgrad_arm7.so, compiled on the tables that
    urdf_model.py generates from the synthetic 7-joint arm (joint frames at
    general angles, so the transforms have no structural zeros).  The per-call
    and large-batch (SoA) workloads as for the iiwa."""
    if not os.path.exists(rbdgrad.lib_path(name)):
        log(f"[model] {rbdgrad.lib_path(name)} not built: skipped")
        return {}
    import urdf_model as um
    tab = um.tables(um.read_urdf(urdf))
    tables = (np.array(tab["links"]),
              np.concatenate([np.array(tab["rot"]).reshape(7, 9), np.array(tab["offset"])], axis=1))
    ctx = rbdgrad.Context(device=0, model=name)
    ctx.autotune(stream)
    res = {}
    x1 = knots(128, seed=101)
    ms, out = run_flat(ctx, x1, steps, warmup, stream)
    r = res["per_call_n128"] = {"kernel": ctx.kernel_name(128), "us_per_launch": ms / steps * 1e3,
                                "knots_per_s": 128 * steps / (ms * 1e-3)}
    CHECKS.add(r, "accuracy", out, x1, model=tables)
    x4 = knots(LARGE, seed=104)
    ms, out = run_flat(ctx, x4, 20, 3, stream, "soa")
    per = ms / 20
    gbs, tfl = roofline(LARGE, per, FLOPS_PER_KNOT_ARM7)
    r = res["large_2p20_fp32_soa"] = {
        "ms_per_launch": per, "knots_per_s": LARGE / (per * 1e-3), "hbm_frac": gbs / HBM_PEAK_GBS,
        "fp32_tflops": tfl, "fp32_frac": tfl / FP32_PEAK_TFLOPS, "flops_per_knot": FLOPS_PER_KNOT_ARM7}
    CHECKS.add(r, "accuracy", out, x4, sample=8192, model=tables)
    log(f"[model] {name}: per call {res['per_call_n128']['us_per_launch']:.2f} us, 2^20 SoA {per * 1e3:.1f} us")
    ctx.close()
    return res


# ---------------------------------------------------------------------------
# CPU leg: the oracle as the CPU baseline, and as the checker of the GPU outputs
# ---------------------------------------------------------------------------
def cpu_baseline(x_large, x1, vol1, seconds):
    """The oracle's FP32 restatement of the reference's ReusableThreads loop
    (time_CPU.cpp:47-54) timed on this host.  Headline: all usable cores (up to
    16, the GPU host's CPU share) on calls of 65,536 knots taken from the
    large-batch input, about `seconds` of calls.  Beside it, in the detail
    record: 4 threads (the reference's CPU_THREADS_GLOBAL), the 128-knot call
    at both thread counts, and the reference's own timed configuration (qdd
    and Minv given, QDD_MINV_PASSED_IN) on the 128-knot call."""
    from oracle import pyoracle
    cores = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())

    def leg(inp, mode, threads, secs):
        times, _ = pyoracle.cpu_bench(inp, threads, iters=10 ** 7, mode=mode, time_cap_s=secs)
        times = times[len(times) // 10:] if len(times) >= 10 else times   # drop the pool's warm-up calls
        mean_us = float(np.mean(times))
        return {"value": len(inp) / (mean_us * 1e-6), "unit": "knots/s", "cores": threads, "calls": int(len(times)),
                "mean_us": mean_us, "std_us": float(np.std(times)), "p50_us": float(np.percentile(times, 50)),
                "p99_us": float(np.percentile(times, 99))}

    sample = np.ascontiguousarray(x_large[:65536])
    head = leg(sample, 0, cores, seconds)
    head.update(kind="port", sample=(f"{head['calls']} calls of 65,536 knots of the 2^20-knot large-batch input "
                                     f"(10% stretched-arm poses), full path (q, qd, u in), FP32, ReusableThreads-style "
                                     f"pool, {cores} threads, ~{seconds:.0f} s"))
    detail = {"large_sample_4_threads": leg(sample, 0, 4, seconds / 2),
              "per_call_n128": [leg(x1, 0, t, seconds / 4) for t in sorted({4, cores})],
              "per_call_n128_reference_mode": [leg(vol1, 1, t, seconds / 4) for t in sorted({4, cores})]}
    log(f"[cpu] {cores} threads: {head['value']:.3e} knots/s; 4 threads: {detail['large_sample_4_threads']['value']:.3e}")
    return head, detail


# ---------------------------------------------------------------------------
def profile_one(args, stream):
    """One workload, for rocprofv3 runs (kernel trace / counters); prints its
    own event timing, so a profile and its timing come from the same run."""
    if args.only == "copy":
        # a plain device-to-device copy of known size, to calibrate FETCH_SIZE/WRITE_SIZE
        nbytes = LARGE * rbdgrad.OUT_PER_KNOT * 4
        a, b = rbdgrad.DeviceBuffer(nbytes), rbdgrad.DeviceBuffer(nbytes)
        a.fill_zero()
        for _ in range(args.steps):
            rbdgrad._hipcheck(rbdgrad.hip().hipMemcpy(b.handle, a.handle, nbytes, rbdgrad.D2D), "hipMemcpy D2H")
        rbdgrad.hip().hipDeviceSynchronize()
        log(f"copy: {args.steps} x {nbytes} bytes device to device")
        return
    if args.only == "c2":
        # the built-in mapping table (a profiler's overhead would skew an autotune run)
        ctx = rbdgrad.Context(device=0, precision=args.precision, model=args.model)
        log(json.dumps(mixed_batch(ctx, stream, steps=args.steps)))
        return
    if args.only == "c4-other":
        other_entries(stream, args.steps, min(args.warmup, 10), knots(LARGE, seed=104, near_singular_frac=0.1))
        return
    if args.only == "c3":
        ctx = rbdgrad.Context(device=0, precision=args.precision, model=args.model)
        mode = {16: "auto", 32: "auto_two_wave"}[args.server_lanes]
        log(json.dumps(served(ctx, [mode], [args.rate or 10000])))
        return
    lanes = 1 if args.only in ("c1-one-lane", "c4-aos", "c4-soa") else (16 if args.only == "c1-lane-group" else 0)
    ctx = rbdgrad.Context(device=0, precision=args.precision, lanes_per_knot=lanes, model=args.model)
    if args.only.startswith("c1"):
        n, x, layout = 128, knots(128, seed=101), "aos"
    else:
        n, layout = LARGE, args.only[3:]
        x = knots(n, seed=104, near_singular_frac=0.1)
    ms, _ = run_flat(ctx, x, args.steps, min(args.warmup, 10), stream, layout)
    kname = ctx.kernel_name(n).replace("_aos_", f"_{layout}_")    # the library names the AoS kernel
    log(f"{args.only} ({ctx.model}, {layout}, {args.precision}): {kname}, n={n}: "
        f"{ms / args.steps * 1e3:.3f} us per launch over {args.steps}")


def compact(line, limit=4000):
    """Keep the printed line under `limit` characters: secondary fields go
    first (they are all in the detail file)."""
    for k in ("other_entries_2p20", "c4_modes", "latency_us", "c3_served", "c1_per_call_n128", "c2_mixed_batch"):
        if len(json.dumps(line)) <= limit:
            break
        line.pop(k, None)
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed launches of the 2^20-knot headline")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--call-steps", type=int, default=2000, help="timed launches of the 128-knot per-call workload")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline time (headline leg)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--full", action="store_true", help="also every serving mode, at more rates")
    ap.add_argument("--detail", default=os.path.join(ROOT, "gpurun_out", "bench_detail.json"),
                    help="where the full record of every workload is written")
    ap.add_argument("--only", choices=["c1", "c1-one-lane", "c1-lane-group", "c4-aos", "c4-soa", "c4-other", "c2", "c3",
                                       "copy"],
                    help="run one workload only and print its timing (for profiler runs)")
    ap.add_argument("--precision", default="fp32", choices=list(rbdgrad.PRECISIONS), help="with --only")
    ap.add_argument("--rate", type=int, default=0, help="with --only c3: requests/s")
    ap.add_argument("--server-lanes", type=int, default=32, choices=[16, 32],
                    help="with --only c3: the server's mapping (16: lane groups, 32: two waves)")
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
    t_start = time.time()
    detail = {}
    # ---- headline: the large batch, every precision mode ----
    legs, x4 = large_batch(stream, args.steps, args.warmup)
    detail["large_batch_2p20"] = legs
    detail["other_entries_2p20"] = other_entries(stream, args.steps, args.warmup, x4)
    # ---- secondary workloads ----
    ctx = rbdgrad.Context(device=0, precision="fp32")
    thr, tune = ctx.autotune(stream)
    bthr, btune = ctx.autotune_batch(stream)
    detail["autotune"] = {"group_threshold_knots": thr, "two_wave_threshold_knots": ctx.two_wave_threshold,
                          "us_per_launch_by_shape": {str(k): v for k, v in tune.items()},
                          "mapping_table": ctx.mapping_table(),
                          "batch_group_threshold_knots": bthr,
                          "batch_us_by_shape": {str(k): v for k, v in btune.items()},
                          "batch_mapping_table": ctx.mapping_table(batched=True)}
    log(f"[autotune] per call {ctx.mapping_table()}; per batch {ctx.mapping_table(batched=True)}")
    detail["per_call_n128"], vol1 = per_call(ctx, stream, args.call_steps, 100)
    x1 = knots(128, seed=101)
    detail["mixed_64_traj"] = mixed_batch(ctx, stream)
    modes = list(SERVED_MODES) if args.full else ["auto_two_wave"]
    rates = (1000, 3000, 10000, 30000, 100000, 300000) if args.full else (1000, 10000, 100000)
    detail["served"] = served(ctx, modes, rates)
    detail["urdf_model_arm7"] = urdf_model(stream, args.call_steps // 4, 100)
    detail["hbm_copy"] = copy_bandwidth(stream)
    ctx.close()
    stream.destroy()
    detail["latency"] = harness_latency(0.25)

    # ---- CPU leg ----
    if args.no_cpu:
        cpu = None
        CHECKS.skip("--no-cpu")
    else:
        cpu, detail["cpu_baseline_detail"] = cpu_baseline(x4, x1, vol1, args.cpu_seconds)
        log("[cpu] checking the GPU outputs against the FP64 oracle")
        CHECKS.run(threads=min(16, os.cpu_count() or 1))
    detail["cpu_baseline"] = cpu

    # ---- the line ----
    key = choose_headline(legs) or "fp32_soa"
    h = legs[key]
    acc = h.get("accuracy", {})
    tflops, gbs = h["fp32_tflops"], h["hbm_gbs"]
    executed = None
    if key == "fp32_soa":      # the counters are of this kernel
        t = h["ms_per_launch"] * 1e-3
        ex_tflops = EXECUTED_FLOPS_PER_KNOT * LARGE / t * 1e-12
        executed = {"flops_per_knot": EXECUTED_FLOPS_PER_KNOT, "valu_instr_per_knot": EXECUTED_VALU_PER_KNOT,
                    "tflops": ex_tflops, "frac": ex_tflops / FP32_PEAK_TFLOPS}
    mixed = detail["mixed_64_traj"]
    pc = detail["per_call_n128"]
    srv = detail["served"]["auto_two_wave"]
    line = {
        "metric": METRIC,
        "value": h["knots_per_s"],
        "unit": "knots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_launch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if h["precision"].startswith("fp32") else h["precision"],
        "data": "synthetic: seeded uniform q~U(-pi,pi), qd~U(-2,2), u~U(-50,50); 10% stretched-arm poses; "
                "iiwa 14 model constants",
        "config": {"workload": "large batch: 2^20 knot points (10% with joints 2, 4, 6 within 1e-3 rad of 0), "
                               "one launch per step",
                   "knots_per_step": LARGE, "layout": h["layout"], "precision": h["precision"], "kernel": h["kernel"],
                   "mapping": "one lane per knot",
                   "settle_ms": SETTLE_MS,
                   "headline_choice": "FP32 holds the bar on every knot" if key == "fp32_soa" else
                                      f"FP32 breaks the bar somewhere: fastest passing mode {key}"},
        "roofline": {"bound": "fp32-valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP32_PEAK_TFLOPS, "traffic": h["traffic_bytes"],
                     "flops_per_knot": FLOPS_PER_KNOT, "knots_per_launch": LARGE,
                     "note": "flops_per_knot: the reference method's count, with the structural zeros of its dense "
                             "6x6 products that the kernel skips; executed work from counters below",
                     "executed": executed},
        "roofline_hbm": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": h["traffic_bytes"], "bytes_per_knot": BYTES_PER_KNOT},
        "accuracy": {"knots_checked": acc.get("knots_checked"), "bar": REL_TOL,
                     "max_rel_err": {b: acc[b]["max"] for b in ("qdd", "dqdd_dq", "dqdd_dqd")} if "qdd" in acc else acc},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "c4_modes": {k: {"us": round(r["ms_per_launch"] * 1e3, 1),
                         "max_err": max((r["accuracy"][b]["max"] for b in ("qdd", "dqdd_dq", "dqdd_dqd")),
                                        default=None) if "qdd" in r.get("accuracy", {}) else None}
                     for k, r in legs.items()},
        "other_entries_2p20": {k: {"us": round(r["us_per_launch"], 1), "hbm_frac": round(r["hbm_frac"], 3),
                                   "max_err": max((v["max"] for v in r.get("accuracy", {}).values()
                                                   if isinstance(v, dict)), default=None)}
                               for k, r in detail["other_entries_2p20"].items() if isinstance(r, dict)},
        "c1_per_call_n128": {"us": round(pc["auto"]["us_per_launch"], 2), "knots_per_s": pc["auto"]["knots_per_s"],
                             "kernel": pc["auto"]["kernel"]},
        "c2_mixed_batch": {"knots": mixed["knots"], "us_direct": round(mixed["direct"]["us_per_batch"], 2),
                           "us_graph": round(mixed["graph"]["us_per_batch"], 2),
                           "us_per_call_kernel": round(mixed["per_call_kernel"]["us_per_batch"], 2),
                           "lanes_per_knot": mixed["direct"]["lanes_per_knot"],
                           "waves_per_group": mixed["direct"]["waves_per_group"],
                           "knots_per_s": mixed["direct"]["knots_per_s"]},
        "c3_served": {"mode": "automatic (zero-copy batches up to 16,384 knots, copies above), two-wave batches, "
                              "4 slots, 256 MiB budget; p50/p99 us, knots/s",
                      "rates": {r: [round(v["p50_latency_us"], 1), round(v["p99_latency_us"], 1),
                                    float(f"{v['knots_per_s']:.3g}")] for r, v in srv.items()}},
        "detail": os.path.relpath(args.detail, ROOT),
    }
    lat = detail.get("latency")
    if lat and "full_compute" in lat:
        line["latency_us"] = {"full_path_compute_p50": {n: round(v["p50"], 2) for n, v in lat["full_compute"].items()},
                              "full_path_compute_io_p50": {n: round(v["p50"], 2)
                                                           for n, v in lat["full_compute_io"].items()}}
    detail["bench_seconds"] = time.time() - t_start
    os.makedirs(os.path.dirname(os.path.abspath(args.detail)), exist_ok=True)
    with open(args.detail, "w") as f:
        json.dump({"line": line, "workloads": detail}, f, indent=1, default=float)
    print(json.dumps(compact(line), default=float), flush=True)


if __name__ == "__main__":
    main()
