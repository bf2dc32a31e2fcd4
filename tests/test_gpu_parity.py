"""GPU parity of librbdgrad.so (HIP, gfx950) against the FP64 oracle and the
committed golden fixtures, through the C ABI.  Run on the GPU host:
    python -u -m pytest tests -m gpu -x -v --timeout 120 --timeout-method thread

Bar (the target for the FP32 product): max relative error per output block
(max|out - ref| / max|ref|) <= 1e-4 against the FP64 oracle on the same FP32
inputs.  Integer/identity properties (layout equivalence, batch independence,
determinism) are checked bit for bit.
"""
import os

import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu

GOLDEN = np.load(os.path.join(common.GOLDEN, "fdgrad_golden.npz"))
SETS = sorted({k.split("__")[0] for k in GOLDEN.files if not k.startswith("refmode")})
PRECISIONS = ["fp32", "fp64_pivot", "fp64_columns", "fp64", "fp32_hw_trig"]
# fp64 keeps only the FP32 rounding of inputs/outputs
# lanes per knot: one lane, groups of 4, 7, 8 and 16 lanes, two waves of 16
MAPPINGS = (1, 4, 7, 8, 16, 32)
BOUND = {"fp32": common.REL_TOL, "fp64_pivot": common.REL_TOL, "fp64_columns": 1e-5, "fp64": 1e-6,
         "fp32_hw_trig": common.REL_TOL}


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def contexts(rbd):
    """(flags, precision) -> context with the mapping chosen per call;
    (flags, precision, lanes) -> context with the mapping forced."""
    ctxs = {}
    for flags in (0, 1, 2, 3):
        for p in PRECISIONS:
            ctxs[(flags, p)] = rbd.Context(flags=flags, precision=p)
            for lanes in MAPPINGS:
                ctxs[(flags, p, lanes)] = rbd.Context(flags=flags, precision=p, lanes_per_knot=lanes)
    yield ctxs
    for c in ctxs.values():
        c.close()


@pytest.mark.parametrize("lanes", MAPPINGS)
@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("name", SETS)
def test_golden_aos(contexts, name, prec, lanes):
    x, flags, ref = GOLDEN[name + "__in"], int(GOLDEN[name + "__flags"]), GOLDEN[name + "__out"]
    say(f"golden {name} n={len(x)} flags={flags} prec={prec} lanes={lanes}")
    out = contexts[(flags, prec, lanes)].run_aos(x)
    err = common.max_block_errors(out, ref)
    assert max(err.values()) < BOUND[prec], err


@pytest.mark.parametrize("lanes", [1, 16])
@pytest.mark.parametrize("prec", ["fp32", "fp64_columns"])
def test_reference_mode_golden(contexts, prec, lanes):
    vol, ref = GOLDEN["refmode_n128__in"], GOLDEN["refmode_n128__out"]
    say(f"refmode n={len(vol)} prec={prec} lanes={lanes}")
    out = contexts[(0, prec, lanes)].run_refmode(vol)
    err = common.dqdd_block_errors(out, ref)
    assert max(float(v.max()) for v in err.values()) < BOUND[prec]


@pytest.mark.parametrize("lanes", MAPPINGS)
@pytest.mark.parametrize("prec", PRECISIONS)
def test_soa_equals_aos_bitwise(contexts, prec, lanes):
    x = common.knots(1000, seed=300)
    say(f"soa vs aos n=1000 prec={prec} lanes={lanes}")
    ctx = contexts[(0, prec, lanes)]
    np.testing.assert_array_equal(ctx.run_soa(x, pad=24), ctx.run_aos(x))


@pytest.mark.parametrize("prec", PRECISIONS)
def test_lane_group_matches_one_lane(contexts, prec):
    """Both knot-to-lane mappings compute the same numbers (to FP32 rounding of
    a different instruction schedule) and every output value is written."""
    x = common.knots(300, seed=306)
    a = contexts[(0, prec, 1)].run_aos(x)
    bound = 2e-5 if prec == "fp32_hw_trig" else 1e-5     # its larger errors differ more between schedules
    for lanes in MAPPINGS[1:]:
        b = contexts[(0, prec, lanes)].run_aos(x)
        assert np.all(np.isfinite(b))
        err = common.max_block_errors(b, a.astype(np.float64))
        assert max(err.values()) < bound, (lanes, err)


def test_autotune_sets_a_crossover(rbd):
    """The autotuner measures every mapping and installs what it measured;
    which mapping wins is a timing, reported by bench.py, not asserted here."""
    ctx = rbd.Context(precision="fp32")
    # the built-in table: two waves, 7 lanes in workgroups of 4 waves, 4 lanes, one lane
    assert [(l, w) for _, l, w in ctx.mapping_table()] == [(32, 1), (7, 4), (4, 1), (1, 1)]
    thr, table = ctx.autotune()
    say("autotune crossover", thr, table)
    assert ctx.group_threshold == thr
    keys = {rbd.shape_key(*sh) for sh in rbd.AUTOTUNE_SHAPES}
    assert len(keys) == 12
    assert all(min(t.values()) > 0 and set(t) == keys for t in table.values())
    assert thr in (0,) + tuple(table) and ctx.two_wave_threshold in (0,) + tuple(table)
    installed = ctx.mapping_table()
    assert [n for n, _, _ in installed] == sorted(table)
    for n, lanes, wpg in installed:                                  # the fastest shape at each size
        assert table[n][rbd.shape_key(lanes, wpg)] == min(table[n].values())
        assert ctx.kernel_name(n) == ctx.kernel_name(n)              # dispatch resolves
    ctx.group_threshold = 1000                                       # by hand: the thresholds decide
    assert ctx.mapping_table() == []
    ctx.close()


def test_handoff_timeout_is_reported(rbd):
    """A two-wave launch whose handoff waits all give up (forced from the host
    with a negative bound; every kernel still runs to its end) is reported:
    rbd_context_status says so after the synchronisation, and further
    launches on the context fail until the status is cleared."""
    ctx = rbd.Context(precision="fp32", lanes_per_knot=32)
    x = common.knots(256, seed=720)
    good = ctx.run_aos(x)
    assert ctx.status() == 0
    ctx.set_handoff_spins(-1)
    ctx.run_aos(x)                                   # launches; the download synchronises
    assert ctx.status() == rbd.ERR_DEVICE_TIMEOUT
    with pytest.raises(rbd.RbdError, match="device wait timed out"):
        ctx.run_aos(x)
    # the batched two-wave kernel reports through the same word
    ctx.clear_status()
    b = rbd.Batch(ctx, 4)
    d_in, d_out = rbd.DeviceBuffer.from_array(x), rbd.DeviceBuffer(len(x) * 105 * 4)
    b.run([(d_in, d_out, len(x))])
    d_out.download((len(x), 105))
    assert ctx.status() == rbd.ERR_DEVICE_TIMEOUT
    with pytest.raises(rbd.RbdError, match="device wait timed out"):
        b.run([(d_in, d_out, len(x))])
    b.close()
    # cleared and bounded again: results as before
    ctx.clear_status()
    ctx.set_handoff_spins(1 << 24)
    np.testing.assert_array_equal(ctx.run_aos(x), good)
    assert ctx.status() == 0
    ctx.close()


@pytest.mark.parametrize("n", [1, 128, 777, 16384, 16385, 40000])
def test_host_pointer_call_equals_device_call(contexts, n):
    """Host buffers: up to 16,384 knots through pinned memory the kernel reads
    and writes directly, above through device buffers; the same kernel as the
    device-pointer call, so bitwise its results, also on repeated calls."""
    x = common.knots(n, seed=301 + n)
    say(f"host-pointer call n={n}")
    ctx = contexts[(0, "fp32")]
    ref = ctx.run_aos(x)
    for _ in range(2):
        np.testing.assert_array_equal(ctx.fdgrad_host(x), ref)


@pytest.mark.parametrize("lanes", [0, 1, 7, 4])
@pytest.mark.parametrize("n", [1, 2, 8, 9, 10, 17, 63, 64, 65, 127, 129, 1000])
def test_ragged_sizes(contexts, oracle, n, lanes):
    """Sizes that end inside a wave's last knot group (64 knots per wave with
    one lane per knot, 9 with 7 lanes, 16 with 4; 0: the mapping the context
    picks); with one lane, SoA equals AoS bit for bit."""
    x = common.knots(n, seed=400 + n)
    say(f"ragged n={n} lanes={lanes}")
    ctx = contexts[(0, "fp32", lanes)] if lanes else contexts[(0, "fp32")]
    out = ctx.run_aos(x)
    err = common.max_block_errors(out, oracle.fdgrad(x.astype(np.float64)))
    assert max(err.values()) < common.REL_TOL, err
    if lanes == 1:
        np.testing.assert_array_equal(ctx.run_soa(x, pad=3), out)


@pytest.mark.parametrize("lanes", [4, 7, 8])
def test_group_workgroup_sizes_are_bitwise_equal(rbd, oracle, lanes):
    """The 4-, 7- and 8-lane groups in workgroups of 1, 2 and 4 waves (the
    waves of a workgroup share nothing): bitwise the same results, AoS and
    SoA, at sizes that end inside a wave, a workgroup and the grid."""
    ctxs = {w: rbd.Context(precision="fp32", lanes_per_knot=lanes, waves_per_group=w) for w in (1, 2, 4)}
    for n in (1, 2, 8, 9, 10, 17, 63, 64, 65, 575, 1000, 9014, 9217, 20000):
        x = common.knots(n, seed=700 + n)
        ref_aos = ctxs[1].run_aos(x)
        ref_soa = ctxs[1].run_soa(x, pad=3)
        for w in (2, 4):
            np.testing.assert_array_equal(ctxs[w].run_aos(x), ref_aos)
            np.testing.assert_array_equal(ctxs[w].run_soa(x, pad=3), ref_soa)
        if n == 1000:
            err = common.max_block_errors(ref_aos, oracle.fdgrad(x.astype(np.float64)))
            assert max(err.values()) < common.REL_TOL, err
        say(f"lanes {lanes}, n={n}: workgroups of 1, 2, 4 waves agree")
    for c in ctxs.values():
        c.close()


def test_zero_knots_is_a_no_op(rbd, contexts):
    ctx = contexts[(0, "fp32")]
    d = rbd.DeviceBuffer(64)
    ctx.fdgrad(0, d, d)
    ctx.fdgrad_soa(0, d, 0, d, 0)
    ctx.dynamics_gradient_kernel(0, d, d)


def test_invalid_arguments(rbd, contexts):
    ctx = contexts[(0, "fp32")]
    d = rbd.DeviceBuffer(21 * 4 * 8)
    with pytest.raises(rbd.RbdError, match="negative"):
        ctx.fdgrad(-1, d, d)
    with pytest.raises(rbd.RbdError, match="leading dimension"):
        ctx.fdgrad_soa(8, d, 4, d, 8)


@pytest.mark.parametrize("lanes", MAPPINGS)
def test_batch_composition_independence(contexts, lanes):
    """For a fixed knot-to-lane mapping each knot's result is bitwise the same
    alone and inside any batch (the 16-lane mapping switches between its
    latency and throughput builds at 1024 waves: 4096 knots straddles it)."""
    x = common.knots(8192, seed=302)
    say(f"batch independence n=8192 lanes={lanes}")
    ctx = contexts[(0, "fp32", lanes)]
    full = ctx.run_aos(x)
    idx = np.random.default_rng(0).choice(8192, 37, replace=False)
    for i in idx:
        np.testing.assert_array_equal(ctx.run_aos(x[i:i + 1])[0], full[i])
    perm = np.random.default_rng(1).permutation(8192)
    np.testing.assert_array_equal(ctx.run_aos(x[perm]), full[perm])


def test_non_finite_input_is_contained(contexts):
    x = common.knots(200, seed=303)
    ctx = contexts[(0, "fp32")]
    clean = ctx.run_aos(x)
    bad = x.copy()
    bad[17, 3] = np.nan
    bad[101, 9] = np.inf
    out = ctx.run_aos(bad)
    ok = np.ones(200, bool)
    ok[[17, 101]] = False
    np.testing.assert_array_equal(out[ok], clean[ok])


@pytest.fixture(scope="module")
def large_batch(oracle):
    """The large-batch workload: 2^20 knots, 10% at stretched-arm poses (joints
    2, 4, 6 within 1e-3 rad of 0), and the FP64 oracle on every knot."""
    x = common.knots(1 << 20, seed=304, near_singular_frac=0.1)
    return x, oracle.fdgrad(x.astype(np.float64), threads=16)


@pytest.mark.parametrize("prec,bar", [("fp32", common.REL_TOL), ("fp64_pivot", common.REL_TOL),
                                      ("fp64_columns", 2e-5), ("fp32_hw_trig", common.REL_TOL)])
def test_large_batch_every_knot(rbd, large_batch, prec, bar):
    """FP32-only against FP64-pivot Minv (the pivots D_k = S^T IA_k S and 1/D_k
    in FP64, minv.h:430) and FP64 derivative columns, at 2^20 knots: the
    largest error over every knot and output block is under the bar (1e-4 for
    the FP32-based modes; 2e-5 with FP64 columns, whose dqdd/dq keeps the
    errors of the FP32 Minv and RNEA: 1.2e-5 at most on the bench's 2^20
    knots).  Also: deterministic, and the first d/dq column exactly zero."""
    x, ref = large_batch
    say(f"large batch n={len(x)} prec={prec}")
    ctx = rbd.Context(precision=prec, lanes_per_knot=1)
    out = ctx.run_soa(x)
    err = common.block_errors(out, ref)
    worst = {k: float(v.max()) for k, v in err.items()}
    say(prec, "max", worst, "p99.9", {k: float(np.percentile(v, 99.9)) for k, v in err.items()})
    assert max(worst.values()) < bar, worst
    assert np.abs(out[:, 7:14]).max() == 0.0
    if prec == "fp32":
        np.testing.assert_array_equal(ctx.run_soa(x), out)          # deterministic
        np.testing.assert_array_equal(ctx.run_aos(x), out)          # AoS and SoA bitwise equal
    ctx.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64_columns"])
def test_one_lane_calls_beyond_the_resident_waves(rbd, oracle, prec):
    """A one-lane call of more 64-knot blocks than the device keeps resident
    runs one round of resident waves, each looping over blocks w, w + G, ...
    Here 6,145 blocks (a ragged size) on 2,048 or 1,536 waves with FP32
    columns, 1,024 with FP64 columns: every knot is written (the output starts
    as NaN), AoS and SoA agree bit for bit, knots computed alone are bitwise
    the same, sampled knots (the last block among them) hold the bar, and the
    reference-compatible layout gives each tiled golden knot its 128-knot
    result."""
    n = 3 * (1 << 17) + 37
    x = common.knots(n, seed=306, near_singular_frac=0.1)
    say(f"one lane beyond resident waves n={n} prec={prec}")
    ctx = rbd.Context(precision=prec, lanes_per_knot=1)
    d_in = rbd.DeviceBuffer.from_array(x)
    d_out = rbd.DeviceBuffer.from_array(np.full((n, rbd.OUT_PER_KNOT), np.nan, np.float32))
    ctx.fdgrad(n, d_in, d_out)
    out = d_out.download((n, rbd.OUT_PER_KNOT))
    d_in.free()
    d_out.free()
    assert np.isfinite(out).all()
    np.testing.assert_array_equal(ctx.run_soa(x), out)
    rng = np.random.default_rng(7)
    sample = np.concatenate([rng.choice(n - 64, 2000, replace=False), np.arange(n - 64, n)])
    for i in sample[::50]:
        np.testing.assert_array_equal(ctx.run_aos(x[i:i + 1])[0], out[i])
    err = common.max_block_errors(out[sample], oracle.fdgrad(x[sample].astype(np.float64)))
    assert max(err.values()) < BOUND[prec], err
    vol = GOLDEN["refmode_n128__in"]
    one = ctx.run_refmode(vol)
    tiled = ctx.run_refmode(np.tile(vol, (n // len(vol) + 1, 1))[:n])
    np.testing.assert_array_equal(tiled, np.tile(one, (n // len(vol) + 1, 1))[:n])
    ctx.close()


def test_fp32_error_statistics_near_singular(contexts, oracle):
    """Every knot at a stretched-arm pose: each precision mode holds its bar at
    the largest error over the knots."""
    x = common.knots(16384, seed=305, near_singular_frac=1.0)
    ref = oracle.fdgrad(x.astype(np.float64), threads=16)
    res = {}
    for p in ("fp32", "fp64_pivot", "fp64_columns", "fp32_hw_trig"):
        e = common.block_errors(contexts[(0, p)].run_aos(x), ref)
        res[p] = {k: (float(v.max()), float(np.percentile(v, 99.9))) for k, v in e.items()}
        say(p, res[p])
    assert max(v[0] for v in res["fp64_columns"].values()) < 1e-5
    for p in ("fp32", "fp64_pivot", "fp32_hw_trig"):
        assert max(v[0] for v in res[p].values()) < common.REL_TOL, (p, res[p])


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_hardware_trig_mode_holds_the_bar(contexts, oracle, flags):
    """RBD_PREC_FP32_HW_TRIG (hardware sin/cos) against the FP64 oracle on
    random poses, stretched-arm poses and angles far outside [-pi, pi]."""
    x = np.concatenate([common.knots(8192, seed=310 + flags), common.knots(8192, seed=320 + flags, near_singular_frac=1.0)])
    big = common.knots(2048, seed=330 + flags)
    big[:, :7] += np.float32(2 * np.pi) * np.random.default_rng(331).integers(-50, 50, (2048, 7)).astype(np.float32)
    x = np.ascontiguousarray(np.concatenate([x, big]), dtype=np.float32)
    ref = oracle.fdgrad(x.astype(np.float64), flags, threads=8)
    for lanes in MAPPINGS:
        e = common.block_errors(contexts[(flags, "fp32_hw_trig", lanes)].run_aos(x), ref)
        err = {k: float(v.max()) for k, v in e.items()}
        say(f"hw trig flags={flags} lanes={lanes}", err)
        assert max(err.values()) < common.REL_TOL, err


def test_secondary_entry_points_beyond_the_resident_waves(rbd):
    """The one-lane kernels of the other entry points (+ Minv, the WAFR
    layout, external forces, forward dynamics) stage their outputs in LDS and,
    for calls of more 64-knot blocks than the device keeps resident, run one
    round of resident waves that loop over the blocks.  A ragged call of
    196,645 knots (3,073 blocks, more than any of them keeps resident) equals,
    bit for bit, the same knots in calls of at most 40,000 knots (fewer blocks
    than resident waves, one block per wave)."""
    n = 3 * (1 << 16) + 37
    x = common.knots(n, seed=307, near_singular_frac=0.1)
    fext = np.random.default_rng(308).uniform(-20, 20, (n, 42)).astype(np.float32)
    chunks = [slice(a, min(a + 40000, n)) for a in range(0, n, 40000)]
    say(f"secondary entry points n={n}")
    ctx = rbd.Context(precision="fp32", lanes_per_knot=1)
    whole = ctx.run_with_minv(x)
    parts = [ctx.run_with_minv(x[c]) for c in chunks]
    for j in range(2):
        np.testing.assert_array_equal(whole[j], np.concatenate([p[j] for p in parts]))
    whole = ctx.run_wafr(x)
    parts = [ctx.run_wafr(x[c]) for c in chunks]
    for j in range(2):
        np.testing.assert_array_equal(whole[j], np.concatenate([p[j] for p in parts]))
    whole = ctx.run_fext(x, fext)
    np.testing.assert_array_equal(whole, np.concatenate([ctx.run_fext(x[c], fext[c]) for c in chunks]))
    whole = ctx.forward_dynamics(x)
    parts = [ctx.forward_dynamics(x[c]) for c in chunks]
    for j in range(2):
        np.testing.assert_array_equal(whole[j], np.concatenate([p[j] for p in parts]))
    assert np.isfinite(whole[1]).all()
    ctx.close()
