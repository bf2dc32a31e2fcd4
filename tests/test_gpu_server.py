"""GPU tests of the serving loop (rbd_server_*): continuous batching, admission
by a memory budget, overlap of batches, and results independent of batching."""
import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx(rbd):
    c = rbd.Context(precision="fp32")
    yield c
    c.close()


@pytest.mark.parametrize("zero_copy,poll", [(0, 0), (1, 0), (1, 1), (2, 1)])
@pytest.mark.parametrize("lanes", [1, 16, 32])
def test_result_is_the_same_alone_and_in_the_mix(rbd, ctx, oracle, lanes, zero_copy, poll):
    # zero_copy = 2 with a 300-knot limit: the probe alone is a zero-copy batch,
    # in the mix most batches are copied
    srv = rbd.Server(ctx, lanes_per_knot=lanes, slots=4, zero_copy=zero_copy, poll_dispatch=poll,
                     zero_copy_max_knots=300)
    rng = np.random.default_rng(21)
    probe = common.knots(128, seed=500)
    alone = srv.wait(srv.submit(probe))
    say("alone done")
    # now the same request among 300 others of mixed length, submitted back to back
    others = [common.knots(int(rng.choice([16, 32, 64, 128, 256])), seed=600 + i) for i in range(300)]
    ids = [srv.submit(x) for x in others[:150]]
    pid = srv.submit(probe)
    ids += [srv.submit(x) for x in others[150:]]
    mixed = srv.wait(pid)
    outs = [srv.wait(i) for i in ids]
    np.testing.assert_array_equal(mixed, alone)
    st = srv.stats()
    say("stats", st)
    assert st["requests_completed"] == 302
    # every one of the 300 against the oracle, and bitwise against the same
    # request served again on its own
    ref = oracle.fdgrad(np.concatenate(others).astype(np.float64), threads=8)
    err = common.max_block_errors(np.concatenate(outs), ref)
    assert max(err.values()) < common.REL_TOL, err
    for x, o in zip(others, outs):
        np.testing.assert_array_equal(srv.wait(srv.submit(x)), o)
    # the same requests through the plain per-call path with the same mapping:
    # identical (the two-wave batch kernel compiles the same arithmetic into
    # another kernel, so there FP32 rounding may differ)
    flat = rbd.Context(precision="fp32", lanes_per_knot=lanes)
    if lanes == 32:
        err = common.max_block_errors(alone, flat.run_aos(probe))
        assert max(err.values()) < 2e-5, err
    else:
        np.testing.assert_array_equal(flat.run_aos(probe), alone)
        for x, o in zip(others, outs):
            np.testing.assert_array_equal(flat.run_aos(x), o)
    flat.close()
    srv.close()


@pytest.mark.parametrize("zero_copy", [0, 1, 2])
def test_memory_budget_is_respected(rbd, ctx, zero_copy):
    """A budget that holds only a few requests: everything still completes, in
    several batches, and the pool never exceeds the budget."""
    per_knot = 21 * 4 + 105 * 4
    budget = 3 * 256 * per_knot + 4096
    srv = rbd.Server(ctx, memory_budget_bytes=budget, slots=2, zero_copy=zero_copy)
    xs = [common.knots(256, seed=700 + i) for i in range(20)]
    ids = [srv.submit(x) for x in xs]
    outs = [srv.wait(i) for i in ids]
    st = srv.stats()
    say("stats", st)
    assert st["requests_completed"] == 20
    assert st["max_budget_bytes_used"] <= budget
    assert st["batches"] >= 7
    flat = rbd.Context(precision="fp32", lanes_per_knot=16)
    for x, o in zip(xs, outs):
        np.testing.assert_array_equal(o, flat.run_aos(x))
    flat.close()
    srv.close()


def test_request_larger_than_budget_is_rejected(rbd, ctx):
    srv = rbd.Server(ctx, memory_budget_bytes=1 << 16)
    with pytest.raises(rbd.RbdError, match="capacity"):
        srv.submit(common.knots(256, seed=1))
    srv.close()


@pytest.mark.parametrize("poll", [0, 1])
@pytest.mark.parametrize("zero_copy", [0, 1, 2])
@pytest.mark.parametrize("slots", [1, 4])
def test_poisson_load(rbd, ctx, slots, zero_copy, poll):
    srv = rbd.Server(ctx, slots=slots, zero_copy=zero_copy, poll_dispatch=poll)
    pool = common.knots(8192, seed=800)
    st = srv.run_poisson(rate=20000, duration_s=0.5, lengths=[16, 32, 64, 128, 256], pool=pool, seed=1)
    say("poisson 20k/s", st)
    assert st["requests_completed"] > 5000
    assert st["max_inflight_batches"] <= slots
    assert 0 < st["p50_latency_us"] <= st["p99_latency_us"] <= st["max_latency_us"]
    assert st["knots_per_s"] > 0
    srv.close()


@pytest.mark.parametrize("zero_copy", [0, 1, 2])
def test_batches_overlap_on_the_device(rbd, ctx, zero_copy):
    """Requests arriving while batches are in flight are admitted into free
    slots: under a 100k req/s load several batches are on the device at once."""
    srv = rbd.Server(ctx, slots=4, zero_copy=zero_copy)
    pool = common.knots(8192, seed=801)
    st = srv.run_poisson(rate=100000, duration_s=0.2, lengths=[16, 32, 64, 128, 256], pool=pool, seed=2)
    say("poisson 100k/s", st)
    assert st["requests_completed"] > 15000
    assert 2 <= st["max_inflight_batches"] <= 4
    tr = srv.trace()
    assert len(tr["admit_us"]) == st["batches"]
    assert np.all(tr["admit_us"] >= tr["oldest_submit_us"])
    assert np.all(tr["enqueued_us"] >= tr["admit_us"])
    assert np.all(tr["done_us"] >= tr["enqueued_us"])
    assert np.all(tr["delivered_us"] >= tr["done_us"])
    assert int(tr["requests"].sum()) == st["requests_completed"]
    srv.close()


def test_server_config_errors(rbd, ctx):
    for bad in ({"zero_copy": 3}, {"zero_copy": -1}, {"zero_copy_max_knots": -5}):
        with pytest.raises(rbd.RbdError):
            rbd.Server(ctx, **bad)


def test_handoff_timeout_fails_the_requests(rbd):
    """Two-wave batches of a server whose context forces every handoff wait to
    give up (a host-set bound; every kernel still runs to its end): each
    request fails with RBD_ERR_DEVICE_TIMEOUT instead of returning the invalid
    results, and the server refuses new requests."""
    ctx = rbd.Context(precision="fp32")
    ctx.set_handoff_spins(-1)
    srv = rbd.Server(ctx, lanes_per_knot=32, slots=2, zero_copy=2, poll_dispatch=1)
    rid = srv.submit(common.knots(64, seed=700))
    with pytest.raises(rbd.RbdError, match="device wait timed out"):
        srv.wait(rid)
    with pytest.raises(rbd.RbdError, match="device wait timed out"):
        srv.submit(common.knots(64, seed=710))
    assert srv.stats()["requests_completed"] == 0
    srv.close()
    ctx.close()
