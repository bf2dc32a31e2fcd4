"""GPU tests of the batched uneven-trajectory path (rbd_batch_*): a persistent
kernel draining a device work queue, with and without hipGraph replay."""
import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_batch(rbd, lengths, seed):
    xs, trajs = [], []
    for i, n in enumerate(lengths):
        x = common.knots(int(n), seed=seed * 1000 + i) if n > 0 else np.zeros((0, 21), np.float32)
        d_in = rbd.DeviceBuffer.from_array(x) if n > 0 else rbd.DeviceBuffer(64)
        d_out = rbd.DeviceBuffer(max(1, int(n)) * 105 * 4)
        xs.append(x)
        trajs.append((d_in, d_out, int(n)))
    return xs, trajs


def download(trajs):
    return [d_out.download((n, 105)) for _, d_out, n in trajs]


def same_as_flat(batched, flat, lanes):
    """A batch's trajectory against the per-call kernel of the same mapping:
    bitwise for one lane and 16-lane groups (the same kernel code); within
    1e-5 per block for the 4-, 7- and 8-lane groups, whose batch kernel is
    another compilation of the same arithmetic (FMA contraction may differ)."""
    if lanes in (1, 16):
        np.testing.assert_array_equal(batched, flat)
    else:
        err = common.max_block_errors(batched, flat.astype(np.float64))
        assert max(err.values()) < 1e-5, err


@pytest.fixture(scope="module")
def ctx(rbd):
    c = rbd.Context(precision="fp32")
    yield c
    c.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_uneven_batch_matches_per_trajectory_calls(rbd, ctx, oracle, use_graph):
    lengths = np.random.default_rng(11).integers(8, 257, 64)        # the mixed-length workload
    xs, trajs = make_batch(rbd, lengths, seed=1)
    say(f"batch 64 trajectories, {lengths.sum()} knots, graph={use_graph}")
    b = rbd.Batch(ctx, 256)
    b.run(trajs, use_graph=use_graph)
    rbd.hip().hipDeviceSynchronize()
    total = int(lengths.sum())
    assert b.last_lanes == next(l for n, l, _ in ctx.mapping_table(batched=True) if total <= n)
    got = download(trajs)
    flat_ctx = rbd.Context(precision="fp32", lanes_per_knot=b.last_lanes)
    for x, g in zip(xs, got):
        same_as_flat(g, flat_ctx.run_aos(x), b.last_lanes)
    allx = np.concatenate(xs)
    err = common.max_block_errors(np.concatenate(got), oracle.fdgrad(allx.astype(np.float64), threads=8))
    assert max(err.values()) < common.REL_TOL, err
    flat_ctx.close()
    b.close()


def test_graph_replays_with_changing_batches(rbd, ctx):
    """Replay the captured graph several times with different trajectory sets,
    with eager work and synchronisation between replays; check every replay."""
    b = rbd.Batch(ctx, 128)
    flat = {lanes: rbd.Context(precision="fp32", lanes_per_knot=lanes) for lanes in rbd.AUTOTUNE_LANES}
    rng = np.random.default_rng(12)
    for rep in range(6):
        lengths = rng.integers(1, 200, rng.integers(1, 100))
        xs, trajs = make_batch(rbd, lengths, seed=100 + rep)
        b.run(trajs, use_graph=True)
        rbd.hip().hipDeviceSynchronize()
        lanes = b.last_lanes                                          # the mapping this batch ran
        flat[1].run_aos(common.knots(33, seed=rep))                   # eager work in between
        for x, g in zip(xs, download(trajs)):
            same_as_flat(g, flat[lanes].run_aos(x), lanes)
        say(f"replay {rep}: {len(lengths)} trajectories, {sum(lengths)} knots, lanes {lanes} ok")
    for c in flat.values():
        c.close()
    b.close()


def test_large_batch_uses_one_lane_mapping(rbd, ctx, oracle):
    lengths = [4000] * 8 + [777, 1, 0, 5000]
    xs, trajs = make_batch(rbd, lengths, seed=3)
    b = rbd.Batch(ctx, 16)
    b.run(trajs, use_graph=True)
    rbd.hip().hipDeviceSynchronize()
    assert b.last_lanes == 1
    got = download(trajs)
    one = rbd.Context(precision="fp32", lanes_per_knot=1)
    for x, g, n in zip(xs, got, lengths):
        if n:
            np.testing.assert_array_equal(g, one.run_aos(x))
    one.close()
    b.close()


def test_small_grid_still_drains_the_queue(rbd, ctx):
    """A persistent grid far smaller than the chunk count: every wave loops
    until the queue is empty, no knot is skipped."""
    lengths = [300, 5, 129, 64, 63]
    xs, trajs = make_batch(rbd, lengths, seed=4)
    for lanes in (16, 1, 7, 4):
        c = rbd.Context(precision="fp32", lanes_per_knot=lanes)
        b = rbd.Batch(c, 8)
        b.set_grid(lanes, 3 if lanes != 1 else 2)
        b.run(trajs)
        rbd.hip().hipDeviceSynchronize()
        assert b.last_lanes == lanes
        flat = rbd.Context(precision="fp32", lanes_per_knot=lanes)
        for x, g in zip(xs, download(trajs)):
            same_as_flat(g, flat.run_aos(x), lanes)
        flat.close()
        b.close()
        c.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_alternating_small_and_large_trajectory_counts(rbd, ctx, use_graph):
    """Batches of up to 64 trajectories find a chunk's trajectory inside the
    wave, larger ones through the plan kernel's prefix sum; the queue cursor is
    returned to zero by each launch, so back-to-back launches of both kinds on
    one batch object, without synchronisation between them, all cover every
    knot (zero-length trajectories included)."""
    b = rbd.Batch(ctx, 256)
    rng = np.random.default_rng(21)
    runs = []
    for rep, n_traj in enumerate([200, 10, 64, 65, 3, 256]):
        lengths = rng.integers(0, 90, n_traj)
        lengths[0] = max(lengths[0], 1)
        xs, trajs = make_batch(rbd, lengths, seed=300 + rep)
        b.run(trajs, use_graph=use_graph)
        runs.append((xs, trajs, b.last_lanes, lengths))
    rbd.hip().hipDeviceSynchronize()
    flat = {lanes: rbd.Context(precision="fp32", lanes_per_knot=lanes) for lanes in rbd.AUTOTUNE_LANES}
    for xs, trajs, lanes, lengths in runs:
        for x, g, n in zip(xs, download(trajs), lengths):
            if n:
                same_as_flat(g, flat[lanes].run_aos(x), lanes)
        say(f"{len(lengths)} trajectories, {int(sum(lengths))} knots, lanes {lanes} ok")
    for c in flat.values():
        c.close()
    b.close()


@pytest.mark.parametrize("lanes,waves", [(16, 100), (1, 70), (16, 2048), (1, 5)])
def test_last_round_counters_return_to_zero(rbd, lanes, waves):
    """Chunk totals above the grid and not a multiple of it, so every launch
    ends with a partial round pulled through the tail counters (64 of them,
    or one per wave for grids under 64); back-to-back launches on one batch
    object with no synchronisation between them each cover every knot, which
    holds only if each launch leaves its counters at zero."""
    c = rbd.Context(precision="fp32")
    c.batch_group_threshold = 1 << 40 if lanes == 16 else 0
    b = rbd.Batch(c, 128)
    b.set_grid(lanes, waves)
    rng = np.random.default_rng(40 + waves)
    chunk = 64 // lanes
    runs = []
    for rep in range(5):
        n_traj = int(rng.integers(1, 128))
        lengths = rng.integers(1, 3 * chunk * waves // n_traj + 2, n_traj)
        xs, trajs = make_batch(rbd, lengths, seed=400 + rep)
        total = int(sum((int(n) + chunk - 1) // chunk for n in lengths))
        b.run(trajs)
        assert b.last_lanes == lanes
        runs.append((xs, trajs, lengths, total))
    rbd.hip().hipDeviceSynchronize()
    flat = rbd.Context(precision="fp32", lanes_per_knot=lanes)
    for xs, trajs, lengths, total in runs:
        for x, g in zip(xs, download(trajs)):
            np.testing.assert_array_equal(g, flat.run_aos(x))
        say(f"lanes {lanes} grid {waves}: {len(lengths)} trajectories, {total} chunks "
            f"({total // waves} full rounds + {total % waves}) ok")
    flat.close()
    b.close()
    c.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_two_wave_batches(rbd, oracle, use_graph):
    """lanes_per_knot 32: 4-knot chunks on workgroups of two waves (Minv on one,
    RNEA and columns on the other), pulled in a fixed order.  Back-to-back
    launches without synchronisation: small batches (descriptors in the
    arguments), large ones (plan kernel), zero-length trajectories, and batches
    of more chunks than the grid's 4 workgroups per CU, so that workgroups loop
    over several chunks.  Every trajectory is bitwise the same as when run
    alone in a batch; against the flat two-wave kernel (the same arithmetic,
    compiled into another kernel, so FP32 rounding may differ) within 2e-5 per
    block, and within the 1e-4 bar of the oracle."""
    c = rbd.Context(precision="fp32", lanes_per_knot=32)
    b = rbd.Batch(c, 256)
    rng = np.random.default_rng(50)
    runs = []
    for rep, (n_traj, hi) in enumerate([(3, 40), (64, 257), (65, 90), (200, 120), (10, 2000), (1, 5)]):
        lengths = rng.integers(0, hi, n_traj)
        lengths[0] = max(lengths[0], 1)
        xs, trajs = make_batch(rbd, lengths, seed=500 + rep)
        b.run(trajs, use_graph=use_graph)
        assert b.last_lanes == 32
        runs.append((xs, trajs, lengths))
    rbd.hip().hipDeviceSynchronize()
    for rep, (xs, trajs, lengths) in enumerate(runs):
        got = download(trajs)
        for t, (x, g, n) in enumerate(zip(xs, got, lengths)):
            if not n:
                continue
            err = common.max_block_errors(g, c.run_aos(x))
            assert max(err.values()) < 2e-5, (rep, t, err)
            if t % 17 == 0:                               # alone in a batch: bitwise the same
                one = [(rbd.DeviceBuffer.from_array(x), rbd.DeviceBuffer(int(n) * 105 * 4), int(n))]
                b.run(one, use_graph=use_graph)
                rbd.hip().hipDeviceSynchronize()
                np.testing.assert_array_equal(download(one)[0], g)
        if rep == 1:
            err = common.max_block_errors(np.concatenate(got), oracle.fdgrad(np.concatenate(xs).astype(np.float64),
                                                                            threads=8))
            assert max(err.values()) < common.REL_TOL, err
        say(f"two-wave batch: {len(lengths)} trajectories, {int(sum(lengths))} knots "
            f"({int(sum((int(n) + 3) // 4 for n in lengths))} chunks) ok")
    b.close()
    c.close()


def test_batch_autotune(rbd):
    c = rbd.Context(precision="fp32")
    # the built-in table: 7 lanes in workgroups of 4 waves, 4 lanes, one lane
    assert [(l, w) for _, l, w in c.mapping_table(batched=True)] == [(7, 4), (4, 1), (1, 1)]
    thr, table = c.autotune_batch()
    sizes = sorted(table)
    assert len(sizes) == 11 and sizes[0] == 8 * 132 and sizes[-1] == 512 * 132 and 64 * 142 in sizes
    keys = {rbd.shape_key(*sh) for sh in rbd.AUTOTUNE_SHAPES}
    assert all(min(t.values()) > 0 and set(t) == keys for t in table.values())
    assert c.batch_group_threshold == thr
    assert thr == max([0] + [n for n in sizes if table[n]["16"] < table[n]["1"]])
    installed = c.mapping_table(batched=True)
    assert [n for n, _, _ in installed] == sizes
    for n, lanes, wpg in installed:
        assert table[n][rbd.shape_key(lanes, wpg)] == min(table[n].values())
    say(f"batch crossover {thr} knots; table {table}")
    c.close()


def test_batch_errors(rbd, ctx):
    b = rbd.Batch(ctx, 2)
    xs, trajs = make_batch(rbd, [5, 5, 5], seed=5)
    with pytest.raises(rbd.RbdError, match="capacity"):
        b.run(trajs)
    b.run([])                                     # empty batch is a no-op
    with pytest.raises(rbd.RbdError):
        b.set_grid(5, 10)
    b.close()


def run_batches(rbd, b, batches, use_graph, stream=None):
    """Launch every batch back to back (buffers allocated and filled first),
    synchronise once, return each batch's outputs and the mappings used."""
    lanes = []
    for xs, trajs in batches:
        b.run(trajs, use_graph=use_graph, stream=stream)
        lanes.append(b.last_lanes)
    rbd.hip().hipDeviceSynchronize()
    return [download(trajs) for _, trajs in batches], lanes


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("lanes", [4, 7, 8])
def test_group_mapping_batches(rbd, oracle, lanes, use_graph):
    """lanes_per_knot 4, 7, 8: chunks of floor(64 / L) knots of the batch's
    flattened knots, so a chunk spans trajectories.  Batches of up to 64
    trajectories (each lane finds its trajectory by a shuffle search) and of
    more (plan kernel's knot prefix sum), zero- and one-knot trajectories,
    back to back without synchronisation, more chunks than the grid.  Every
    trajectory is bitwise the same in another batch composition (reversed
    order), within 1e-5 per block of the flat kernel of the same mapping and
    within the 1e-4 bar of the oracle."""
    c = rbd.Context(precision="fp32", lanes_per_knot=lanes)
    b = rbd.Batch(c, 256)
    rng = np.random.default_rng(60 + lanes)
    specs = [(64, 257), (3, 40), (200, 120), (65, 90), (10, 2000), (1, 5), (64, 30)]
    batches = []
    for rep, (n_traj, hi) in enumerate(specs):
        lengths = rng.integers(0, hi, n_traj)
        lengths[0] = max(lengths[0], 1)
        if n_traj > 3:
            lengths[1:3] = [0, 1]
        batches.append(make_batch(rbd, lengths, seed=600 + 10 * lanes + rep))
    got, used = run_batches(rbd, b, batches, use_graph)
    assert used == [lanes] * len(specs)
    # the same trajectories in reverse order, another composition of every chunk
    rev = [(xs[::-1], [(d_in, rbd.DeviceBuffer(max(1, n) * 105 * 4), n) for d_in, _, n in trajs[::-1]])
           for xs, trajs in batches]
    got_rev, _ = run_batches(rbd, b, rev, use_graph)
    for (xs, trajs), g, gr in zip(batches, got, got_rev):
        for x, a, r in zip(xs, g, gr[::-1]):
            if len(x):
                np.testing.assert_array_equal(a, r)
                err = common.max_block_errors(a, c.run_aos(x))
                assert max(err.values()) < 1e-5, err
        allx = np.concatenate([x for x in xs if len(x)])
        err = common.max_block_errors(np.concatenate([a for x, a in zip(xs, g) if len(x)]),
                                      oracle.fdgrad(allx.astype(np.float64), threads=8))
        assert max(err.values()) < common.REL_TOL, err
        say(f"lanes {lanes}: {len(xs)} trajectories, {len(allx)} knots ok")
    b.close()
    c.close()


@pytest.mark.parametrize("lanes", [4, 7, 8])
def test_group_batches_in_larger_workgroups(rbd, lanes):
    """The batch kernels of the 4-, 7- and 8-lane groups in workgroups of 2
    and 4 waves give bitwise the results of workgroups of one wave: batches of
    up to 64 trajectories and more, back to back without synchronisation,
    direct and replayed, on the default grid and on grids of 3 and 5 waves
    (not multiples of the workgroup, with more chunks than waves, so the last
    partial round goes through the tail counters on every launch)."""
    rng = np.random.default_rng(80 + lanes)
    specs = [(64, 257), (3, 40), (200, 120), (65, 90), (1, 5), (64, 30)]
    batches = []
    for rep, (n_traj, hi) in enumerate(specs):
        lengths = rng.integers(0, hi, n_traj)
        lengths[0] = max(lengths[0], 1)
        if n_traj > 3:
            lengths[1:3] = [0, 1]
        batches.append(make_batch(rbd, lengths, seed=800 + 10 * lanes + rep))
    ref_ctx = rbd.Context(precision="fp32", lanes_per_knot=lanes, waves_per_group=1)
    rb = rbd.Batch(ref_ctx, 256)
    ref, used = run_batches(rbd, rb, batches, False)
    assert used == [lanes] * len(specs)
    rb.close()
    ref_ctx.close()
    for wpg in (2, 4):
        c = rbd.Context(precision="fp32", lanes_per_knot=lanes, waves_per_group=wpg)
        for grid in (None, 3, 5):
            for use_graph in (False, True):
                b = rbd.Batch(c, 256)
                if grid:
                    b.set_grid(lanes, grid)
                got, _ = run_batches(rbd, b, batches, use_graph)
                for g, r in zip(got, ref):
                    for a, e in zip(g, r):
                        np.testing.assert_array_equal(a, e)
                b.close()
            say(f"lanes {lanes}: workgroups of {wpg} waves, grid {grid or 'default'}: bitwise ok")
        c.close()


@pytest.mark.parametrize("lanes", [1, 7, 16])
def test_small_graph_replays_back_to_back(rbd, lanes):
    """Graph replays of small batches (up to 64 trajectories: the batch goes
    into the one kernel node's arguments before each replay) enqueued back to
    back on one non-default stream with no host synchronisation between them:
    each replay runs with its own arguments and grid.  Buffers are allocated
    and filled first, so nothing else synchronises.  Then 100,000 replays of one
    list, and three rounds of 20,000 calls alternating between two lists (a
    new capture each), leave the device's free memory where it was to within
    the runtime's pool steps: the replays within 8 MiB, all captures within 64
    MiB (the runtime's pools grow in steps of 10-22 MiB while earlier graphs
    wait to be destroyed, in any of the rounds), and at least one round
    within 8 MiB of the one before, so nothing is allocated per replay or per
    capture (rewriting a graph's kernel arguments instead kept about 34 KB per
    replay, 3.9 GB over 100,000; a leak of 420 bytes or more per capture would
    take more than 8 MiB in every round)."""
    c = rbd.Context(precision="fp32", lanes_per_knot=lanes)
    b = rbd.Batch(c, 64)
    s = rbd.Stream()
    rng = np.random.default_rng(70 + lanes)
    batches = [make_batch(rbd, rng.integers(1, 150, int(rng.integers(1, 65))), seed=700 + 10 * lanes + i)
               for i in range(10)]
    got, used = run_batches(rbd, b, batches, True, stream=s)
    assert used == [lanes] * 10
    for (xs, trajs), g in zip(batches, got):
        # the same batch launched directly (no graph): bitwise the same
        b.run(trajs, use_graph=False, stream=s)
        s.synchronize()
        for a, d in zip(g, download(trajs)):
            np.testing.assert_array_equal(a, d)
    tl = [rbd.TrajList(batches[0][1]), rbd.TrajList(batches[1][1])]
    for i in range(100):
        b.run(tl[i % 2], use_graph=True, stream=s)
    s.synchronize()
    free0 = rbd.mem_info()[0]
    for i in range(100000):
        b.run(tl[0], use_graph=True, stream=s)
        if i % 5000 == 0:
            s.synchronize()
    s.synchronize()
    free1 = rbd.mem_info()[0]
    free = []
    for _ in range(3):
        for i in range(20000):
            b.run(tl[i % 2], use_graph=True, stream=s)
        s.synchronize()
        free.append(rbd.mem_info()[0])
    say(f"lanes {lanes}: free device memory {free0} -> {free1} (100,000 replays) -> {free} "
        f"(20,000 captures each)")
    losses = [before - after for before, after in zip([free1] + free[:-1], free)]
    assert free1 >= free0 - (8 << 20)
    assert free[-1] >= free1 - (64 << 20), losses
    assert min(losses) <= (8 << 20), losses
    s.destroy()
    b.close()
    c.close()
