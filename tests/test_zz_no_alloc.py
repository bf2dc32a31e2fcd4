"""GPU test of rbd_fdgrad_batched_noalloc_f32, the batched call that allocates
nothing (its queue counters live in a file-scope device array of the
library): through the C ABI, on two calls in a row -- the second with fewer
trajectories and knots -- its outputs are bitwise those of
rbd_fdgrad_batched_f32 on the same trajectories, for every mapping; more
than 64 trajectories are refused."""
import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu


def inputs(rbd, lengths, seed):
    xs, d_ins = [], []
    for i, n in enumerate(lengths):
        x = common.knots(int(n), seed=seed * 1000 + i) if n > 0 else np.zeros((0, 21), np.float32)
        xs.append(x)
        d_ins.append(rbd.DeviceBuffer.from_array(x) if n > 0 else rbd.DeviceBuffer(64))
    return xs, d_ins


def outputs(rbd, lengths):
    return [rbd.DeviceBuffer(max(1, int(n)) * 105 * 4) for n in lengths]


@pytest.mark.parametrize("lanes", [0, 1, 4, 7, 8, 16, 32])
def test_noalloc_matches_the_batched_call(rbd, oracle, lanes):
    ctx = rbd.Context(precision="fp32", lanes_per_knot=lanes)
    batch = rbd.Batch(ctx, 64)
    rng = np.random.default_rng(80 + lanes)
    for call, (n_traj, hi) in enumerate([(64, 257), (9, 40)]):       # the second call smaller
        lengths = rng.integers(0, hi, n_traj)
        lengths[0] = max(int(lengths[0]), 1)
        xs, d_ins = inputs(rbd, lengths, seed=800 + 10 * lanes + call)
        out_a, out_b = outputs(rbd, lengths), outputs(rbd, lengths)
        ctx.batched_noalloc([(i, o, int(n)) for i, o, n in zip(d_ins, out_a, lengths)])
        batch.run([(i, o, int(n)) for i, o, n in zip(d_ins, out_b, lengths)])
        rbd.hip().hipDeviceSynchronize()
        got_a = [o.download((int(n), 105)) for o, n in zip(out_a, lengths) if n]
        got_b = [o.download((int(n), 105)) for o, n in zip(out_b, lengths) if n]
        for a, b in zip(got_a, got_b):
            np.testing.assert_array_equal(a, b)
        allx = np.concatenate([x for x in xs if len(x)])
        err = common.max_block_errors(np.concatenate(got_a), oracle.fdgrad(allx.astype(np.float64), threads=8))
        assert max(err.values()) < common.REL_TOL, err
        print(f"lanes {lanes} call {call}: {n_traj} trajectories, {len(allx)} knots, mapping {batch.last_lanes}",
              flush=True)
    too_many = [(rbd.DeviceBuffer(64), rbd.DeviceBuffer(64), 0)] * 65
    with pytest.raises(rbd.RbdError, match="at most 64"):
        ctx.batched_noalloc(too_many)
    batch.close()
    ctx.close()
