"""GPU tests of the reference's staged variants: the fused and split
inverse-dynamics-gradient kernels, the v/a/f (start) stage and the finish
stage, against the FP64 oracle."""
import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu


def err_max(out, ref):
    return max(float(v.max()) for v in common.dqdd_block_errors(out, ref).values())


@pytest.mark.parametrize("flags", [0, 1, 2])
@pytest.mark.parametrize("lanes", [1, 16])
def test_fused_and_split_kernels(rbd, oracle, flags, lanes):
    n = 1000
    x = common.knots(n, seed=600 + flags, near_singular_frac=0.1).astype(np.float64)
    vol = np.concatenate([x, oracle.fdgrad(x, flags)[:, :7]], axis=1)
    vol32 = vol.astype(np.float32)
    ref = oracle.id_gradient(vol32.astype(np.float64), flags)
    print(f"staged n={n} flags={flags} lanes={lanes}", flush=True)
    for precision in ("fp32", "fp64_columns"):
        ctx = rbd.Context(flags=flags, precision=precision, lanes_per_knot=lanes)
        assert err_max(ctx.id_gradient(vol32), ref) < common.REL_TOL
        split = ctx.gradient_setup_split(vol32)
        ref_split = oracle.vaf(np.concatenate([vol32.astype(np.float64), np.zeros((n, 126))], axis=1), flags)
        np.testing.assert_allclose(split, ref_split, rtol=0, atol=2e-5 * np.abs(ref_split).max())
        assert err_max(ctx.id_gradient_split(split), ref) < common.REL_TOL
        ctx.close()


def test_finish_stage_gives_the_full_path(rbd, oracle):
    n = 500
    x = common.knots(n, seed=610)
    ctx = rbd.Context(precision="fp32")
    qdd, minv = ctx.forward_dynamics(x)
    vol = np.concatenate([x, qdd], axis=1)
    cdu = ctx.id_gradient(vol)
    dqdd = ctx.gradient_finish(minv.transpose(0, 2, 1).reshape(n, 49), cdu)
    ref = oracle.fdgrad(x.astype(np.float64))[:, 7:]
    assert err_max(dqdd, ref) < common.REL_TOL
    # the finish stage alone, against the oracle's product on the same FP32 operands
    m32 = minv.transpose(0, 2, 1).reshape(n, 49)
    ref_f = oracle.finish(m32.astype(np.float64), cdu.astype(np.float64))
    assert err_max(dqdd, ref_f) < 1e-6
    ctx.close()


def test_staged_variants_of_a_large_ragged_call(rbd, oracle):
    """One lane per knot over 3 * 2^16 + 37 knots (3,073 blocks of 64, the
    last one ragged), outputs prefilled with NaN, FP32: for the fused and
    split kernels and the finish stage every knot is written, each output
    equals bit for bit the same knots in calls of at most 40,000 knots, and a
    sample holds the bar against the oracle."""
    n = 3 * (1 << 16) + 37
    x = common.knots(n, seed=620, near_singular_frac=0.1)
    print(f"staged variants n={n}", flush=True)
    ctx = rbd.Context(precision="fp32", lanes_per_knot=1)
    qdd, minv = ctx.forward_dynamics(x)
    vol = np.ascontiguousarray(np.concatenate([x, qdd], axis=1), dtype=np.float32)
    m = np.ascontiguousarray(minv.transpose(0, 2, 1).reshape(n, 49))     # column-major per knot
    split = ctx.gradient_setup_split(vol)

    def call(fn, *inputs):
        k = len(inputs[0])
        d_in = [rbd.DeviceBuffer.from_array(a) for a in inputs]
        d_out = rbd.DeviceBuffer.from_array(np.full((k, rbd.DQDD_PER_KNOT), np.nan, np.float32))
        fn(k, d_out, *d_in)
        out = d_out.download((k, rbd.DQDD_PER_KNOT))
        for b in d_in + [d_out]:
            b.free()
        return out

    chunks = [slice(a, min(a + 40000, n)) for a in range(0, n, 40000)]
    cdu = call(ctx.dynamics_gradient_kernel_fused, vol)
    cdu_split = call(ctx.dynamics_gradient_kernel_split, split)
    dqdd = call(lambda k, o, c, mm: ctx.gradient_finish_dev(k, o, c, mm), cdu, m)
    for out in (cdu, cdu_split, dqdd):
        assert np.isfinite(out).all()
    np.testing.assert_array_equal(cdu, np.concatenate([call(ctx.dynamics_gradient_kernel_fused, vol[c])
                                                       for c in chunks]))
    np.testing.assert_array_equal(cdu_split, np.concatenate([call(ctx.dynamics_gradient_kernel_split, split[c])
                                                             for c in chunks]))
    np.testing.assert_array_equal(dqdd, np.concatenate([call(lambda k, o, c, mm: ctx.gradient_finish_dev(k, o, c, mm),
                                                             cdu[c], m[c]) for c in chunks]))
    sample = np.concatenate([np.random.default_rng(8).choice(n - 64, 500, replace=False), np.arange(n - 64, n)])
    ref_cdu = oracle.id_gradient(vol[sample].astype(np.float64))
    assert err_max(cdu[sample], ref_cdu) < common.REL_TOL
    assert err_max(cdu_split[sample], ref_cdu) < common.REL_TOL
    assert err_max(dqdd[sample], oracle.fdgrad(x[sample].astype(np.float64))[:, 7:]) < common.REL_TOL
    ctx.close()
