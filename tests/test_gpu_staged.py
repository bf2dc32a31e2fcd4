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
