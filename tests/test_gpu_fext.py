"""GPU tests of the external-force entry points (rbd_fdgrad_traj_fext_f32,
rbd_forward_dynamics_fext_f32) against the FP64 oracle."""
import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "fp64_columns"])
@pytest.mark.parametrize("lanes", [1, 16])
@pytest.mark.parametrize("per_knot", [False, True])
def test_fdgrad_with_external_forces(rbd, oracle, precision, lanes, per_knot):
    n = 777
    x = common.knots(n, seed=500 + lanes, near_singular_frac=0.1)
    rng = np.random.default_rng(501)
    fext = rng.uniform(-20, 20, (n, 42) if per_knot else 42).astype(np.float32)
    print(f"fext n={n} precision={precision} lanes={lanes} per_knot={per_knot}", flush=True)
    for flags in (0, 1, 2):
        ctx = rbd.Context(flags=flags, precision=precision, lanes_per_knot=lanes)
        got = ctx.run_fext(x, fext)
        ref = oracle.fdgrad(x.astype(np.float64), flags, fext=fext.astype(np.float64))
        err = common.max_block_errors(got, ref)
        assert max(err.values()) < common.REL_TOL, (flags, err)
        ctx.close()


def test_forward_dynamics_with_external_forces(rbd, oracle):
    n = 300
    x = common.knots(n, seed=502)
    fext = np.random.default_rng(503).uniform(-20, 20, (n, 42)).astype(np.float32)
    ctx = rbd.Context(precision="fp32")
    qdd, minv = ctx.forward_dynamics(x, fext=fext)
    for i in range(0, n, 17):
        q_ref, m_ref = oracle.fd(*(x[i, s].astype(np.float64) for s in (slice(0, 7), slice(7, 14), slice(14, 21))),
                                 fext=fext[i].astype(np.float64))
        np.testing.assert_allclose(qdd[i], q_ref, rtol=0, atol=2e-6 * np.abs(q_ref).max())
        np.testing.assert_allclose(minv[i], m_ref, rtol=0, atol=2e-6 * np.abs(m_ref).max())
    # zero forces give the plain entry points' results
    zero = np.zeros(42, np.float32)
    q0, m0 = ctx.forward_dynamics(x, fext=zero)
    q1, m1 = ctx.forward_dynamics(x)
    np.testing.assert_array_equal(m0, m1)
    np.testing.assert_allclose(q0, q1, rtol=0, atol=1e-5 * np.abs(q1).max())
    ctx.close()


def test_external_force_argument_errors(rbd):
    ctx = rbd.Context(precision="fp32")
    x = common.knots(4, seed=504)
    d_in = rbd.DeviceBuffer.from_array(x)
    d_out = rbd.DeviceBuffer(4 * 105 * 4)
    d_f = rbd.DeviceBuffer(42 * 4 * 4)
    lib = rbd.lib()
    assert lib.rbd_fdgrad_traj_fext_f32(ctx.handle, 4, d_in.handle, None, 0, d_out.handle, None) != 0
    assert lib.rbd_fdgrad_traj_fext_f32(ctx.handle, 4, d_in.handle, d_f.handle, 41, d_out.handle, None) != 0
    assert lib.rbd_fdgrad_traj_fext_f32(ctx.handle, 4, d_in.handle, d_f.handle, 42, d_out.handle, None) == 0
    rbd.hip().hipDeviceSynchronize()
    ctx.close()
