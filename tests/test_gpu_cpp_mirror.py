"""Runs the C++ parity program tests/cpp/test_mirror (built by
__graft_entry__.build()), which calls the library through the reference-shaped
C++ interface fast-rbd-gradients_amd/rbdgrad.hpp the way the reference's
time_GPU.cu harness calls its kernel."""
import os
import subprocess

import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu

BIN = os.path.join(common.ROOT, "tests", "cpp", "test_mirror")


def close_f32(a, b, rel=1e-5):
    """Equal to FP32 rounding: separate kernels may contract different products
    into FMAs, so the same value computed by two entry points can differ in the
    last bits."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = np.abs(b).reshape(len(b), -1).max(axis=1)
    err = np.abs(a - b).reshape(len(b), -1).max(axis=1) / np.maximum(scale, 1e-6)
    assert err.max() < rel, err.max()


def test_cpp_mirror_program(rbd, oracle):
    assert os.path.exists(BIN), "tests/cpp/test_mirror is not built (run __graft_entry__.build())"
    ref = oracle.fdgrad(np.full((1, 21), np.float32(0.1)).astype(np.float64))[0]
    res = subprocess.run([BIN] + [repr(float(v)) for v in ref], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    assert "TEST_FOR_EQUIVALENCE" in res.stdout and res.stdout.strip().endswith("OK")


def test_forward_dynamics_entry_point(rbd, oracle):
    x = common.knots(333, seed=900)
    ctx = rbd.Context(precision="fp32")
    qdd, minv = ctx.forward_dynamics(x)
    full = ctx.run_aos(x)
    close_f32(qdd, full[:, :7])
    for i in range(0, 333, 37):
        q_ref, m_ref = oracle.fd(*(x[i, s].astype(np.float64) for s in (slice(0, 7), slice(7, 14), slice(14, 21))))
        np.testing.assert_allclose(minv[i], m_ref, rtol=0, atol=2e-6 * np.abs(m_ref).max())
        np.testing.assert_allclose(qdd[i], q_ref, rtol=0, atol=2e-6 * np.abs(q_ref).max())
    ctx.close()


@pytest.mark.parametrize("lanes", [1, 16])
def test_dqdd_du_output(rbd, lanes):
    x = common.knots(200, seed=901)
    ctx = rbd.Context(precision="fp32", lanes_per_knot=lanes)
    out, minv = ctx.run_with_minv(x)
    close_f32(out, ctx.run_aos(x), rel=2e-4)
    _, minv_fd = ctx.forward_dynamics(x)
    close_f32(minv, minv_fd)
    np.testing.assert_array_equal(minv, minv.transpose(0, 2, 1))
    ctx.close()


@pytest.mark.parametrize("lanes", [1, 16])
def test_wafr_layout(rbd, oracle, lanes):
    """rbd_fdgrad_wafr_f32: per knot [dqdd/dq | dqdd/dqd | dqdd/du = Minv]
    (147 floats, each 7x7 column-major, as the WAFR harness reads d_dqdd,
    time_GPU_WAFR.cu:18) and qdd to its own, optional buffer; the values of
    the 105 + 49 entry point (another kernel instantiation, so to FP32
    rounding) in the WAFR places, and within the bar of the oracle."""
    x = common.knots(300, seed=902)
    ctx = rbd.Context(precision="fp32", lanes_per_knot=lanes)
    out, minv = ctx.run_with_minv(x)                      # minv [knot][row][col]
    qdd, w = ctx.run_wafr(x)
    err = common.max_block_errors(np.hstack([qdd, w[:, :98]]), out)
    assert max(err.values()) < 1e-5, err
    close_f32(w[:, 98:].reshape(-1, 7, 7).transpose(0, 2, 1), minv)   # entry (r, c) at 7c + r
    q2, w2 = ctx.run_wafr(x, with_qdd=False)
    assert q2 is None
    np.testing.assert_array_equal(w2, w)
    err = common.max_block_errors(np.hstack([qdd, w[:, :98]]), oracle.fdgrad(x.astype(np.float64)))
    assert max(err.values()) < common.REL_TOL, err
    ctx.close()
