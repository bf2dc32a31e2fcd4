"""CPU check of the kernel arithmetic: fast-rbd-gradients_amd/csrc/fdgrad_knot.hpp
compiled for the host (tests/hostcheck) against the URDF data and the FP64
oracle, in every precision mode, before anything runs on the GPU."""
import ctypes
import os

import numpy as np
import pytest

import common



@pytest.fixture(scope="module")
def host():
    return common.hostcheck_lib()


def run(host, x, flags=0, prec=0, lanes=0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.full((x.shape[0], 105), np.nan, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    host.host_fdgrad_f32(x.shape[0], x.ctypes.data_as(fp), out.ctypes.data_as(fp), flags, prec, lanes)
    return out


def test_device_model_tables_match_urdf(host):
    model = common.UrdfModel()
    buf = np.zeros(36)
    dp = ctypes.POINTER(ctypes.c_double)
    for k in range(7):
        host.host_model_inertia(k, buf.ctypes.data_as(dp))
        np.testing.assert_allclose(buf.reshape(6, 6).T, model.inertia(k), atol=1e-7)   # float32 constants
        for q in (-2.5, -0.3, 0.0, 1.1, 3.0):
            host.host_model_transform(k, q, buf.ctypes.data_as(dp))
            np.testing.assert_allclose(buf.reshape(6, 6).T, model.transform(k, q), atol=1e-15)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_host_build_matches_oracle(host, oracle, flags, prec):
    x = common.knots(400, seed=200 + flags)
    ref = oracle.fdgrad(x.astype(np.float64), flags)
    err = common.max_block_errors(run(host, x, flags, prec), ref)
    bound = common.REL_TOL if prec < 3 else 1e-6
    assert max(err.values()) < bound, err


def test_fp64_columns_fix_near_singular_poses(host, oracle):
    """Stretched-arm poses: FP32 sits near the bar on dqdd/dqd; FP64 columns are
    far below it (the measured reason for RBD_PREC_FP64_COLUMNS)."""
    x = common.knots(3000, seed=210, near_singular_frac=1.0)
    ref = oracle.fdgrad(x.astype(np.float64), threads=8)
    e32 = common.max_block_errors(run(host, x, 0, 0), ref)
    e64c = common.max_block_errors(run(host, x, 0, 2), ref)
    assert max(e64c.values()) < 1e-5, e64c
    assert e64c["dqdd_dqd"] < e32["dqdd_dqd"] / 4


def test_reference_mode(host, oracle):
    g = np.load(os.path.join(common.GOLDEN, "fdgrad_golden.npz"))
    vol, ref = g["refmode_n128__in"], g["refmode_n128__out"]
    out = np.zeros((vol.shape[0], 98), dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    for lanes in (0, 1):
        out[:] = np.nan
        host.host_fdgrad_refmode_f32(vol.shape[0], np.ascontiguousarray(vol).ctypes.data_as(fp),
                                     out.ctypes.data_as(fp), 0, 0, lanes)
        err = common.dqdd_block_errors(out, ref)
        assert max(float(v.max()) for v in err.values()) < common.REL_TOL


def test_golden_sets(host):
    g = np.load(os.path.join(common.GOLDEN, "fdgrad_golden.npz"))
    names = sorted({k.split("__")[0] for k in g.files if not k.startswith("refmode")})
    for name in names:
        x, flags, ref = g[name + "__in"], int(g[name + "__flags"]), g[name + "__out"]
        err = common.max_block_errors(run(host, x, flags, 2), ref)
        assert max(err.values()) < common.REL_TOL, (name, err)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("prec", [0, 2])
def test_lane_group_mapping_matches_single_lane(host, oracle, flags, prec):
    """The lane-group mapping writes every output value (none left NaN) and
    agrees with the one-lane mapping and with the oracle."""
    x = common.knots(200, seed=220 + flags)
    one = run(host, x, flags, prec, lanes=0)
    grp = run(host, x, flags, prec, lanes=1)
    assert np.all(np.isfinite(grp))
    np.testing.assert_allclose(grp, one, rtol=1e-5, atol=1e-6 * np.abs(one).max())
    err = common.max_block_errors(grp, oracle.fdgrad(x.astype(np.float64), flags))
    assert max(err.values()) < common.REL_TOL, err


@pytest.mark.parametrize("flags", [0, 1, 2])
def test_forward_dynamics_and_minv_outputs(host, oracle, flags):
    x = common.knots(100, seed=230 + flags)
    fp = ctypes.POINTER(ctypes.c_float)
    xc = np.ascontiguousarray(x)
    qdd = np.zeros((100, 7), np.float32)
    minv = np.zeros((100, 49), np.float32)
    host.host_fd_f32(100, xc.ctypes.data_as(fp), qdd.ctypes.data_as(fp), minv.ctypes.data_as(fp), flags)
    full = run(host, x, flags)
    for lanes in (0, 1):
        out2 = np.full((100, 105), np.nan, np.float32)
        minv2 = np.full((100, 49), np.nan, np.float32)
        host.host_fdgrad_minv_f32(100, xc.ctypes.data_as(fp), out2.ctypes.data_as(fp), minv2.ctypes.data_as(fp),
                                  flags, lanes)
        np.testing.assert_array_equal(minv2, minv)             # same Minv arithmetic
        assert np.all(np.isfinite(out2))
    np.testing.assert_array_equal(qdd, full[:, :7])
    for i in range(0, 100, 9):
        q_ref, m_ref = oracle.fd(x[i, :7].astype(np.float64), x[i, 7:14].astype(np.float64),
                                 x[i, 14:].astype(np.float64), flags)
        np.testing.assert_allclose(minv[i].reshape(7, 7).T, m_ref, rtol=0, atol=1e-5 * np.abs(m_ref).max())
        np.testing.assert_allclose(qdd[i], q_ref, rtol=0, atol=1e-5 * np.abs(q_ref).max())


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_wafr_layout(host, oracle, flags):
    """The WAFR harness's gradient layout (time_GPU_WAFR.cu:18; finish_dqdd_v2,
    helpers_WAFR/dynamics_WAFR.cuh:2555-2579): per knot 147 floats, dqdd/dq,
    dqdd/dqd and dqdd/du = Minv, each 7x7 column-major (entry (r, c) at 7c + r),
    and qdd on its own: the full path's values placed there, and Minv against
    the oracle's forward dynamics."""
    n = 60
    x = np.ascontiguousarray(common.knots(n, seed=240 + flags))
    fp = ctypes.POINTER(ctypes.c_float)
    full = run(host, x, flags)
    for lanes in (0, 1):
        qdd = np.full((n, 7), np.nan, np.float32)
        w = np.full((n, 147), np.nan, np.float32)
        host.host_fdgrad_wafr_f32(n, x.ctypes.data_as(fp), qdd.ctypes.data_as(fp), w.ctypes.data_as(fp), flags,
                                  lanes)
        ref = full if lanes == 0 else run(host, x, flags, lanes=1)
        np.testing.assert_array_equal(qdd, ref[:, :7])
        np.testing.assert_array_equal(w[:, :98], ref[:, 7:])
        w2 = np.full((n, 147), np.nan, np.float32)
        host.host_fdgrad_wafr_f32(n, x.ctypes.data_as(fp), None, w2.ctypes.data_as(fp), flags, lanes)   # no qdd
        np.testing.assert_array_equal(w2, w)
        for i in range(0, n, 7):
            _, m_ref = oracle.fd(x[i, :7].astype(np.float64), x[i, 7:14].astype(np.float64),
                                 x[i, 14:].astype(np.float64), flags)
            m = w[i, 98:].reshape(7, 7).T                                   # column-major -> [row][col]
            np.testing.assert_allclose(m, m_ref, rtol=0, atol=1e-5 * np.abs(m_ref).max())


@pytest.mark.parametrize("flags", [0, 1, 2])
@pytest.mark.parametrize("per_knot", [False, True])
def test_external_forces(host, oracle, flags, per_knot):
    n = 120
    x = np.ascontiguousarray(common.knots(n, seed=260 + flags, near_singular_frac=0.1))
    rng = np.random.default_rng(261)
    fext = rng.uniform(-20, 20, (n, 42) if per_knot else 42).astype(np.float32)
    stride = 42 if per_knot else 0
    fp = ctypes.POINTER(ctypes.c_float)
    ref = oracle.fdgrad(x.astype(np.float64), flags, fext=fext.astype(np.float64))
    for lanes in (0, 1):
        out = np.full((n, 105), np.nan, np.float32)
        host.host_fdgrad_fext_f32(n, x.ctypes.data_as(fp), fext.ctypes.data_as(fp), stride, out.ctypes.data_as(fp),
                                  flags, lanes)
        err = common.max_block_errors(out, ref)
        assert max(err.values()) < common.REL_TOL, (lanes, err)
    qdd = np.zeros((n, 7), np.float32)
    minv = np.zeros((n, 49), np.float32)
    host.host_fd_fext_f32(n, x.ctypes.data_as(fp), fext.ctypes.data_as(fp), stride, qdd.ctypes.data_as(fp),
                          minv.ctypes.data_as(fp), flags)
    for i in range(0, n, 13):
        f = fext[i] if per_knot else fext
        q_ref, _ = oracle.fd(*(x[i, s].astype(np.float64) for s in (slice(0, 7), slice(7, 14), slice(14, 21))),
                             flags, fext=f.astype(np.float64))
        np.testing.assert_allclose(qdd[i], q_ref, rtol=0, atol=1e-5 * np.abs(q_ref).max())
    # the forces matter: the result without them is far off
    plain = run(host, x, flags)
    assert np.abs(plain - out).max() > 1e-2


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


@pytest.mark.parametrize("flags", [0, 1, 2])
def test_staged_variants(host, oracle, flags):
    """Fused and split inverse-dynamics gradient, the v/a/f stage and the
    finish stage against the oracle's."""
    n = 150
    x = common.knots(n, seed=270 + flags, near_singular_frac=0.1).astype(np.float64)
    qdd = oracle.fdgrad(x, flags)[:, :7]
    vol = np.concatenate([x, qdd], axis=1)
    ref_cdu = oracle.id_gradient(vol, flags)
    vol32 = np.ascontiguousarray(vol, dtype=np.float32)
    split32 = np.ascontiguousarray(np.concatenate([vol32, np.zeros((n, 126), np.float32)], axis=1))
    host.host_vaf_f32(n, _p(split32), flags)
    ref_split = oracle.vaf(np.concatenate([vol32.astype(np.float64), np.zeros((n, 126))], axis=1), flags)
    np.testing.assert_allclose(split32, ref_split, rtol=0, atol=2e-5 * np.abs(ref_split).max())
    for split, inp in ((0, vol32), (1, split32)):
        for lanes in (0, 1):
            cdu = np.full((n, 98), np.nan, np.float32)
            host.host_id_gradient_f32(n, _p(inp), _p(cdu), flags, split, lanes)
            err = {k: float(v.max()) for k, v in common.dqdd_block_errors(cdu, ref_cdu).items()}
            assert max(err.values()) < common.REL_TOL, (split, lanes, err)
    minv = np.stack([oracle.fd(v[:7], v[7:14], v[14:21], flags)[1].T.ravel() for v in x])
    minv32 = np.ascontiguousarray(minv, dtype=np.float32)
    cdu32 = np.ascontiguousarray(ref_cdu, dtype=np.float32)
    dqdd = np.full((n, 98), np.nan, np.float32)
    host.host_finish_f32(n, _p(minv32), _p(cdu32), _p(dqdd))
    ref = oracle.finish(minv32.astype(np.float64), cdu32.astype(np.float64))
    err = {k: float(v.max()) for k, v in common.dqdd_block_errors(dqdd, ref).items()}
    assert max(err.values()) < 1e-6, err


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_two_wave_mapping(host, oracle, flags, prec):
    """The two-wave latency mapping (Minv on one wave, RNEA and columns with
    the bias accelerations plus qdd corrections on the other) against the
    oracle, at random and stretched-arm poses."""
    n = 300
    x = np.ascontiguousarray(common.knots(n, seed=280 + flags, near_singular_frac=0.1))
    out = np.full((n, 105), np.nan, np.float32)
    host.host_fdgrad_two_wave_f32(n, _p(x), _p(out), flags, prec)
    assert np.all(np.isfinite(out))
    err = common.max_block_errors(out, oracle.fdgrad(x.astype(np.float64), flags))
    bar = common.REL_TOL if prec in (0, 1) else 1e-5
    assert max(err.values()) < bar, err
    # the same values as the one-knot-per-lane-group mapping up to rounding
    ref = run(host, x, flags, prec, lanes=1)
    assert np.abs(out - ref).max() <= 1e-3 * np.abs(ref).max()


def test_minv_by_columns_equals_the_whole_recursion(host):
    """compute_minv_column (one column per lane, the two-wave mapping's Minv)
    gives each upper entry exactly as compute_minv does."""
    rng = np.random.default_rng(290)
    q = rng.uniform(-np.pi, np.pi, (200, 7))
    q[::4, [1, 3, 5]] = rng.uniform(-1e-3, 1e-3, (50, 3))
    full = np.full(200 * 49, np.nan)
    cols = np.full(200 * 49, np.nan)
    dp = ctypes.POINTER(ctypes.c_double)
    host.host_minv_columns(200, np.ascontiguousarray(q).ctypes.data_as(dp), full.ctypes.data_as(dp),
                           cols.ctypes.data_as(dp))
    upper = np.triu(np.ones((7, 7), bool)).ravel()
    full, cols = full.reshape(200, 49)[:, upper], cols.reshape(200, 49)[:, upper]
    assert np.all(np.isfinite(full))
    np.testing.assert_array_equal(cols, full)


@pytest.mark.parametrize("spins,set_after,seen,polls,pauses", [
    (10, 0, 1, 1, 0),         # flag already set: one poll, no sleep
    (10, 3, 1, 4, 3),         # set after three clear polls: seen on the fourth
    (10, 10, 1, 11, 10),      # seen on the last poll the bound allows
    (10, 11, 0, 11, 10),      # one poll too late: the wait gives up
    (10, -1, 0, 11, 10),      # never set: gives up after spins + 1 polls
    (0, -1, 0, 1, 0),         # spins 0: one poll
    (-1, 0, 0, 0, 0),         # spins < 0: gives up at once without polling
])
def test_handoff_wait_bound(host, spins, set_after, seen, polls, pauses):
    """The two-wave mapping's LDS handoff wait (bounded_wait, the loop of
    lds_wait): it returns as soon as the flag is seen, and gives up -- which
    makes the kernel write its context's failure word -- after spins + 1
    polls, or at once for spins < 0."""
    p, z = ctypes.c_int(-1), ctypes.c_int(-1)
    assert host.host_bounded_wait(spins, set_after, ctypes.byref(p), ctypes.byref(z)) == seen
    assert (p.value, z.value) == (polls, pauses)


@pytest.mark.parametrize("lanes", [4, 7, 8])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
@pytest.mark.parametrize("flags", [0, 3])
def test_group_mapping(host, oracle, flags, prec, lanes):
    """The L-lane group mappings (4, 7, 8 lanes per knot: each lane a share of
    the joints' sin/cos, of Minv and of the 14 column tasks) give every output
    exactly as the 16-lane groups do -- the same arithmetic, split another way
    -- and every output is written; against the oracle at random and
    stretched-arm poses."""
    n = 200
    x = np.ascontiguousarray(common.knots(n, seed=295 + flags, near_singular_frac=0.2))
    out = np.full((n, 105), np.nan, np.float32)
    host.host_fdgrad_group_f32(n, _p(x), _p(out), flags, prec, lanes)
    np.testing.assert_array_equal(out, run(host, x, flags, prec, lanes=1))
    err = common.max_block_errors(out, oracle.fdgrad(x.astype(np.float64), flags))
    assert max(err.values()) < (common.REL_TOL if prec in (0, 1) else 1e-5), err
