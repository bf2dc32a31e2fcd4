"""Pins the FP64 oracle (oracle/rbd_oracle.hpp) -- CPU only.

The reference holds no golden vectors, so every function of the oracle is
checked against something it does not share code with:
  * model constants and transforms   <- URDF fields + textbook spatial algebra
  * RNEA                              <- textbook RNEA on the URDF model
  * gravity torque                    <- finite-difference gradient of potential energy
  * Minv                              <- M built column-wise from RNEA, Minv M = I, symmetry, SPD
  * qdd = FD(q, qd, u)                <- RNEA(q, qd, qdd) = u
  * dqdd/dq, dqdd/dqd                 <- complex-step derivatives of FD, central differences,
                                         qd = 0 => dqdd/dqd = 0 (no damping)
  * reference-compatible mode         <- the full path on the same knot
  * FP32 instantiation                <- FP64 within the FP32 bar
Each check would fail for a plausible mistake (a dropped term, a wrong sign or
index, a transposed operand): see test_oracle_mutation_is_caught.
"""
import numpy as np
import pytest

import common


@pytest.fixture(scope="module")
def model():
    return common.UrdfModel()


def test_inertias_match_urdf(oracle, model):
    I = oracle.inertias()
    for k in range(7):
        np.testing.assert_allclose(I[k], model.inertia(k), rtol=0, atol=1e-15)
        np.testing.assert_allclose(I[k], I[k].T, atol=0)


def test_transforms_match_urdf(oracle, model):
    rng = np.random.default_rng(7)
    for _ in range(20):
        q = rng.uniform(-np.pi, np.pi, 7)
        X = oracle.transforms(q)
        for k in range(7):
            np.testing.assert_allclose(X[k], model.transform(k, q[k]), atol=1e-14)
            E = X[k][:3, :3]
            np.testing.assert_allclose(E @ E.T, np.eye(3), atol=1e-14)
            np.testing.assert_array_equal(X[k][:3, 3:], 0)


@pytest.mark.parametrize("flags", [0, 2])
def test_rnea_matches_textbook(oracle, model, flags):
    rng = np.random.default_rng(8)
    g = 0.0 if flags & 2 else 9.81
    for _ in range(20):
        q, qd, qdd = rng.uniform(-np.pi, np.pi, 7), rng.uniform(-2, 2, 7), rng.uniform(-30, 30, 7)
        np.testing.assert_allclose(oracle.rnea(q, qd, qdd, flags), model.rnea(q, qd, qdd, gravity=g),
                                   rtol=1e-12, atol=1e-11)


def test_gravity_torque_is_potential_gradient(oracle, model):
    rng = np.random.default_rng(9)
    h = 1e-6
    for _ in range(10):
        q = rng.uniform(-np.pi, np.pi, 7)
        g = oracle.rnea(q, np.zeros(7), np.zeros(7))
        grad = np.array([(model.potential(q + h * e) - model.potential(q - h * e)) / (2 * h) for e in np.eye(7)])
        np.testing.assert_allclose(g, grad, rtol=1e-7, atol=1e-7)


def test_power_balance(oracle):
    """qd^T (tau(q,qd,0) - g(q)) = 1/2 qd^T Mdot qd, Mdot along qd by central differences."""
    rng = np.random.default_rng(10)
    h = 1e-6
    for _ in range(10):
        q, qd = rng.uniform(-np.pi, np.pi, 7), rng.uniform(-2, 2, 7)
        z = np.zeros(7)
        c = oracle.rnea(q, qd, z) - oracle.rnea(q, z, z)

        def mass(qq):
            g0 = oracle.rnea(qq, z, z)
            return np.stack([oracle.rnea(qq, z, e) - g0 for e in np.eye(7)], axis=1)
        Mdot = (mass(q + h * qd) - mass(q - h * qd)) / (2 * h)
        assert abs(qd @ c - 0.5 * qd @ Mdot @ qd) < 1e-6 * max(1.0, abs(qd @ c))


def test_minv_is_inverse_of_mass_matrix(oracle):
    rng = np.random.default_rng(11)
    z = np.zeros(7)
    for i in range(30):
        q = rng.uniform(-np.pi, np.pi, 7)
        if i % 3 == 0:
            q[[1, 3, 5]] = rng.uniform(-1e-3, 1e-3, 3)      # stretched arm
        g0 = oracle.rnea(q, z, z)
        M = np.stack([oracle.rnea(q, z, e) - g0 for e in np.eye(7)], axis=1)
        _, Minv = oracle.fd(q, z, z)
        np.testing.assert_allclose(Minv @ M, np.eye(7), atol=1e-10)
        np.testing.assert_array_equal(Minv, Minv.T)
        assert np.all(np.linalg.eigvalsh(Minv) > 0)
        assert np.all(oracle.pivots(q) > 0)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_forward_dynamics_inverts_rnea(oracle, flags):
    x = common.knots(50, seed=12).astype(np.float64)
    out = oracle.fdgrad(x, flags)
    for i in range(50):
        q, qd, u = x[i, :7], x[i, 7:14], x[i, 14:]
        np.testing.assert_allclose(oracle.rnea(q, qd, out[i, :7], flags), u, rtol=0, atol=1e-10)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("near_singular", [0.0, 1.0])
def test_gradient_matches_complex_step(oracle, flags, near_singular):
    x = common.knots(40, seed=13, near_singular_frac=near_singular).astype(np.float64)
    out = oracle.fdgrad(x, flags)
    for i in range(40):
        cs = oracle.fd_complex_step(x[i], flags=flags)
        err = common.max_block_errors(out[i], cs)
        assert max(err.values()) < 1e-11, err


def test_gradient_matches_central_differences(oracle):
    x = common.knots(5, seed=14).astype(np.float64)
    h = 1e-6
    for i in range(5):
        out = oracle.fdgrad(x[i])[0]
        for j in range(14):
            e = np.zeros(21)
            e[j] = h
            d = (oracle.fdgrad(x[i] + e)[0, :7] - oracle.fdgrad(x[i] - e)[0, :7]) / (2 * h)
            scale = np.abs(out[7:]).max()
            np.testing.assert_allclose(out[7 + 7 * j: 14 + 7 * j], d, rtol=1e-6, atol=1e-7 * scale)


def test_zero_velocity_has_zero_velocity_gradient(oracle):
    x = common.knots(20, seed=15).astype(np.float64)
    x[:, 7:14] = 0.0
    out = oracle.fdgrad(x)
    assert np.abs(out[:, 56:]).max() < 1e-12
    # with damping the velocity gradient is exactly -0.5 Minv
    out_d = oracle.fdgrad(x, oracle.VEL_DAMPING)
    for i in range(20):
        _, Minv = oracle.fd(x[i, :7], x[i, 7:14], x[i, 14:], oracle.VEL_DAMPING)
        np.testing.assert_allclose(out_d[i, 56:].reshape(7, 7).T, -0.5 * Minv, atol=1e-12)


def test_first_joint_column_is_zero(oracle):
    """Rotating joint 1 about the (vertical) gravity axis changes nothing: dqdd/dq_0 = 0."""
    out = oracle.fdgrad(common.knots(20, seed=16).astype(np.float64))
    assert np.abs(out[:, 7:14]).max() < 1e-12


def test_reference_mode_matches_full_path(oracle):
    x = common.knots(30, seed=17).astype(np.float64)
    full = oracle.fdgrad(x)
    vol = np.zeros((30, 77))
    for i in range(30):
        qdd, Minv = oracle.fd(x[i, :7], x[i, 7:14], x[i, 14:])
        vol[i, :21], vol[i, 21:28], vol[i, 28:] = x[i], qdd, Minv.T.reshape(-1)
    np.testing.assert_allclose(oracle.fdgrad_refmode(vol), full[:, 7:], rtol=1e-12, atol=1e-10)


def test_fp32_instantiation_within_bar(oracle):
    x = common.knots(500, seed=18)
    ref = oracle.fdgrad(x.astype(np.float64))
    err = common.max_block_errors(oracle.fdgrad_f32(x), ref)
    assert max(err.values()) < common.REL_TOL, err


def test_multithreaded_oracle_matches_serial(oracle):
    x = common.knots(333, seed=19).astype(np.float64)
    np.testing.assert_array_equal(oracle.fdgrad(x, threads=4), oracle.fdgrad(x))


@pytest.mark.parametrize("flags", [0, 2])
def test_external_forces_do_virtual_power(oracle, model, flags):
    """A force fext_k on link k (its own frame) changes the joint torques by
    -J^T fext: the power it delivers, sum_k fext_k . v_k, equals -tau_ext . qd
    for every qd; with qd = e_j that gives tau_ext[j] = -sum_k fext_k . v_k(e_j),
    with the link velocities v_k from the textbook model."""
    rng = np.random.default_rng(41)
    for _ in range(5):
        q, qd, qdd = rng.uniform(-3, 3, 7), rng.uniform(-2, 2, 7), rng.uniform(-5, 5, 7)
        fext = rng.uniform(-20, 20, 42)
        d_tau = oracle.rnea(q, qd, qdd, flags, fext=fext) - oracle.rnea(q, qd, qdd, flags)
        want = np.array([-sum(fext[6 * k:6 * k + 6] @ v for k, v in enumerate(model.link_velocities(q, e)))
                         for e in np.eye(7)])
        np.testing.assert_allclose(d_tau, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_forward_dynamics_with_external_forces_inverts_rnea(oracle):
    rng = np.random.default_rng(42)
    for flags in (0, 1, 2):
        q, qd, u = rng.uniform(-3, 3, 7), rng.uniform(-2, 2, 7), rng.uniform(-50, 50, 7)
        fext = rng.uniform(-20, 20, 42)
        qdd, _ = oracle.fd(q, qd, u, flags, fext=fext)
        tau = oracle.rnea(q, qd, qdd, flags, fext=fext)
        np.testing.assert_allclose(tau, u, rtol=0, atol=1e-10 * np.abs(u).max())


@pytest.mark.parametrize("flags", [0, 1, 2])
def test_external_force_gradient_matches_complex_step(oracle, flags):
    x = common.knots(12, seed=43, near_singular_frac=0.25).astype(np.float64)
    fext = np.random.default_rng(44).uniform(-20, 20, 42)
    out = oracle.fdgrad(x, flags, fext=fext)
    for i in range(len(x)):
        cs = oracle.fd_complex_step(x[i], flags=flags, fext=fext)
        err = common.max_block_errors(out[i], cs)
        assert max(err.values()) < 1e-11, err
    # fext changes the result (the test above is not vacuous)
    assert np.abs(out - oracle.fdgrad(x, flags)).max() > 1e-3


def test_per_knot_external_forces(oracle):
    x = common.knots(6, seed=45).astype(np.float64)
    fext = np.random.default_rng(46).uniform(-20, 20, (6, 42))
    out = oracle.fdgrad(x, fext=fext)
    for i in range(6):
        np.testing.assert_array_equal(out[i], oracle.fdgrad(x[i:i + 1], fext=fext[i])[0])


def fused_inputs(n, seed, oracle, flags=0):
    """(n, 28) [q|qd|u|qdd] with qdd = FD(q, qd, u), as the reference prepares them."""
    x = common.knots(n, seed=seed, near_singular_frac=0.2).astype(np.float64)
    qdd = oracle.fdgrad(x, flags)[:, :7]
    return np.concatenate([x, qdd], axis=1)


@pytest.mark.parametrize("flags", [0, 1, 2])
def test_id_gradient_matches_complex_step(oracle, flags):
    vol = fused_inputs(20, 47, oracle, flags)
    vol[::3, 21:] = np.random.default_rng(48).uniform(-30, 30, (7, 7))   # also off the dynamics manifold
    cdu = oracle.id_gradient(vol, flags)
    for i in range(len(vol)):
        cs = oracle.rnea_complex_step(vol[i], flags=flags)
        scale = np.abs(cs).max()
        np.testing.assert_allclose(cdu[i], cs, rtol=0, atol=1e-12 * scale)


def test_split_stage_matches_fused(oracle, model):
    vol = fused_inputs(10, 49, oracle)
    split = oracle.vaf(np.concatenate([vol, np.zeros((10, 126))], axis=1))
    np.testing.assert_array_equal(oracle.id_gradient_split(split), oracle.id_gradient(vol))
    # v is the textbook link velocity; f[k][2] after the backward pass is tau
    for i in range(10):
        q, qd, qdd = vol[i, :7], vol[i, 7:14], vol[i, 21:]
        np.testing.assert_allclose(split[i, 28:70].reshape(7, 6), model.link_velocities(q, qd), rtol=0, atol=1e-12)
        np.testing.assert_allclose(split[i, 112:154].reshape(7, 6)[:, 2], model.rnea(q, qd, qdd), rtol=0,
                                   atol=1e-10 * np.abs(model.rnea(q, qd, qdd)).max())


@pytest.mark.parametrize("flags", [0, 1])
def test_finish_of_id_gradient_is_the_full_path(oracle, flags):
    """dqdd = -Minv dtau/dx at qdd = FD(q, qd, u): the staged pipeline gives the
    full path's gradient."""
    vol = fused_inputs(15, 50, oracle, flags)
    minv = np.stack([oracle.fd(v[:7], v[7:14], v[14:21], flags)[1].T.ravel() for v in vol])
    dqdd = oracle.finish(minv, oracle.id_gradient(vol, flags))
    full = oracle.fdgrad(vol[:, :21], flags)[:, 7:]
    np.testing.assert_allclose(dqdd, full, rtol=0, atol=1e-11 * np.abs(full).max())


def test_op_count_is_frozen(oracle):
    """The algorithmic FLOP count per knot used for the roofline (bench.py)."""
    c = oracle.op_count(0)
    assert c["add"] + c["mul"] + c["div"] == 31340 and c["trig"] == 14 and c["div"] == 7
    s = oracle.op_count(0, skip_zeros=True)
    assert s["add"] + s["mul"] + s["div"] == 14459


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5])
def test_oracle_mutation_is_caught(oracle, model, mutation):
    """Each planted mistake (a dropped term, a flipped sign, a transposed operand,
    a skipped step) makes one of the checks above fail."""
    x = common.knots(4, seed=20).astype(np.float64)
    try:
        oracle.set_mutation(mutation)
        out = oracle.fdgrad(x)
        cs = np.array([oracle.fd_complex_step(x[i]) for i in range(4)])
        grad_err = max(common.max_block_errors(out, cs).values())
        q, qd, qdd = x[0, :7], x[0, 7:14], out[0, :7]
        rnea_err = np.abs(oracle.rnea(q, qd, qdd) - model.rnea(q, qd, qdd)).max()
        z = np.zeros(7)
        g0 = oracle.rnea(q, z, z)
        M = np.stack([oracle.rnea(q, z, e) - g0 for e in np.eye(7)], axis=1)
        minv_err = np.abs(oracle.fd(q, z, z)[1] @ M - np.eye(7)).max()
    finally:
        oracle.set_mutation(0)
    assert max(grad_err, rnea_err, minv_err) > 1e-4, (grad_err, rnea_err, minv_err)
