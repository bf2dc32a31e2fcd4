"""Robots from a URDF (fast-rbd-gradients_amd/urdf_model.py) -- CPU only.

The reference hard-codes the iiwa's tables (helpers_CPU/dynamicsGradient.h:123-229,
helpers_GPU/dynamicsGradient.cuh:211-310); its README.md:1 points to GRiD for
other robots.  Here a URDF becomes a model header that the kernels are compiled on.
What pins it:
  * on the iiwa's URDF data the generator reproduces csrc/iiwa_model.hpp
    exactly: the kernel arithmetic built on the generated header gives
    bit-identical outputs;
  * on tests/golden/synthetic_arm7.urdf (joint frames at general angles, axes
    other than +z, fixed links lumped into the chain, a joint hung off a fixed
    link, a dropped pedestal) the oracle, computing on the generated tables,
    agrees with a textbook RNEA over the whole URDF tree in the URDF's own
    frames (common.UrdfTree); its gravity torque is the gradient of the
    textbook potential energy, its Minv inverts the textbook mass matrix, and
    its analytic gradient matches complex-step derivatives;
  * the kernel arithmetic (host build) on the synthetic arm matches the oracle
    in every precision mode and knot-to-lane mapping;
  * URDFs the kernels cannot take are refused with a message.
"""
import ctypes
import os

import numpy as np
import pytest

import common

ARM7 = os.path.join(common.GOLDEN, "synthetic_arm7.urdf")
IIWA_JSON = os.path.join(common.GOLDEN, "iiwa_14_urdf.json")


@pytest.fixture(scope="module")
def arm7():
    return common.generate_model(ARM7, "arm7")


@pytest.fixture(scope="module")
def tree():
    return common.UrdfTree(ARM7)


@pytest.fixture
def on_arm7(oracle, arm7):
    with oracle.using_model(*arm7):
        yield oracle


@pytest.fixture(scope="module")
def host_arm7(arm7):
    return common.hostcheck_lib("arm7")


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def run(host, x, flags=0, prec=0, lanes=0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.full((x.shape[0], 105), np.nan, dtype=np.float32)
    host.host_fdgrad_f32(x.shape[0], _p(x), _p(out), flags, prec, lanes)
    return out


def run_two_wave(host, x, flags=0, prec=0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.full((x.shape[0], 105), np.nan, dtype=np.float32)
    host.host_fdgrad_two_wave_f32(x.shape[0], _p(x), _p(out), flags, prec)
    return out


def test_generated_iiwa_header_is_bit_identical():
    common.generate_model(IIWA_JSON, "iiwa_gen")
    gen = common.hostcheck_lib("iiwa_gen")
    base = common.hostcheck_lib()
    assert gen.host_model_name() == b"iiwa_gen" and base.host_model_name() == b"iiwa14"
    x = common.knots(300, seed=600, near_singular_frac=0.1)
    for flags in (0, 1, 2, 3):
        for prec in (0, 1, 2, 3):
            for lanes in (0, 1):
                np.testing.assert_array_equal(run(gen, x, flags, prec, lanes), run(base, x, flags, prec, lanes))
            np.testing.assert_array_equal(run_two_wave(gen, x, flags, prec), run_two_wave(base, x, flags, prec))


@pytest.mark.parametrize("gravity", [9.81, 0.0])
def test_oracle_on_urdf_model_matches_textbook_tree(on_arm7, tree, gravity):
    flags = 0 if gravity else on_arm7.NO_GRAVITY
    rng = np.random.default_rng(601)
    for _ in range(30):
        q, qd, qdd = rng.uniform(-np.pi, np.pi, 7), rng.uniform(-2, 2, 7), rng.uniform(-30, 30, 7)
        ref = tree.rnea(q, qd, qdd, gravity=gravity)
        np.testing.assert_allclose(on_arm7.rnea(q, qd, qdd, flags), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_oracle_gravity_torque_is_potential_gradient(on_arm7, tree):
    rng = np.random.default_rng(602)
    h = 1e-6
    z = np.zeros(7)
    for _ in range(10):
        q = rng.uniform(-np.pi, np.pi, 7)
        grad = np.array([(tree.potential(q + h * e) - tree.potential(q - h * e)) / (2 * h) for e in np.eye(7)])
        np.testing.assert_allclose(on_arm7.rnea(q, z, z), grad, rtol=1e-7, atol=1e-7)


def test_oracle_minv_inverts_textbook_mass_matrix(on_arm7, tree):
    rng = np.random.default_rng(603)
    z = np.zeros(7)
    for _ in range(20):
        q = rng.uniform(-np.pi, np.pi, 7)
        _, Minv = on_arm7.fd(q, z, z)
        np.testing.assert_allclose(Minv @ tree.mass_matrix(q), np.eye(7), atol=1e-10)
        assert np.all(on_arm7.pivots(q) > 0)


@pytest.mark.parametrize("flags", [0, 2])
def test_oracle_forward_dynamics_inverts_textbook_rnea(on_arm7, tree, flags):
    x = common.knots(40, seed=604).astype(np.float64)
    out = on_arm7.fdgrad(x, flags)
    g = 0.0 if flags & 2 else 9.81
    for i in range(40):
        np.testing.assert_allclose(tree.rnea(x[i, :7], x[i, 7:14], out[i, :7], gravity=g), x[i, 14:],
                                   rtol=0, atol=1e-9)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_oracle_gradient_matches_complex_step_on_urdf_model(on_arm7, flags):
    x = common.knots(30, seed=605).astype(np.float64)
    out = on_arm7.fdgrad(x, flags)
    for i in range(30):
        err = common.max_block_errors(out[i], on_arm7.fd_complex_step(x[i], flags=flags))
        assert max(err.values()) < 1e-11, err


def test_device_tables_on_urdf_model(host_arm7, on_arm7):
    """The compile-time inertias and transforms of the generated header equal
    the oracle's on the same tables."""
    buf = np.zeros(36)
    dp = ctypes.POINTER(ctypes.c_double)
    I = on_arm7.inertias()
    assert host_arm7.host_model_name() == b"arm7"
    for k in range(7):
        host_arm7.host_model_inertia(k, buf.ctypes.data_as(dp))
        np.testing.assert_allclose(buf.reshape(6, 6).T, I[k], rtol=1e-7, atol=1e-9)   # float32 constants
        for q in (-2.5, 0.0, 1.1):
            host_arm7.host_model_transform(k, q, buf.ctypes.data_as(dp))
            np.testing.assert_allclose(buf.reshape(6, 6).T, on_arm7.transforms(np.full(7, q))[k], atol=1e-15)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_host_build_on_urdf_model_matches_oracle(host_arm7, on_arm7, flags, prec):
    x = common.knots(300, seed=610 + flags)
    ref = on_arm7.fdgrad(x.astype(np.float64), flags)
    bar = common.REL_TOL if prec < 2 else 1e-5
    for lanes in (0, 1):
        out = run(host_arm7, x, flags, prec, lanes)
        assert np.all(np.isfinite(out))
        err = common.max_block_errors(out, ref)
        assert max(err.values()) < bar, (lanes, err)
    err = common.max_block_errors(run_two_wave(host_arm7, x, flags, prec), ref)
    assert max(err.values()) < bar, ("two-wave", err)


@pytest.mark.parametrize("flags", [0, 1])
def test_host_build_on_urdf_model_external_forces(host_arm7, on_arm7, flags):
    n = 100
    x = np.ascontiguousarray(common.knots(n, seed=620 + flags))
    fext = np.random.default_rng(621).uniform(-20, 20, (n, 42)).astype(np.float32)
    ref = on_arm7.fdgrad(x.astype(np.float64), flags, fext=fext.astype(np.float64))
    for lanes in (0, 1):
        out = np.full((n, 105), np.nan, np.float32)
        host_arm7.host_fdgrad_fext_f32(n, _p(x), _p(fext), 42, _p(out), flags, lanes)
        err = common.max_block_errors(out, ref)
        assert max(err.values()) < common.REL_TOL, (lanes, err)


def _edit(tmp_path, old, new, count=1):
    with open(ARM7) as f:
        text = f.read()
    assert old in text
    p = tmp_path / "robot.urdf"
    p.write_text(text.replace(old, new, count))
    return str(p)


@pytest.mark.parametrize("old,new,msg", [
    ('<joint name="j5" type="continuous">', '<joint name="j5" type="prismatic">', "revolute joints only"),
    ('<joint name="j7" type="revolute">', '<joint name="j7" type="fixed">', "chain of 7 revolute joints"),
    ('<parent link="l5"/>\n    <child link="l6"/>', '<parent link="l4"/>\n    <child link="l6"/>', "one serial chain"),
    ('<axis xyz="0.6 0 0.8"/>', '<axis xyz="0 0 0"/>', "zero axis"),
    ('<link name="sensor"/>', '<link name="sensor"/>\n  <link name="l9"/>\n  <joint name="j8" type="revolute">'
     '<parent link="l7"/><child link="l9"/><axis xyz="0 0 1"/></joint>', "chain of 7 revolute joints"),
])
def test_unsupported_urdfs_are_refused(tmp_path, old, new, msg):
    import urdf_model
    with pytest.raises(urdf_model.ModelError, match=msg):
        urdf_model.read_urdf(_edit(tmp_path, old, new))


def test_massless_chain_link_is_refused(tmp_path):
    import urdf_model
    text = open(ARM7).read()
    a = text.index('<link name="l6">')
    b = text.index("</link>", a) + len("</link>")
    p = tmp_path / "robot.urdf"
    p.write_text(text[:a] + '<link name="l6"/>' + text[b:])
    with pytest.raises(urdf_model.ModelError, match="l6 has no inertial"):
        urdf_model.read_urdf(str(p))


def test_op_count_on_urdf_model(on_arm7):
    """The algorithmic FLOP count per knot bench.py prices the synthetic arm at."""
    import bench
    c = on_arm7.op_count(0)
    assert c["add"] + c["mul"] + c["div"] == bench.FLOPS_PER_KNOT_ARM7


@pytest.mark.parametrize("axis", [(0, 0, 1), (0, 0, -1), (1, 0, 0), (0, 1, 0), (0.6, 0, 0.8), (1, 2, -3),
                                  (0, 0, 5), (1e-9, 0, -1)])
def test_axis_frame_takes_z_to_the_axis(axis):
    import urdf_model
    R = np.array(urdf_model.axis_frame(axis))
    a = np.array(axis, float) / np.linalg.norm(axis)
    np.testing.assert_allclose(R @ [0, 0, 1], a, atol=1e-12)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    assert abs(np.linalg.det(R) - 1) < 1e-12


def test_rpy_matrix_is_the_urdf_convention():
    """R = Rz(yaw) Ry(pitch) Rx(roll), against the textbook model's own rotation."""
    import urdf_model
    rng = np.random.default_rng(640)
    for _ in range(20):
        r, p, y = rng.uniform(-np.pi, np.pi, 3)
        np.testing.assert_allclose(np.array(urdf_model.rpy_matrix(r, p, y)), common._rpy(r, p, y), atol=1e-14)
