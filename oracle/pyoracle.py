"""TEST INFRASTRUCTURE ONLY: ctypes binding of oracle/liboracle.so.

The CPU restatement of the reference's iiwa forward-dynamics gradient
(helpers_CPU/dynamicsGradient.h, helpers_CPU/minv.h) in FP64 / FP32, the
complex-step derivative used to pin it, and the CPU baseline timing loop.
Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

NQ = 7
IN_PER_KNOT = 21       # q, qd, u
OUT_PER_KNOT = 105     # qdd, dqdd/dq, dqdd/dqd
VOL_PER_KNOT = 77      # q, qd, u, qdd, Minv (reference COMPLETELY_FUSED layout)
DQDD_PER_KNOT = 98     # dqdd/dq, dqdd/dqd

VEL_DAMPING = 1
NO_GRAVITY = 2

_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        L.oracle_fdgrad_f64.argtypes = [ctypes.c_long, dp, dp, ctypes.c_int]
        L.oracle_fdgrad_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_int]
        L.oracle_fdgrad_f64_mt.argtypes = [ctypes.c_int, ctypes.c_long, dp, dp, ctypes.c_int]
        L.oracle_fdgrad_refmode_f64.argtypes = [ctypes.c_long, dp, dp, ctypes.c_int]
        L.oracle_fdgrad_refmode_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_int]
        L.oracle_fd_complex_step.argtypes = [dp, ctypes.c_double, dp, ctypes.c_int]
        L.oracle_id_gradient_f64.argtypes = [ctypes.c_long, dp, dp, ctypes.c_int]
        L.oracle_vaf_f64.argtypes = [ctypes.c_long, dp, ctypes.c_int]
        L.oracle_id_gradient_split_f64.argtypes = [ctypes.c_long, dp, dp, ctypes.c_int]
        L.oracle_finish_f64.argtypes = [ctypes.c_long, dp, dp, dp]
        L.oracle_rnea_complex_step.argtypes = [dp, ctypes.c_double, dp, ctypes.c_int]
        L.oracle_fd_complex_step_fext.argtypes = [dp, dp, ctypes.c_double, dp, ctypes.c_int]
        L.oracle_fdgrad_fext_f64.argtypes = [ctypes.c_long, dp, dp, ctypes.c_long, dp, ctypes.c_int]
        L.oracle_rnea_fext.argtypes = [dp, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_fd_fext.argtypes = [dp, dp, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_inertias.argtypes = [dp]
        L.oracle_transforms.argtypes = [dp, dp]
        L.oracle_rnea.argtypes = [dp, dp, dp, dp, ctypes.c_int]
        L.oracle_fd.argtypes = [dp, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_pivots.argtypes = [dp, dp]
        L.oracle_op_count.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
        L.oracle_cpu_bench.argtypes = [ctypes.c_int, ctypes.c_long, fp, fp, ctypes.c_long, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double, dp]
        L.oracle_cpu_bench.restype = ctypes.c_long
        L.oracle_set_mutation.argtypes = [ctypes.c_int]
        L.oracle_set_model.argtypes = [dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _c64(a, shape_last):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.shape[-1] == shape_last
    return a


FEXT_PER_SET = 42   # 7 links x spatial force (moment, force) in the link frame


def _fext(fext, n):
    """fext: None, (42,) for every knot (stride 0) or (n, 42) per knot (stride 42)."""
    if fext is None:
        return None, 0
    f = np.ascontiguousarray(fext, dtype=np.float64)
    if f.ndim == 1:
        assert f.shape == (FEXT_PER_SET,)
        return f, 0
    assert f.shape == (n, FEXT_PER_SET)
    return f, FEXT_PER_SET


def fdgrad(x, flags=0, threads=1, fext=None):
    """FP64 full path. x: (n, 21) [q|qd|u] -> (n, 105) [qdd|dqdd/dq|dqdd/dqd].
    fext: external link forces, (42,) for all knots or (n, 42)."""
    x = _c64(np.atleast_2d(x), IN_PER_KNOT)
    out = np.empty((x.shape[0], OUT_PER_KNOT))
    if fext is not None:
        f, stride = _fext(fext, x.shape[0])
        lib().oracle_fdgrad_fext_f64(x.shape[0], _dp(x), _dp(f), stride, _dp(out), flags)
        return out
    if threads > 1:
        lib().oracle_fdgrad_f64_mt(threads, x.shape[0], _dp(x), _dp(out), flags)
    else:
        lib().oracle_fdgrad_f64(x.shape[0], _dp(x), _dp(out), flags)
    return out


def fdgrad_f32(x, flags=0):
    """The same method in FP32 (the CPU baseline's arithmetic)."""
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
    out = np.empty((x.shape[0], OUT_PER_KNOT), dtype=np.float32)
    lib().oracle_fdgrad_f32(x.shape[0], _fp(x), _fp(out), flags)
    return out


def fdgrad_refmode(vol, flags=0):
    """FP64 reference-compatible mode: vol (n, 77) -> (n, 98)."""
    vol = _c64(np.atleast_2d(vol), VOL_PER_KNOT)
    out = np.empty((vol.shape[0], DQDD_PER_KNOT))
    lib().oracle_fdgrad_refmode_f64(vol.shape[0], _dp(vol), _dp(out), flags)
    return out


def fd_complex_step(x, h=1e-20, flags=0, fext=None):
    """[qdd | dqdd/dq | dqdd/dqd] of one knot by complex-step differentiation of FD."""
    x = _c64(x, IN_PER_KNOT)
    out = np.empty(OUT_PER_KNOT)
    if fext is None:
        lib().oracle_fd_complex_step(_dp(x), h, _dp(out), flags)
    else:
        f = _c64(fext, FEXT_PER_SET)
        lib().oracle_fd_complex_step_fext(_dp(x), _dp(f), h, _dp(out), flags)
    return out


FUSED_PER_KNOT = 28    # q, qd, u, qdd
SPLIT_PER_KNOT = 154   # q, qd, u, qdd, v (42), a (42), f (42)


def id_gradient(vol, flags=0):
    """Inverse-dynamics gradient cdu (n, 98) = [dtau/dq | dtau/dqd] from the
    fused layout (n, 28) [q|qd|u|qdd]."""
    vol = _c64(np.atleast_2d(vol), FUSED_PER_KNOT)
    out = np.empty((vol.shape[0], DQDD_PER_KNOT))
    lib().oracle_id_gradient_f64(vol.shape[0], _dp(vol), _dp(out), flags)
    return out


def vaf(vol, flags=0):
    """Split layout (n, 154) with v, a, f filled from the first 28 columns."""
    vol = np.array(np.atleast_2d(vol), dtype=np.float64)
    assert vol.shape[-1] == SPLIT_PER_KNOT
    lib().oracle_vaf_f64(vol.shape[0], _dp(vol), flags)
    return vol


def id_gradient_split(vol, flags=0):
    vol = _c64(np.atleast_2d(vol), SPLIT_PER_KNOT)
    out = np.empty((vol.shape[0], DQDD_PER_KNOT))
    lib().oracle_id_gradient_split_f64(vol.shape[0], _dp(vol), _dp(out), flags)
    return out


def finish(minv, cdu):
    """dqdd = -Minv cdu: minv (n, 49) column-major (upper triangle read), cdu (n, 98)."""
    minv = _c64(np.atleast_2d(minv), NQ * NQ)
    cdu = _c64(np.atleast_2d(cdu), DQDD_PER_KNOT)
    out = np.empty_like(cdu)
    lib().oracle_finish_f64(cdu.shape[0], _dp(minv), _dp(cdu), _dp(out))
    return out


def rnea_complex_step(vol, h=1e-20, flags=0):
    """cdu of one knot (vol: 28 values) by complex-step differentiation of RNEA."""
    vol = _c64(vol, FUSED_PER_KNOT)
    out = np.empty(DQDD_PER_KNOT)
    lib().oracle_rnea_complex_step(_dp(vol), h, _dp(out), flags)
    return out


def inertias():
    out = np.empty(7 * 36)
    lib().oracle_inertias(_dp(out))
    return out.reshape(7, 6, 6).transpose(0, 2, 1)   # [k][row][col]


def transforms(q):
    q = _c64(q, NQ)
    out = np.empty(7 * 36)
    lib().oracle_transforms(_dp(q), _dp(out))
    return out.reshape(7, 6, 6).transpose(0, 2, 1)


def rnea(q, qd, qdd, flags=0, fext=None):
    q, qd, qdd = (_c64(a, NQ) for a in (q, qd, qdd))
    tau = np.empty(NQ)
    if fext is None:
        lib().oracle_rnea(_dp(q), _dp(qd), _dp(qdd), _dp(tau), flags)
    else:
        f = _c64(fext, FEXT_PER_SET)
        lib().oracle_rnea_fext(_dp(q), _dp(qd), _dp(qdd), _dp(f), _dp(tau), flags)
    return tau


def fd(q, qd, u, flags=0, fext=None):
    """Returns (qdd, Minv dense [row][col])."""
    q, qd, u = (_c64(a, NQ) for a in (q, qd, u))
    qdd = np.empty(NQ)
    minv = np.empty(NQ * NQ)
    if fext is None:
        lib().oracle_fd(_dp(q), _dp(qd), _dp(u), _dp(qdd), _dp(minv), flags)
    else:
        f = _c64(fext, FEXT_PER_SET)
        lib().oracle_fd_fext(_dp(q), _dp(qd), _dp(u), _dp(f), _dp(qdd), _dp(minv), flags)
    return qdd, minv.reshape(NQ, NQ).T


def pivots(q):
    q = _c64(q, NQ)
    d = np.empty(NQ)
    lib().oracle_pivots(_dp(q), _dp(d))
    return d


def op_count(mode=0, skip_zeros=False):
    c = (ctypes.c_longlong * 4)()
    lib().oracle_op_count(mode, int(bool(skip_zeros)), c)
    return {"add": c[0], "mul": c[1], "div": c[2], "trig": c[3]}


def set_model(links=None, joints=None):
    """Compute on another 7-joint serial chain from now on (None: the iiwa 14).
    links (7, 10): m, cx, cy, cz, ixx, ixy, ixz, iyy, iyz, izz (inertia about
    the COM, link frame); joints (7, 12): the joint frame's fixed rotation R
    (row-major; frame = R Rz(q)) then its offset in the parent frame."""
    if links is None:
        lib().oracle_set_model(None, None)
        return
    links = np.ascontiguousarray(links, dtype=np.float64).reshape(NQ, 10)
    joints = np.ascontiguousarray(joints, dtype=np.float64).reshape(NQ, 12)
    lib().oracle_set_model(_dp(links), _dp(joints))


class using_model:
    """with using_model(links, joints): ... -- restores the iiwa 14 afterwards."""

    def __init__(self, links, joints):
        self.links, self.joints = links, joints

    def __enter__(self):
        set_model(self.links, self.joints)

    def __exit__(self, *exc):
        set_model(None)


def set_mutation(m):
    """Plant mistake m in the oracle (test of the tests); 0 restores it."""
    lib().oracle_set_mutation(int(m))


def cpu_bench(x, threads, iters, mode=0, flags=0, time_cap_s=10.0):
    """Time the reference's ReusableThreads loop on the FP32 restatement.
    x: (n, 21) for mode 0, (n, 77) for mode 1.  Returns per-call times in us."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    out = np.empty((n, OUT_PER_KNOT if mode == 0 else DQDD_PER_KNOT), dtype=np.float32)
    times = np.empty(int(iters))
    done = lib().oracle_cpu_bench(int(threads), n, _fp(x), _fp(out), int(iters), mode, flags,
                                  float(time_cap_s), _dp(times))
    return times[:done], out
