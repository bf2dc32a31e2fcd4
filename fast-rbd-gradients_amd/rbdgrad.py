"""Python binding of librbdgrad.so (include/rbdgrad.h) plus the few HIP runtime
calls the tests and bench.py need (device buffers, streams, events).

The product is the C ABI; this module only forwards to it.  There is no CPU
fallback: importing the binding fails loudly when the HIP library is missing,
and every call raises RbdError when the library reports an error.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbdgrad.so")


def lib_path(model=None):
    """librbdgrad.so (the iiwa 14), or librbdgrad_MODEL.so built from a URDF by
    `make -C fast-rbd-gradients_amd MODEL=... URDF=...`."""
    return LIB_PATH if not model else os.path.join(_HERE, f"librbdgrad_{model}.so")

NQ = 7
IN_PER_KNOT = 21
OUT_PER_KNOT = 105
VOL_PER_KNOT = 77
DQDD_PER_KNOT = 98

FLAG_VEL_DAMPING = 1
FLAG_NO_GRAVITY = 2

VEL_DAMPING = 1        # RBD_FLAG_VEL_DAMPING
NO_GRAVITY = 2         # RBD_FLAG_NO_GRAVITY (the reference's MPC_MODE)
PREC_FP32 = 0
PREC_FP64_PIVOT = 1
PREC_FP64_COLUMNS = 2
PREC_FP64 = 3
PREC_FP32_HW_TRIG = 4
PRECISIONS = {"fp32": PREC_FP32, "fp64_pivot": PREC_FP64_PIVOT, "fp64_columns": PREC_FP64_COLUMNS,
              "fp64": PREC_FP64, "fp32_hw_trig": PREC_FP32_HW_TRIG}

# every function the header declares, for the export check
EXPORTED = [
    "rbd_config_default", "rbd_create", "rbd_destroy", "rbd_status_string", "rbd_last_error",
    "rbd_abi_version", "rbd_fdgrad_traj_f32", "rbd_fdgrad_traj_soa_f32",
    "rbd_dynamics_gradient_kernel_f32", "rbd_fdgrad_traj_host_f32", "rbd_kernel_name",
    "rbd_autotune", "rbd_group_threshold", "rbd_set_group_threshold", "rbd_two_wave_threshold",
    "rbd_set_two_wave_threshold",
    "rbd_autotune_batch", "rbd_batch_group_threshold", "rbd_set_batch_group_threshold",
    "rbd_batch_create", "rbd_batch_destroy", "rbd_fdgrad_batched_f32", "rbd_batch_set_grid",
    "rbd_batch_last_lanes",
    "rbd_server_config_default", "rbd_server_create", "rbd_server_destroy", "rbd_server_submit",
    "rbd_server_poll", "rbd_server_wait", "rbd_server_get_stats", "rbd_server_reset_stats",
    "rbd_server_run_poisson", "rbd_server_get_trace", "rbd_fdgrad_minv_traj_f32", "rbd_forward_dynamics_f32",
    "rbd_fdgrad_traj_fext_f32", "rbd_forward_dynamics_fext_f32",
    "rbd_dynamics_gradient_kernel_fused_f32", "rbd_dynamics_gradient_kernel_split_f32",
    "rbd_gradient_setup_split_f32", "rbd_gradient_finish_f32", "rbd_model_name", "rbd_fdgrad_wafr_f32",
    "rbd_context_status", "rbd_context_clear_status", "rbd_set_handoff_spins", "rbd_mapping_table",
    "rbd_fdgrad_batched_noalloc_f32",
]

# the mappings the autotuners compare, in the order of their table rows (lanes per knot)
AUTOTUNE_LANES = (1, 16, 32, 4, 7, 8)
# the launch shapes the autotuners time, in their table order: (lanes per
# knot, waves per workgroup); the 4-, 7- and 8-lane groups also run in
# workgroups of 2 and 4 waves
AUTOTUNE_SHAPES = tuple((l, 1) for l in AUTOTUNE_LANES) + tuple((l, w) for w in (2, 4) for l in (4, 7, 8))


def shape_key(lanes, wpg=1):
    """Key of a launch shape in the autotune timings: "7" for 7 lanes per knot
    in workgroups of one wave, "7x4" in workgroups of four."""
    return str(lanes) if wpg == 1 else f"{lanes}x{wpg}"

ERR_DEVICE_TIMEOUT = 7   # RBD_ERR_DEVICE_TIMEOUT


class RbdError(RuntimeError):
    pass


class HipError(RuntimeError):
    pass


class Traj(ctypes.Structure):
    _fields_ = [("d_in", ctypes.c_void_p), ("d_out", ctypes.c_void_p), ("n_knots", ctypes.c_int64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("lanes_per_knot", ctypes.c_int), ("memory_budget_bytes", ctypes.c_int64),
                ("max_batch_trajectories", ctypes.c_int64), ("max_batch_knots", ctypes.c_int64),
                ("batch_window_us", ctypes.c_double), ("use_graph", ctypes.c_int), ("zero_copy", ctypes.c_int),
                ("poll_dispatch", ctypes.c_int), ("zero_copy_max_knots", ctypes.c_int)]


class BatchRecord(ctypes.Structure):
    _fields_ = [("oldest_submit_us", ctypes.c_double), ("admit_us", ctypes.c_double),
                ("enqueued_us", ctypes.c_double), ("done_us", ctypes.c_double), ("delivered_us", ctypes.c_double),
                ("requests", ctypes.c_int64), ("knots", ctypes.c_int64), ("slot", ctypes.c_int),
                ("reserved", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests_completed", ctypes.c_int64), ("knots_completed", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("max_inflight_batches", ctypes.c_int64),
                ("max_batch_knots", ctypes.c_int64), ("max_budget_bytes_used", ctypes.c_int64),
                ("queued", ctypes.c_int64), ("p50_latency_us", ctypes.c_double), ("p99_latency_us", ctypes.c_double),
                ("mean_latency_us", ctypes.c_double), ("max_latency_us", ctypes.c_double),
                ("knots_per_s", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("flags", ctypes.c_uint), ("precision", ctypes.c_int),
                ("lanes_per_knot", ctypes.c_int), ("waves_per_group", ctypes.c_int),
                ("reserved", ctypes.c_int * 3)]


_libs = {}
_hip = None


def lib(model=None):
    """Load librbdgrad.so, or librbdgrad_MODEL.so (raises if it is not built)."""
    key = model or ""
    if key not in _libs:
        path = lib_path(model)
        if not os.path.exists(path):
            raise RbdError(f"{path} is not built: run __graft_entry__.build() or "
                           f"`make -C fast-rbd-gradients_amd" + (f" MODEL={model} URDF=...`" if model else "`"))
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.rbd_config_default.argtypes = [ctypes.POINTER(Config)]
        L.rbd_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
        L.rbd_destroy.argtypes = [vp]
        L.rbd_status_string.argtypes = [ctypes.c_int]
        L.rbd_status_string.restype = ctypes.c_char_p
        L.rbd_last_error.restype = ctypes.c_char_p
        L.rbd_fdgrad_traj_f32.argtypes = [vp, i64, vp, vp, vp]
        L.rbd_fdgrad_traj_soa_f32.argtypes = [vp, i64, vp, i64, vp, i64, vp]
        L.rbd_dynamics_gradient_kernel_f32.argtypes = [vp, i64, vp, vp, vp, vp]
        L.rbd_fdgrad_traj_host_f32.argtypes = [vp, i64, vp, vp]
        L.rbd_fdgrad_minv_traj_f32.argtypes = [vp, i64, vp, vp, vp, vp]
        L.rbd_fdgrad_wafr_f32.argtypes = [vp, i64, vp, vp, vp, vp]
        L.rbd_forward_dynamics_f32.argtypes = [vp, i64, vp, vp, vp, vp]
        L.rbd_fdgrad_traj_fext_f32.argtypes = [vp, i64, vp, vp, i64, vp, vp]
        L.rbd_dynamics_gradient_kernel_fused_f32.argtypes = [vp, i64, vp, vp, vp, vp]
        L.rbd_dynamics_gradient_kernel_split_f32.argtypes = [vp, i64, vp, vp, vp, vp]
        L.rbd_gradient_setup_split_f32.argtypes = [vp, i64, vp, vp]
        L.rbd_gradient_finish_f32.argtypes = [vp, i64, vp, vp, vp, vp]
        L.rbd_forward_dynamics_fext_f32.argtypes = [vp, i64, vp, vp, i64, vp, vp, vp]
        L.rbd_kernel_name.argtypes = [vp, ctypes.c_int, i64]
        L.rbd_kernel_name.restype = ctypes.c_char_p
        L.rbd_autotune.argtypes = [vp, vp, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        L.rbd_group_threshold.argtypes = [vp]
        L.rbd_group_threshold.restype = i64
        L.rbd_set_group_threshold.argtypes = [vp, i64]
        L.rbd_two_wave_threshold.argtypes = [vp]
        L.rbd_two_wave_threshold.restype = i64
        L.rbd_set_two_wave_threshold.argtypes = [vp, i64]
        L.rbd_autotune_batch.argtypes = [vp, vp, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        L.rbd_batch_group_threshold.argtypes = [vp]
        L.rbd_batch_group_threshold.restype = i64
        L.rbd_set_batch_group_threshold.argtypes = [vp, i64]
        L.rbd_batch_create.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.rbd_batch_destroy.argtypes = [vp]
        L.rbd_fdgrad_batched_f32.argtypes = [vp, ctypes.POINTER(Traj), ctypes.c_int, ctypes.c_int, vp]
        L.rbd_batch_set_grid.argtypes = [vp, ctypes.c_int, ctypes.c_uint]
        L.rbd_fdgrad_batched_noalloc_f32.argtypes = [vp, ctypes.POINTER(Traj), ctypes.c_int, vp]
        L.rbd_batch_last_lanes.argtypes = [vp]
        L.rbd_server_config_default.argtypes = [ctypes.POINTER(ServerConfig)]
        L.rbd_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.rbd_server_destroy.argtypes = [vp]
        L.rbd_server_submit.argtypes = [vp, vp, vp, i64, ctypes.POINTER(ctypes.c_uint64)]
        L.rbd_server_poll.argtypes = [vp, ctypes.c_uint64]
        L.rbd_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_double]
        L.rbd_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.rbd_server_reset_stats.argtypes = [vp]
        L.rbd_server_get_trace.argtypes = [vp, ctypes.POINTER(BatchRecord), i64, ctypes.POINTER(i64)]
        L.rbd_server_run_poisson.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(i64), ctypes.c_int,
                                             vp, i64, ctypes.c_uint64, ctypes.POINTER(ServerStats)]
        L.rbd_model_name.restype = ctypes.c_char_p
        L.rbd_context_status.argtypes = [vp]
        L.rbd_context_clear_status.argtypes = [vp]
        L.rbd_set_handoff_spins.argtypes = [vp, ctypes.c_int]
        L.rbd_mapping_table.argtypes = [vp, ctypes.c_int, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        _libs[key] = L
    return _libs[key]


def hip():
    """The HIP runtime librbdgrad.so is linked against (same process instance)."""
    global _hip
    if _hip is None:
        lib()
        H = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        H.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        H.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        H.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipDeviceSynchronize.argtypes = []
        H.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        H.hipGetErrorString.restype = ctypes.c_char_p
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        H.hipSetDevice.argtypes = [ctypes.c_int]
        H.hipHostMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_uint]
        H.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        H.hipHostFree.argtypes = [vp]
        _hip = H
    return _hip


def _hipcheck(err, what):
    if err != 0:
        raise HipError(f"{what}: {hip().hipGetErrorString(err).decode()} ({err})")


def _check(status, what, L=None):
    if status != 0:
        L = L or lib()
        raise RbdError(f"{what}: {L.rbd_status_string(status).decode()} -- {L.rbd_last_error().decode()}")


def device_count():
    n = ctypes.c_int(0)
    err = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if err == 0 else 0


def mem_info():
    """(free, total) bytes of the current device (hipMemGetInfo)."""
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    _hipcheck(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """hipMalloc'd memory; upload/download numpy arrays."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _hipcheck(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if stream is None:
            _hipcheck(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        else:
            _hipcheck(hip().hipMemcpyAsync(self.ptr, a.ctypes.data, a.nbytes, H2D, stream.handle),
                      "hipMemcpyAsync H2D")

    def download(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _hipcheck(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def fill_zero(self):
        _hipcheck(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def handle(self):
        return self.ptr


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _hipcheck(hip().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        _hipcheck(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle is not None:
            hip().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _hipcheck(hip().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        _hipcheck(hip().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        _hipcheck(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _hipcheck(hip().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle is not None:
            hip().hipEventDestroy(self.handle)
            self.handle = None


def _sh(stream):
    return stream.handle if stream is not None else None


class _Bound:
    """An object of one model library (librbdgrad.so or librbdgrad_MODEL.so)."""
    _L = None

    def _check(self, status, what):
        _check(status, what, self._L)


class Context(_Bound):
    """rbd_context: device, model flags, precision mode.  model=None is the
    KUKA iiwa 14; model="name" loads librbdgrad_name.so, built from a URDF."""

    def __init__(self, device=0, flags=0, precision=PREC_FP32, lanes_per_knot=0, model=None, waves_per_group=0):
        L = self._L = lib(model)
        self.model = L.rbd_model_name().decode()
        cfg = Config()
        L.rbd_config_default(ctypes.byref(cfg))
        cfg.device = device
        cfg.flags = flags
        cfg.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        cfg.lanes_per_knot = int(lanes_per_knot)
        cfg.waves_per_group = int(waves_per_group)
        self._h = ctypes.c_void_p()
        self._check(L.rbd_create(ctypes.byref(cfg), ctypes.byref(self._h)), "rbd_create")
        self.flags = flags
        self.precision = cfg.precision

    def close(self):
        if self._h is not None and self._h.value:
            self._L.rbd_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # --- device-pointer entry points (the drop-in calls) ---
    def fdgrad(self, n, d_in, d_out, stream=None):
        self._check(self._L.rbd_fdgrad_traj_f32(self._h, int(n), d_in.handle, d_out.handle, _sh(stream)),
               "rbd_fdgrad_traj_f32")

    def fdgrad_soa(self, n, d_in, ld_in, d_out, ld_out, stream=None):
        self._check(self._L.rbd_fdgrad_traj_soa_f32(self._h, int(n), d_in.handle, int(ld_in), d_out.handle, int(ld_out),
                                             _sh(stream)), "rbd_fdgrad_traj_soa_f32")

    def dynamics_gradient_kernel(self, n, d_dqdd, d_vol, d_const=None, stream=None):
        self._check(self._L.rbd_dynamics_gradient_kernel_f32(self._h, int(n), d_dqdd.handle, d_vol.handle,
                                                      d_const.handle if d_const is not None else None,
                                                      _sh(stream)), "rbd_dynamics_gradient_kernel_f32")

    def fdgrad_minv(self, n, d_in, d_out, d_minv, stream=None):
        self._check(self._L.rbd_fdgrad_minv_traj_f32(self._h, int(n), d_in.handle, d_out.handle, d_minv.handle,
                                                     _sh(stream)), "rbd_fdgrad_minv_traj_f32")

    def fdgrad_wafr(self, n, d_in, d_qdd, d_dqdd, stream=None):
        self._check(self._L.rbd_fdgrad_wafr_f32(self._h, int(n), d_in.handle, d_qdd.handle if d_qdd else None,
                                                d_dqdd.handle, _sh(stream)), "rbd_fdgrad_wafr_f32")

    def fdgrad_fext(self, n, d_in, d_fext, fext_stride, d_out, stream=None):
        self._check(self._L.rbd_fdgrad_traj_fext_f32(self._h, int(n), d_in.handle, d_fext.handle, int(fext_stride),
                                                     d_out.handle, _sh(stream)), "rbd_fdgrad_traj_fext_f32")

    def dynamics_gradient_kernel_fused(self, n, d_cdu, d_vol, stream=None):
        self._check(self._L.rbd_dynamics_gradient_kernel_fused_f32(self._h, int(n), d_cdu.handle, d_vol.handle, None,
                                                                   _sh(stream)),
                    "rbd_dynamics_gradient_kernel_fused_f32")

    def dynamics_gradient_kernel_split(self, n, d_cdu, d_vol, stream=None):
        self._check(self._L.rbd_dynamics_gradient_kernel_split_f32(self._h, int(n), d_cdu.handle, d_vol.handle, None,
                                                                   _sh(stream)),
                    "rbd_dynamics_gradient_kernel_split_f32")

    def gradient_setup_split_dev(self, n, d_vol, stream=None):
        self._check(self._L.rbd_gradient_setup_split_f32(self._h, int(n), d_vol.handle, _sh(stream)),
                    "rbd_gradient_setup_split_f32")

    def gradient_finish_dev(self, n, d_dqdd, d_cdu, d_minv, stream=None):
        self._check(self._L.rbd_gradient_finish_f32(self._h, int(n), d_dqdd.handle, d_cdu.handle, d_minv.handle,
                                                    _sh(stream)), "rbd_gradient_finish_f32")

    def forward_dynamics_dev(self, n, d_in, d_qdd, d_minv, stream=None):
        self._check(self._L.rbd_forward_dynamics_f32(self._h, int(n), d_in.handle, d_qdd.handle, d_minv.handle,
                                                     _sh(stream)), "rbd_forward_dynamics_f32")

    def status(self):
        """rbd_context_status: 0, or ERR_DEVICE_TIMEOUT once a two-wave kernel
        of this context gave up a handoff wait (synchronise first)."""
        return self._L.rbd_context_status(self._h)

    def clear_status(self):
        self._check(self._L.rbd_context_clear_status(self._h), "rbd_context_clear_status")

    def set_handoff_spins(self, spins):
        """Bound of the two-wave handoff waits (< 0: every wait gives up at once, a test hook)."""
        self._check(self._L.rbd_set_handoff_spins(self._h, int(spins)), "rbd_set_handoff_spins")

    def batched_noalloc(self, trajs, stream=None):
        """rbd_fdgrad_batched_noalloc_f32: up to 64 trajectories (a TrajList or
        [(d_in, d_out, n_knots), ...]) with nothing allocated by the library."""
        if not isinstance(trajs, TrajList):
            trajs = TrajList(trajs)
        self._check(self._L.rbd_fdgrad_batched_noalloc_f32(self._h, trajs.arr, trajs.n, _sh(stream)),
                    "rbd_fdgrad_batched_noalloc_f32")

    def kernel_name(self, n_knots, reference_mode=False):
        return self._L.rbd_kernel_name(self._h, int(bool(reference_mode)), int(n_knots)).decode()

    def _timings(self, fn, what, stream):
        thr = ctypes.c_int64()
        row = 1 + len(AUTOTUNE_SHAPES)
        tab = (ctypes.c_float * (16 * row))()
        self._check(fn(self._h, _sh(stream), ctypes.byref(thr), tab, 16 * row), what)
        table = {int(tab[i]): {shape_key(*sh): tab[i + 1 + j] for j, sh in enumerate(AUTOTUNE_SHAPES)}
                 for i in range(0, 16 * row, row) if tab[i] > 0}
        return thr.value, table

    def autotune(self, stream=None):
        """Time the twelve launch shapes (AUTOTUNE_SHAPES) on the device and
        install the context's mapping table; returns the 16-lane crossover and
        the timings {n: {shape_key: us}} (the two-wave limit is then
        self.two_wave_threshold, the installed table self.mapping_table())."""
        return self._timings(self._L.rbd_autotune, "rbd_autotune", stream)

    def mapping_table(self, batched=False):
        """The installed mapping table: [(largest knots, lanes per knot, waves
        per workgroup), ...] ascending."""
        n = (ctypes.c_int64 * 16)()
        lanes = (ctypes.c_int * 16)()
        wpg = (ctypes.c_int * 16)()
        c = self._L.rbd_mapping_table(self._h, int(bool(batched)), n, lanes, wpg, 16)
        return [(int(n[i]), int(lanes[i]), int(wpg[i])) for i in range(max(c, 0))]

    @property
    def two_wave_threshold(self):
        return self._L.rbd_two_wave_threshold(self._h)

    @two_wave_threshold.setter
    def two_wave_threshold(self, n):
        self._check(self._L.rbd_set_two_wave_threshold(self._h, int(n)), "rbd_set_two_wave_threshold")

    def autotune_batch(self, stream=None):
        """Time the twelve launch shapes of the batched path and install the
        batch mapping table; returns the 16-lane crossover (total knots) and
        the timings {total knots: {shape_key: us}}."""
        return self._timings(self._L.rbd_autotune_batch, "rbd_autotune_batch", stream)

    @property
    def batch_group_threshold(self):
        return self._L.rbd_batch_group_threshold(self._h)

    @batch_group_threshold.setter
    def batch_group_threshold(self, n):
        self._check(self._L.rbd_set_batch_group_threshold(self._h, int(n)), "rbd_set_batch_group_threshold")

    @property
    def group_threshold(self):
        return self._L.rbd_group_threshold(self._h)

    @group_threshold.setter
    def group_threshold(self, n):
        self._check(self._L.rbd_set_group_threshold(self._h, int(n)), "rbd_set_group_threshold")

    # --- host-array conveniences (upload, run, download) ---
    def fdgrad_host(self, x):
        """x: (n, 21) float32 [q|qd|u] -> (n, 105) float32, via rbd_fdgrad_traj_host_f32."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        assert x.shape[1] == IN_PER_KNOT
        out = np.empty((x.shape[0], OUT_PER_KNOT), dtype=np.float32)
        self._check(self._L.rbd_fdgrad_traj_host_f32(self._h, x.shape[0], x.ctypes.data, out.ctypes.data),
               "rbd_fdgrad_traj_host_f32")
        return out

    def run_aos(self, x):
        """Upload x (n, 21), run rbd_fdgrad_traj_f32 on the default stream, download (n, 105)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        n = x.shape[0]
        d_in = DeviceBuffer.from_array(x)
        d_out = DeviceBuffer(n * OUT_PER_KNOT * 4)
        self.fdgrad(n, d_in, d_out)
        return d_out.download((n, OUT_PER_KNOT))

    def run_soa(self, x, pad=0):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        n = x.shape[0]
        ld = n + pad
        xs = np.zeros((IN_PER_KNOT, ld), dtype=np.float32)
        xs[:, :n] = x.T
        d_in = DeviceBuffer.from_array(xs)
        d_out = DeviceBuffer(OUT_PER_KNOT * ld * 4)
        self.fdgrad_soa(n, d_in, ld, d_out, ld)
        return d_out.download((OUT_PER_KNOT, ld))[:, :n].T.copy()

    def run_with_minv(self, x):
        """Full path plus dqdd/du = Minv: returns ((n, 105), (n, 7, 7) Minv [knot][row][col])."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        n = x.shape[0]
        d_in = DeviceBuffer.from_array(x)
        d_out = DeviceBuffer(n * OUT_PER_KNOT * 4)
        d_m = DeviceBuffer(n * 49 * 4)
        self.fdgrad_minv(n, d_in, d_out, d_m)
        return d_out.download((n, OUT_PER_KNOT)), d_m.download((n, 7, 7)).transpose(0, 2, 1).copy()

    def run_wafr(self, x, with_qdd=True):
        """Full path in the WAFR baseline's gradient layout: returns (qdd (n, 7)
        or None, dqdd (n, 147) = [dqdd/dq | dqdd/dqd | Minv] per knot, each
        7x7 column-major)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        n = x.shape[0]
        d_in = DeviceBuffer.from_array(x)
        d_q = DeviceBuffer(n * 7 * 4) if with_qdd else None
        d_d = DeviceBuffer(n * 147 * 4)
        self.fdgrad_wafr(n, d_in, d_q, d_d)
        return (d_q.download((n, 7)) if d_q else None), d_d.download((n, 147))

    def forward_dynamics(self, x, fext=None):
        """qdd (n, 7) and Minv (n, 7, 7) [knot][row][col] from x (n, 21);
        fext: external link forces, (42,) for every knot or (n, 42)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        n = x.shape[0]
        d_in = DeviceBuffer.from_array(x)
        d_q = DeviceBuffer(n * 7 * 4)
        d_m = DeviceBuffer(n * 49 * 4)
        if fext is None:
            self.forward_dynamics_dev(n, d_in, d_q, d_m)
        else:
            d_f, stride = _fext_buffer(fext, n)
            self._check(self._L.rbd_forward_dynamics_fext_f32(self._h, n, d_in.handle, d_f.handle, stride, d_q.handle,
                                                       d_m.handle, None), "rbd_forward_dynamics_fext_f32")
        return d_q.download((n, 7)), d_m.download((n, 7, 7)).transpose(0, 2, 1).copy()

    def run_fext(self, x, fext):
        """Full path with external link forces: x (n, 21), fext (42,) or (n, 42) -> (n, 105)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        n = x.shape[0]
        d_in = DeviceBuffer.from_array(x)
        d_out = DeviceBuffer(n * OUT_PER_KNOT * 4)
        d_f, stride = _fext_buffer(fext, n)
        self.fdgrad_fext(n, d_in, d_f, stride, d_out)
        return d_out.download((n, OUT_PER_KNOT))

    def id_gradient(self, vol):
        """Inverse-dynamics gradient (n, 98) from the fused layout (n, 28) [q|qd|u|qdd]."""
        vol = np.ascontiguousarray(np.atleast_2d(vol), dtype=np.float32)
        n = vol.shape[0]
        d_vol = DeviceBuffer.from_array(vol)
        d_cdu = DeviceBuffer(n * DQDD_PER_KNOT * 4)
        self.dynamics_gradient_kernel_fused(n, d_cdu, d_vol)
        return d_cdu.download((n, DQDD_PER_KNOT))

    def id_gradient_split(self, vol):
        """Inverse-dynamics gradient (n, 98) from the split layout (n, 154)."""
        vol = np.ascontiguousarray(np.atleast_2d(vol), dtype=np.float32)
        n = vol.shape[0]
        d_vol = DeviceBuffer.from_array(vol)
        d_cdu = DeviceBuffer(n * DQDD_PER_KNOT * 4)
        self.dynamics_gradient_kernel_split(n, d_cdu, d_vol)
        return d_cdu.download((n, DQDD_PER_KNOT))

    def gradient_setup_split(self, vol28):
        """Split layout (n, 154) with v, a, f computed on the device from (n, 28)."""
        vol28 = np.ascontiguousarray(np.atleast_2d(vol28), dtype=np.float32)
        n = vol28.shape[0]
        vol = np.zeros((n, 154), np.float32)
        vol[:, :28] = vol28
        d_vol = DeviceBuffer.from_array(vol)
        self.gradient_setup_split_dev(n, d_vol)
        return d_vol.download((n, 154))

    def gradient_finish(self, minv, cdu):
        """-Minv cdu: minv (n, 49) column-major (upper triangle read), cdu (n, 98) -> (n, 98)."""
        minv = np.ascontiguousarray(np.atleast_2d(minv), dtype=np.float32)
        cdu = np.ascontiguousarray(np.atleast_2d(cdu), dtype=np.float32)
        n = cdu.shape[0]
        d_m, d_c = DeviceBuffer.from_array(minv), DeviceBuffer.from_array(cdu)
        d_out = DeviceBuffer(n * DQDD_PER_KNOT * 4)
        self.gradient_finish_dev(n, d_out, d_c, d_m)
        return d_out.download((n, DQDD_PER_KNOT))

    def run_refmode(self, vol):
        vol = np.ascontiguousarray(np.atleast_2d(vol), dtype=np.float32)
        n = vol.shape[0]
        d_vol = DeviceBuffer.from_array(vol)
        d_out = DeviceBuffer(n * DQDD_PER_KNOT * 4)
        self.dynamics_gradient_kernel(n, d_out, d_vol)
        return d_out.download((n, DQDD_PER_KNOT))


FEXT_PER_SET = 42


def _fext_buffer(fext, n):
    """Device copy of external forces and its stride (0: one set for all knots)."""
    f = np.ascontiguousarray(fext, dtype=np.float32)
    if f.ndim == 1:
        if f.shape != (FEXT_PER_SET,):
            raise ValueError("fext must have 42 values (7 links x 6)")
        return DeviceBuffer.from_array(f), 0
    if f.shape != (n, FEXT_PER_SET):
        raise ValueError("per-knot fext must have shape (n, 42)")
    return DeviceBuffer.from_array(f), FEXT_PER_SET


class TrajList:
    """An rbd_traj array built once from [(d_in, d_out, n_knots), ...]; keeps the buffers alive."""

    def __init__(self, trajs):
        self.n = len(trajs)
        self.arr = (Traj * max(1, self.n))()
        for i, (di, do, n) in enumerate(trajs):
            self.arr[i] = Traj(di.handle.value if di is not None else None,
                               do.handle.value if do is not None else None, int(n))
        self._keep = list(trajs)


class Batch(_Bound):
    """rbd_batch: runs many trajectories of uneven length with one persistent-kernel launch."""

    def __init__(self, ctx, max_trajectories):
        self.ctx = ctx
        self._L = ctx._L
        self._h = ctypes.c_void_p()
        self._check(self._L.rbd_batch_create(ctx.handle, int(max_trajectories), ctypes.byref(self._h)),
               "rbd_batch_create")

    def run(self, trajs, use_graph=False, stream=None):
        """trajs: list of (d_in DeviceBuffer, d_out DeviceBuffer, n_knots), or a
        TrajList built from one (no per-call Python work on the list)."""
        if not isinstance(trajs, TrajList):
            trajs = TrajList(trajs)
        self._check(self._L.rbd_fdgrad_batched_f32(self._h, trajs.arr, trajs.n, int(bool(use_graph)), _sh(stream)),
                    "rbd_fdgrad_batched_f32")

    def set_grid(self, lanes_per_knot, waves):
        self._check(self._L.rbd_batch_set_grid(self._h, int(lanes_per_knot), int(waves)), "rbd_batch_set_grid")

    @property
    def last_lanes(self):
        return self._L.rbd_batch_last_lanes(self._h)

    def close(self):
        if self._h is not None and self._h.value:
            self._L.rbd_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server(_Bound):
    """rbd_server: continuous batching of gradient requests (one trajectory each)."""

    def __init__(self, ctx, **cfg):
        self._L = ctx._L
        c = ServerConfig()
        self._L.rbd_server_config_default(ctypes.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.ctx = ctx
        self.config = c
        self._h = ctypes.c_void_p()
        self._keep = {}                      # request id -> (input, output) arrays, alive until collected
        self._check(self._L.rbd_server_create(ctx.handle, ctypes.byref(c), ctypes.byref(self._h)), "rbd_server_create")

    def submit(self, x):
        """Queue one trajectory x (n, 21); returns the request id."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        out = np.empty((x.shape[0], OUT_PER_KNOT), dtype=np.float32)
        rid = ctypes.c_uint64()
        self._check(self._L.rbd_server_submit(self._h, x.ctypes.data, out.ctypes.data, x.shape[0], ctypes.byref(rid)),
               "rbd_server_submit")
        self._keep[rid.value] = (x, out)
        return rid.value

    def poll(self, rid):
        return self._L.rbd_server_poll(self._h, rid)

    def wait(self, rid, timeout_s=60.0):
        """Block until done; returns the (n, 105) result."""
        self._check(self._L.rbd_server_wait(self._h, rid, timeout_s), "rbd_server_wait")
        return self._keep.pop(rid)[1]

    def stats(self):
        st = ServerStats()
        self._check(self._L.rbd_server_get_stats(self._h, ctypes.byref(st)), "rbd_server_get_stats")
        return st.as_dict()

    def trace(self, max_records=65536):
        """The most recent batch records, oldest first, as a structured array."""
        buf = (BatchRecord * int(max_records))()
        n = ctypes.c_int64()
        self._check(self._L.rbd_server_get_trace(self._h, buf, int(max_records), ctypes.byref(n)), "rbd_server_get_trace")
        names = [f for f, _ in BatchRecord._fields_ if f != "reserved"]
        return {f: np.array([getattr(buf[i], f) for i in range(n.value)]) for f in names}

    def reset_stats(self):
        self._check(self._L.rbd_server_reset_stats(self._h), "rbd_server_reset_stats")

    def run_poisson(self, rate, duration_s, lengths, pool, seed=0):
        lens = (ctypes.c_int64 * len(lengths))(*[int(v) for v in lengths])
        pool = np.ascontiguousarray(pool, dtype=np.float32)
        st = ServerStats()
        self._check(self._L.rbd_server_run_poisson(self._h, float(rate), float(duration_s), lens, len(lengths),
                                            pool.ctypes.data, pool.shape[0], int(seed), ctypes.byref(st)),
               "rbd_server_run_poisson")
        return st.as_dict()

    def close(self):
        if self._h is not None and self._h.value:
            self._L.rbd_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
