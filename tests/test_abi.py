"""The C-ABI library loads and exports every function include/rbdgrad.h declares
(no GPU needed: nothing here launches work)."""
import ctypes
import os
import re
import subprocess

import pytest

import common


HEADER = os.path.join(common.ROOT, "include", "rbdgrad.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_the_binding_lists():
    import rbdgrad
    assert declared_functions() == sorted(rbdgrad.EXPORTED)


# librbdgrad.so (the iiwa 14) and the library built from tests/golden/synthetic_arm7.urdf
MODELS = [None, "arm7"]


@pytest.mark.parametrize("model", MODELS)
def test_library_exports_every_declared_symbol(model):
    import rbdgrad
    lib = rbdgrad.lib(model)                  # raises if the library is missing
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", rbdgrad.lib_path(model)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (rbd_\w+)", out))
    assert set(declared_functions()) <= exported
    assert lib.rbd_model_name().decode() == (model or "iiwa14")


@pytest.mark.parametrize("model", MODELS)
def test_library_is_built_for_gfx950(model):
    import rbdgrad
    path = rbdgrad.lib_path(model)
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", path], capture_output=True, text=True)
    if out.returncode != 0:
        # fall back to scanning the offload bundle's target triple
        data = open(path, "rb").read()
        assert b"gfx950" in data
    else:
        assert "gfx950" in out.stdout


def test_status_strings_and_config_default():
    import rbdgrad
    L = rbdgrad.lib()
    assert L.rbd_abi_version() == 3
    for s, text in [(0, "ok"), (1, "invalid argument"), (3, "no HIP device"), (7, "device wait timed out")]:
        assert L.rbd_status_string(s).decode() == text
    cfg = rbdgrad.Config()
    L.rbd_config_default(ctypes.byref(cfg))
    assert (cfg.device, cfg.flags, cfg.precision, cfg.lanes_per_knot, cfg.waves_per_group) == (0, 0, 0, 0, 0)


def test_invalid_arguments_are_reported_without_a_device():
    import rbdgrad
    L = rbdgrad.lib()
    h = ctypes.c_void_p()
    # null output pointer and bad config are rejected before any device call
    assert L.rbd_create(None, None) == 1
    cfg = rbdgrad.Config()
    L.rbd_config_default(ctypes.byref(cfg))
    cfg.precision = 9
    assert L.rbd_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert "precision" in L.rbd_last_error().decode()
    cfg.precision = 0
    cfg.flags = 64
    assert L.rbd_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    # null context on every entry point
    assert L.rbd_fdgrad_traj_f32(None, 1, None, None, None) == 1
    assert L.rbd_fdgrad_traj_soa_f32(None, 1, None, 1, None, 1, None) == 1
    assert L.rbd_dynamics_gradient_kernel_f32(None, 1, None, None, None, None) == 1
    assert L.rbd_fdgrad_traj_host_f32(None, 1, None, None) == 1
    assert L.rbd_context_status(None) == 1
    assert L.rbd_context_clear_status(None) == 1
    assert L.rbd_set_handoff_spins(None, 0) == 1
    assert L.rbd_mapping_table(None, 0, None, None, None, 0) == -1
    cfg.flags = 0
    cfg.lanes_per_knot = 5
    assert L.rbd_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert "lanes_per_knot" in L.rbd_last_error().decode()
    cfg.lanes_per_knot = 7
    cfg.waves_per_group = 3
    assert L.rbd_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert "waves_per_group" in L.rbd_last_error().decode()
    cfg.waves_per_group = 0
    cfg.reserved[0] = 1
    assert L.rbd_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert "reserved" in L.rbd_last_error().decode()
    assert L.rbd_destroy(None) == 0
