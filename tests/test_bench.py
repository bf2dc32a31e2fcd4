"""bench.py without a GPU: every name it calls resolves, and it refuses more
than one GPU with a message (CPU only)."""
import ast
import builtins
import os
import subprocess
import sys

import common

BENCH = os.path.join(common.ROOT, "bench.py")


def test_every_called_name_is_defined():
    """A workload function lost from bench.py shows up only at run time, minutes
    into a GPU run; here it fails at once."""
    tree = ast.parse(open(BENCH).read())
    defined = set(dir(builtins)) | {"__file__", "__name__", "__doc__"}
    for n in ast.walk(tree):
        if isinstance(n, (ast.FunctionDef, ast.ClassDef)):
            defined.add(n.name)
        if isinstance(n, ast.arg):
            defined.add(n.arg)
        elif isinstance(n, ast.Name) and isinstance(n.ctx, ast.Store):
            defined.add(n.id)
        elif isinstance(n, (ast.Import, ast.ImportFrom)):
            defined.update((a.asname or a.name).split(".")[0] for a in n.names)
    loaded = {n.id for n in ast.walk(tree) if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load)}
    assert sorted(loaded - defined) == []


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr


def test_oracle_check_kinds_of_the_other_entry_points():
    """The CPU leg's checks of the other entry points' outputs: exact FP64
    oracle outputs (rounded to FP32) pass in every layout the bench records,
    and outputs with a mistake planted on the CPU fail: the external forces
    dropped, a Minv column moved, qdd off by a sign."""
    import numpy as np
    sys.path.insert(0, common.ROOT)
    import bench
    from oracle import pyoracle
    x = bench.knots(64, seed=3, near_singular_frac=0.1)
    x64 = x.astype(np.float64)
    fd = [pyoracle.fd(k[:7], k[7:14], k[14:21]) for k in x64]
    qdd = np.array([q for q, _ in fd])
    minv = np.array([m.T.reshape(49) for _, m in fd])        # column-major, as the library writes it
    full = pyoracle.fdgrad(x64)
    fext = np.random.default_rng(1).uniform(-20.0, 20.0, (64, 42)).astype(np.float32)
    with_fext = pyoracle.fdgrad(x64, fext=fext.astype(np.float64))
    moved = np.roll(minv.reshape(-1, 7, 7), 1, axis=1).reshape(-1, 49)
    vol28 = np.concatenate([x, qdd.astype(np.float32)], axis=1)
    cdu = pyoracle.id_gradient(vol28.astype(np.float64))
    split = pyoracle.vaf(np.concatenate([vol28.astype(np.float64), np.zeros((64, 126))], axis=1))
    cases = {
        "fdgrad_minv": (np.concatenate([full, minv], 1), {"kind": "fdgrad_minv"}, True),
        "fd": (np.concatenate([qdd, minv], 1), {"kind": "fd"}, True),
        "fext": (with_fext, {"fext": fext}, True),
        "fext_dropped": (full, {"fext": fext}, False),
        "minv_moved": (np.concatenate([full, moved], 1), {"kind": "fdgrad_minv"}, False),
        "qdd_sign": (np.concatenate([-qdd, minv], 1), {"kind": "fd"}, False),
        "dqdd": (full[:, 7:], {"kind": "dqdd"}, True),
        "dqdd_shifted": (full[:, 6:104], {"kind": "dqdd"}, False),
        "cdu": (cdu, {"kind": "cdu", "x": vol28}, True),
        "cdu_sign": (-cdu, {"kind": "cdu", "x": vol28}, False),
        "vaf": (split, {"kind": "vaf", "x": vol28}, True),
        "vaf_f_dropped": (np.concatenate([split[:, :112], np.zeros((64, 42))], 1), {"kind": "vaf", "x": vol28}, False),
    }
    checks, res = bench.OracleChecks(), {}
    for name, (out, kw, _) in cases.items():
        res[name] = {}
        kw = dict(kw)
        checks.add(res[name], "accuracy", out.astype(np.float32), kw.pop("x", x), sample=48, **kw)
    checks.run(threads=2)
    for name, (_, _, good) in cases.items():
        acc = res[name]["accuracy"]
        assert acc["knots_checked"] == 48
        worst = max(v["max"] for v in acc.values() if isinstance(v, dict))
        assert (worst < 1e-6) if good else (worst > 1e-2), (name, acc)
    assert set(res["fd"]["accuracy"]) == {"qdd", "minv", "knots_checked"}
    assert set(res["fdgrad_minv"]["accuracy"]) == {"qdd", "dqdd_dq", "dqdd_dqd", "minv", "knots_checked"}
