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
