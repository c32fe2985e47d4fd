"""Static checks of the driver scripts (CPU): every global name a function of
bench.py or __graft_entry__.py reads is defined at module level or is a
builtin, so no code path ends in a NameError; --gpus above 1 is refused."""
import builtins
import os
import subprocess
import symtable
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _unresolved(path):
    src = open(path).read()
    top = symtable.symtable(src, path, "exec")
    module = {s.get_name() for s in top.get_symbols() if s.is_assigned() or s.is_imported()}
    bad = []

    def walk(t):
        for s in t.get_symbols():
            if t is not top and s.is_global() and s.is_referenced():
                n = s.get_name()
                if n not in module and not hasattr(builtins, n):
                    bad.append((t.get_name(), n))
        for c in t.get_children():
            walk(c)
    walk(top)
    return bad


@pytest.mark.parametrize("name", ["bench.py", "__graft_entry__.py"])
def test_global_names_resolve(name):
    assert _unresolved(os.path.join(ROOT, name)) == []


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
