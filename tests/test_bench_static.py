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


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_static_copy", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)  # defines functions only; main() is not run
    return m


def test_committed_traffic_matches_operator_form():
    """roofline.traffic is read from the committed counter summary only for
    the workload's own size and quadrature-data form; every committed summary
    states both figures."""
    b = _bench_module()
    for name, w in b.WORKLOADS.items():
        path = os.path.join(ROOT, "profiles", f"pa_hex_{name}_hbm_traffic.md")
        if not os.path.exists(path):
            continue
        text = open(path).read()
        assert "traffic_bytes_per_apply:" in text and "qdata_per_point:" in text
        import re
        sq = int(re.search(r"qdata_per_point:\s*([0-9]+)", text).group(1))
        t = b.committed_traffic(name, w["n"], w["p"], sq)
        assert t is not None and t > 0
        assert b.committed_traffic(name, w["n"], w["p"], sq + 1) is None  # other qdata form
        assert b.committed_traffic(name, w["n"] + 1, w["p"], sq) is None  # other size
