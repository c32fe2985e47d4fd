"""The C-ABI library loads and exports every entry point include/femgpu.h
declares (no GPU needed: no compute call is made here)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fem-rs_amd", "femgpu", "libfemgpu.so")


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:int|const char \*)\s*(femgpu_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_header():
    assert os.path.exists(LIB), "build libfemgpu.so first (make -C fem-rs_amd)"
    lib = ctypes.CDLL(LIB)
    declared = _declared()
    assert len(declared) > 25
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (femgpu_\w+)", out))
    assert exported == declared, exported ^ declared
    lib.femgpu_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.femgpu_version()


def test_gfx950_code_object():
    """The embedded device code objects target gfx950 (and nothing else)."""
    blob = open(LIB, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx\w+)", blob))
    assert targets == {b"gfx950"}, targets


def test_python_mirror_imports():
    import femgpu as fg
    for name in ["CsrMatrix", "PAMassOperator", "PADiffusionOperator", "LumpedMassOperator",
                 "HexPAOperator", "SolverConfig", "SolveResult", "solve_cg_operator",
                 "solve_pcg_jacobi", "chebyshev_smooth", "estimate_spectral_radius",
                 "solve_pcg_chebyshev", "solve_ir_mixed", "ConvergenceFailed", "DimensionMismatch",
                 "Linger"]:
        assert hasattr(fg, name)
    cfg = fg.SolverConfig()
    assert (cfg.rtol, cfg.atol, cfg.max_iter) == (1e-8, 0.0, 1000)  # lib.rs:104-108


def _arity(params):
    params = params.strip()
    if params in ("", "void"):
        return 0
    depth, n = 0, 1
    for ch in params:
        if ch in "(<":
            depth += 1
        elif ch in ")>":
            depth -= 1
        elif ch == "," and depth == 0:
            n += 1
    return n


def test_rust_sys_crate_matches_header():
    """The committed fem-gpu-sys crate declares every C entry point with the
    header's name and parameter count, and nothing else."""
    hdr = open(os.path.join(ROOT, "include", "femgpu.h")).read()
    c_decl = {m.group(1): _arity(m.group(2)) for m in
              re.finditer(r"^\s*(?:int|const char \*)\s*(femgpu_\w+)\s*\(([^;]*?)\)\s*;", hdr, re.M | re.S)}
    rs = open(os.path.join(ROOT, "fem-rs_amd", "rust", "fem-gpu-sys", "src", "lib.rs")).read()
    block = rs[rs.index('extern "C" {'):]
    r_decl = {m.group(1): _arity(m.group(2)) for m in
              re.finditer(r"pub fn (femgpu_\w+)\s*\(([^;]*?)\)\s*->", block, re.S)}
    assert c_decl == r_decl, (set(c_decl) ^ set(r_decl), {k for k in c_decl if c_decl[k] != r_decl.get(k)})
    assert set(c_decl) == _declared()
