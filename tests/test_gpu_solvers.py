"""GPU parity of the device solvers against the oracle and the reference's
own known-answer tests (solver/src/lib.rs:1197-1221, amg/src/lib.rs:254-286,
examples/mfem_ex1_poisson.rs:184-237)."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _lap1d(n):
    rp = [0]; ci = []; v = []
    for i in range(n):
        if i > 0: ci.append(i - 1); v.append(-1.0)
        ci.append(i); v.append(2.0)
        if i < n - 1: ci.append(i + 1); v.append(-1.0)
        rp.append(len(ci))
    return np.array(rp, np.uint64), np.array(ci, np.uint32), np.array(v)


def test_cg_laplacian(fg, fr, ctx):
    # lib.rs:1197-1210 (CG on the 1-D Laplacian, |Ax - b| < 1e-6)
    rp, ci, v = _lap1d(50)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    b = np.ones(50); x = np.zeros(50)
    res = fg.solve_cg_operator(50, 50, A, b, x, fg.SolverConfig())
    assert res.converged
    ref = fr.CsrRef(rp, ci, v)
    assert np.linalg.norm(ref.spmv(x) - b) < 1e-6
    xo, ro = fr.solve_cg_operator(fr._Op(50, csr=ref), b, np.zeros(50))
    assert res.iterations == ro.iterations  # in exact arithmetic <= 25 (symmetry)
    assert res.iterations <= 26
    assert np.abs(x - xo).max() < 1e-10 * np.abs(xo).max()
    assert abs(res.final_residual - ro.final_residual) <= 1e-6 * max(ro.final_residual, 1e-300) + 1e-14


def test_pcg_jacobi_laplacian(fg, ctx):
    # lib.rs:1212-1221: converged, iterations < 60
    A = fg.CsrMatrix(*_lap1d(50), ctx=ctx)
    x = np.zeros(50)
    res = fg.solve_pcg_jacobi(A, np.ones(50), x)
    assert res.converged and res.iterations < 60


def test_p1_poisson_256(fg, fr, ctx):
    """256^2 P1 Poisson, PCG-Jacobi to 1e-10 (mfem_ex1_poisson.rs:71-114)."""
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(256)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    x = np.zeros(len(rhs))
    res = fg.solve_pcg_jacobi(A, rhs, x, fg.SolverConfig(rtol=1e-10, max_iter=5000))
    ref = fr.CsrRef(rp, ci, v)
    xo, ro = fr.pcg_jacobi(fr._Op(len(rhs), csr=ref), ref.diagonal(), rhs, np.zeros(len(rhs)),
                           rtol=1e-10, max_iter=5000)
    assert res.converged and ro.converged
    assert abs(res.iterations - ro.iterations) <= 2
    assert res.final_residual < 1e-10
    assert np.abs(x - xo).max() < 1e-8 * np.abs(xo).max()
    # discretisation error vs sin(pi x) sin(pi y): nodal max error ~ O(h^2)
    X, Y = coords[0::2], coords[1::2]
    assert np.abs(x - np.sin(np.pi * X) * np.sin(np.pi * Y)).max() < 5e-5


def test_ex1_small_golden(fg, fr, ctx):
    z = np.load(os.path.join(GOLD, "csr_fixtures.npz"))
    A = fg.CsrMatrix(z["p8_rp"], z["p8_ci"], z["p8_v"], ctx=ctx)
    x = np.zeros(len(z["p8_b"]))
    res = fg.solve_pcg_jacobi(A, z["p8_b"], x, fg.SolverConfig(rtol=1e-10, max_iter=5000))
    assert res.converged and res.final_residual < 1e-8  # ex1 test thresholds
    assert res.iterations == int(z["p8_iters"][0])
    assert np.abs(x - z["p8_x"]).max() < 1e-12


def test_chebyshev_reference_mode(fg, fr, ctx):
    """chebyshev_smooth / estimate_spectral_radius (amg/src/lib.rs:254-286)."""
    z = np.load(os.path.join(GOLD, "csr_fixtures.npz"))
    rp, ci, v = _lap1d(50)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    rho30 = fg.estimate_spectral_radius(A, 30)
    assert 3.5 < rho30 < 4.1
    assert abs(rho30 - z["cheb_rho30"][0]) < 1e-14
    rho = fg.estimate_spectral_radius(A, 20)
    assert abs(rho - z["cheb_rho20"][0]) < 1e-14
    lmax = 1.1 * rho; lmin = lmax / 30
    x = np.zeros(50)
    fg.chebyshev_smooth(A, x, np.ones(50), lmin, lmax, 20)
    assert np.abs(x - z["cheb_x20"]).max() <= 1e-12 * np.abs(z["cheb_x20"]).max()
    r = np.ones(50) - fr.CsrRef(rp, ci, v).spmv(x)
    assert np.linalg.norm(r) / math.sqrt(50) < 0.9


def test_cg_errors(fg, ctx):
    A = fg.CsrMatrix(*_lap1d(20), ctx=ctx)
    with pytest.raises(fg.DimensionMismatch):
        fg.solve_cg_operator(20, 20, A, np.ones(19), np.zeros(20))
    with pytest.raises(fg.ConvergenceFailed) as ei:
        fg.solve_cg_operator(20, 20, A, np.ones(20), np.zeros(20), fg.SolverConfig(max_iter=2))
    assert ei.value.max_iter == 2
    # zero operator: p^T A p = 0 -> breakdown (lib.rs:245-249)
    Z = fg.CsrMatrix(np.arange(21, dtype=np.uint64), np.arange(20, dtype=np.uint32), np.zeros(20), ctx=ctx)
    with pytest.raises(fg.Linger):
        fg.solve_cg_operator(20, 20, Z, np.ones(20), np.zeros(20))
    # PCG returns converged = false on max_iter (par_solver.rs:340-345)
    res = fg.solve_pcg_jacobi(A, np.ones(20), np.zeros(20), fg.SolverConfig(max_iter=1))
    assert not res.converged and res.iterations == 1
    # zero right-hand side: converged in 0 iterations
    res = fg.solve_pcg_jacobi(A, np.zeros(20), np.zeros(20))
    assert res.converged and res.iterations == 0


def _hex_problem(fg, fr, ctx, p, n, kappa, dirichlet=True, dtype=None):
    op = fg.HexPAOperator(p, p + 1, n, form=fg.DIFFUSION, kappa=kappa, dirichlet=dirichlet,
                          dtype=fg.F64 if dtype is None else dtype, ctx=ctx)
    return op


def test_pcg_chebyshev_hex(fg, fr, ctx):
    """Config-2 family at 12^3 Q2: PCG + Chebyshev-Jacobi(4) vs the oracle."""
    n, p = 12, 2
    op = _hex_problem(fg, fr, ctx, p, n, "sinprod")
    b = op.rhs_unit()
    x = np.zeros(op.n_dofs())
    res, lmax = fg.solve_pcg_chebyshev(op, b, x, fg.SolverConfig(rtol=1e-8), degree=4)
    assert res.converged
    d = fr.hex_desc(p, p + 1, n, kind=fr.FR_DIFFUSION, kappa=fr.coef_sinprod(), dirichlet=True)
    c = fr.hex_vertices(n, n, n, 0.0, 0)
    Dq, Mq = fr.hex_qdata(d, c)
    ob = fr._Op(op.n_dofs(), hexop=(d, Dq, Mq, 4))
    diag = fr.hex_diagonal(d, Dq, Mq)
    assert np.abs(b - fr.hex_rhs_unit(d, c)).max() < 1e-15
    lmax_o = 1.1 * fr.estimate_lmax_jacobi(ob, diag, 20)
    assert abs(lmax - lmax_o) < 1e-10 * lmax_o
    xo, ro = fr.pcg_chebyshev(ob, diag, 4, lmax / 30, lmax, b, np.zeros(op.n_dofs()), rtol=1e-8)
    assert abs(res.iterations - ro.iterations) <= 1
    # exact solution of the discrete system (direct solve) within 1e-10 relative
    rp, ci, v = fr.hex_assemble_csr(d, c)
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    A = sp.csr_matrix((v, ci.astype(np.int64), rp.astype(np.int64)), shape=(len(b), len(b)))
    xe = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-7
    x2 = np.zeros_like(x)
    res2, _ = fg.solve_pcg_chebyshev(op, b, x2, fg.SolverConfig(rtol=1e-12), degree=4)
    assert np.linalg.norm(x2 - xe) / np.linalg.norm(xe) < 1e-10


def test_pcg_jacobi_on_pa(fg, fr, ctx):
    op = _hex_problem(fg, fr, ctx, 3, 6, dict(kind="grf", seed=4))
    b = op.rhs_random(9)
    x = np.zeros(op.n_dofs())
    res = fg.solve_pcg_jacobi(op, b, x, fg.SolverConfig(rtol=1e-10, max_iter=5000))
    assert res.converged
    r = b - op.matvec(x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10


def test_ir_mixed(fg, fr, ctx):
    """FP32 inner PCG + FP64 refinement reaches 1e-10 and matches the FP64 solve
    (Q3 with 4^3 points on the unit cube: both precisions run the quad-per-element
    kernel)."""
    n, p = 10, 3
    op64 = _hex_problem(fg, fr, ctx, p, n, dict(kind="grf", seed=1))
    op32 = op64.convert(fg.F32)
    assert op64.element_kernel() == op32.element_kernel() == "quad-per-element"
    b = op64.rhs_random(5)
    x = np.zeros(op64.n_dofs())
    res, inner = fg.solve_ir_mixed(op64, op32, b, x, fg.SolverConfig(rtol=1e-10), inner_rtol=1e-3, mode=0)
    assert res.converged and res.final_residual < 1e-10 and inner > 0
    x64 = np.zeros_like(x)
    r64, _ = fg.solve_pcg_chebyshev(op64, b, x64, fg.SolverConfig(rtol=1e-13, max_iter=5000))
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-9


@pytest.mark.parametrize("delta,interval", [(1e-2, 0), (1e-3, 0), (1e-2, 1), (1e-2, 5)])
def test_ir_mixed_continued(fg, fr, ctx, delta, interval):
    """Continued refinement (one FP32 PCG, FP64 residual replaced after each
    reduction by delta, or every `interval` iterations) reaches 1e-10 on the
    FP64 residual and matches the FP64 solve, with about as many FP32
    iterations as the FP64 PCG needs."""
    n, p = 10, 3
    op64 = _hex_problem(fg, fr, ctx, p, n, dict(kind="grf", seed=1))
    op32 = op64.convert(fg.F32)
    b = op64.rhs_random(5)
    x = np.zeros(op64.n_dofs())
    res, inner = fg.solve_ir_mixed(op64, op32, b, x, fg.SolverConfig(rtol=1e-10, max_iter=5000), mode=1,
                                   inner_rtol=delta, update_interval=interval)
    assert res.converged and res.final_residual < 1e-10
    assert np.linalg.norm(b - op64.matvec(x)) / np.linalg.norm(b) < 1e-10
    x64 = np.zeros_like(x)
    r64, _ = fg.solve_pcg_chebyshev(op64, b, x64, fg.SolverConfig(rtol=1e-13, max_iter=5000))
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-9
    xp = np.zeros_like(x)
    rp, _ = fg.solve_pcg_chebyshev(op64, b, xp, fg.SolverConfig(rtol=1e-10, max_iter=5000))
    assert inner <= 1.3 * rp.iterations + 5, (inner, rp.iterations)


def test_csr_apply_dirichlet(fg, fr, ctx):
    """apply_dirichlet on the device matrix (constraints.rs:31-42) equals the
    oracle's row zeroing bitwise (values and rhs), including prescribed
    non-zero values, a repeated DOF (last value wins) and the cached diagonal;
    a homogeneous system built this way solves as the oracle's does."""
    n = 16
    coords, conn = fr.unit_square_tri(n)
    N = (n + 1) ** 2
    rp, ci, v = fr.simplex_assemble_csr(2, 1, coords, conn, conn.copy(), N, fr.FR_DIFFUSION, 1.0, 3)
    rhs = fr.p1_source_rhs(n, coords, conn, 3)
    bnd = fr.unit_square_boundary(n)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    d0 = A.diagonal()
    b = rhs.copy()
    dofs = np.concatenate([bnd, bnd[:3]])          # repeat three DOFs with new values
    dvals = np.concatenate([np.zeros(len(bnd)), [7.0, 8.0, 9.0]])
    fg.apply_dirichlet(A, b, dofs, dvals)
    last = np.zeros(len(bnd)); last[:3] = [7.0, 8.0, 9.0]
    vr, br = fr.dirichlet_row_zeroing(rp, ci, v, bnd, rhs, last)
    assert np.array_equal(A.values(), vr)
    assert np.array_equal(b, br)
    d1 = A.diagonal()
    assert np.all(d1[bnd] == 1.0) and np.array_equal(np.delete(d1, bnd), np.delete(d0, bnd))
    # row zeroing keeps CG valid only for homogeneous values with x0 = 0 (the
    # matrix is no longer symmetric): solve the homogeneous system
    A0 = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    b0 = rhs.copy()
    fg.apply_dirichlet(A0, b0, bnd, np.zeros(len(bnd)))
    x = np.zeros(N)
    res = fg.solve_pcg_jacobi(A0, b0, x, fg.SolverConfig(rtol=1e-10, max_iter=5000))
    assert res.converged and np.abs(x[bnd]).max() == 0.0
    vr0, br0 = fr.dirichlet_row_zeroing(rp, ci, v, bnd, rhs)
    ref = fr.CsrRef(rp, ci, vr0)
    xo, ro = fr.pcg_jacobi(fr._Op(N, csr=ref), ref.diagonal(), br0, np.zeros(N), rtol=1e-10, max_iter=5000)
    assert np.abs(x - xo).max() < 1e-10 * np.abs(xo).max()
    with pytest.raises(fg.FemGpuError):
        fg.apply_dirichlet(A, None, [N], [0.0])
    with pytest.raises(ValueError):
        fg.apply_dirichlet(A, None, [0, 1], [0.0])


def _pcg_jacobi_both(fg, ctx, rp, ci, v, b, cfg, monkeypatch):
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("FEMGPU_PCG_FUSED", fused)
        A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
        x = np.zeros(len(b))
        res = fg.solve_pcg_jacobi(A, b, x, cfg)
        out[fused] = (x, res)
    return out["1"], out["0"]


@pytest.mark.parametrize("max_iter", [5000, 37])
def test_pcg_jacobi_fused_matches_separate_kernels(fg, fr, ctx, monkeypatch, max_iter):
    """The two-kernel CSR PCG-Jacobi iteration (search direction formed inside
    the SpMV, scalars finalised redundantly per workgroup) gives the iterates
    of the six-kernel loop bit for bit: same x, iteration count and residual,
    converged (config-1 system at 64^2) and stopped at max_iter."""
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(64)
    cfg = fg.SolverConfig(rtol=1e-10, max_iter=max_iter)
    (xf, rf), (xs, rs) = _pcg_jacobi_both(fg, ctx, rp, ci, v, rhs, cfg, monkeypatch)
    print(f"max_iter {max_iter}: fused {rf}, separate {rs}")
    assert rf == rs
    assert np.array_equal(xf, xs)
    assert rf.converged == (max_iter == 5000)
    if max_iter == 37:
        assert rf.iterations == 37


def test_pcg_jacobi_fused_grid_stride(fg, fr, ctx, monkeypatch):
    """Above 262,144 rows each thread of the fixed 1024-workgroup grid owns
    several rows; the fused loop still matches the separate kernels bitwise,
    and the oracle within rounding (1-D Laplacian, 300,001 rows, 60 iterations)."""
    n = 300001
    rp = np.zeros(n + 1, np.uint64)
    cnt = np.full(n, 3); cnt[0] = cnt[-1] = 2
    rp[1:] = np.cumsum(cnt)
    ci = np.empty(int(rp[-1]), np.uint32); v = np.empty(int(rp[-1]))
    rows = np.repeat(np.arange(n), cnt)
    off = np.arange(int(rp[-1])) - rp[rows].astype(np.int64)
    first = rows == 0
    ci[:] = np.where(first, rows + off, rows - 1 + off).astype(np.uint32)
    v[:] = np.where(ci == rows, 2.0 + 0.01 * (rows % 7), -1.0)
    b = np.sin(np.arange(n) * 0.001) + 0.5
    cfg = fg.SolverConfig(rtol=1e-12, max_iter=60)
    (xf, rf), (xs, rs) = _pcg_jacobi_both(fg, ctx, rp, ci, v, b, cfg, monkeypatch)
    assert rf == rs and rf.iterations == 60 and not rf.converged
    assert np.array_equal(xf, xs)
    ref = fr.CsrRef(rp, ci, v)
    xo, ro = fr.pcg_jacobi(fr._Op(n, csr=ref), ref.diagonal(), b, np.zeros(n), rtol=1e-12, max_iter=60)
    assert np.abs(xf - xo).max() < 1e-10 * np.abs(xo).max()


def _both(monkeypatch, run):
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("FEMGPU_PCG_FUSED", fused)
        out[fused] = run()
    return out["1"], out["0"]


@pytest.mark.parametrize("degree,max_iter", [(4, 5000), (4, 7), (1, 5000), (2, 5000)])
def test_pcg_chebyshev_fused_matches_separate_kernels(fg, fr, ctx, monkeypatch, degree, max_iter):
    """PCG + Chebyshev-Jacobi with the fused vector kernels (alpha, beta
    finalised in every workgroup, the first Chebyshev step inside the update)
    equals the separate-kernel loop bit for bit: hex Q2 GRF operator with
    essential DOFs, converged and stopped at max_iter, degrees 1, 2, 4."""
    op = _hex_problem(fg, fr, ctx, 2, 9, dict(kind="grf", seed=3))
    b = op.rhs_random(2)
    cfg = fg.SolverConfig(rtol=1e-10, max_iter=max_iter)

    def run():
        x = np.zeros(op.n_dofs())
        res, lmax = fg.solve_pcg_chebyshev(op, b, x, cfg, degree=degree)
        return x, res, lmax
    (xf, rf, lf), (xs, rs, ls) = _both(monkeypatch, run)
    print(f"degree {degree} max_iter {max_iter}: fused {rf}, separate {rs}")
    assert lf == ls and rf == rs
    assert np.array_equal(xf, xs)
    assert rf.converged == (max_iter == 5000)


def test_pcg_chebyshev_fused_csr_and_continued(fg, fr, ctx, monkeypatch):
    """The fused loop on a CSR operator (config-1 matrix at 48^2, a grid-stride
    free size) and inside continued mixed-precision refinement (segments that
    stop before the preconditioner, then resume with the last direction)
    equals the separate kernels bit for bit."""
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(48)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    cfg = fg.SolverConfig(rtol=1e-10, max_iter=5000)

    def run_csr():
        x = np.zeros(len(rhs))
        res, _ = fg.solve_pcg_chebyshev(A, rhs, x, cfg, degree=3)
        return x, res
    (xf, rf), (xs, rs) = _both(monkeypatch, run_csr)
    assert rf == rs and rf.converged and np.array_equal(xf, xs)

    op64 = _hex_problem(fg, fr, ctx, 3, 8, dict(kind="grf", seed=1))
    op32 = op64.convert(fg.F32)
    b = op64.rhs_random(5)

    def run_ir():
        x = np.zeros(op64.n_dofs())
        res, inner = fg.solve_ir_mixed(op64, op32, b, x, cfg, mode=1, inner_rtol=1e-2, update_interval=16)
        return x, res, inner
    (xf, rf, nf), (xs, rs, ns) = _both(monkeypatch, run_ir)
    assert rf == rs and nf == ns and rf.converged and np.array_equal(xf, xs)
