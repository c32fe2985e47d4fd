"""GPU tests of the boundary additions that mirror reference interfaces:

- solve_cg_operator on a caller-supplied operator (the `apply: F` closure,
  solver/src/lib.rs:198-208), with the closure on the host and on device
  pointers, against the oracle's operator CG and the CSR path
  (solver/tests/operator_cg.rs:73-97);
- the solver reporting of SolverConfig.verbose / print_level
  (solver/src/lib.rs:79-142, parallel/src/par_solver.rs:312-313): per-iteration
  residuals recorded on the device, returned by residual_history and printed.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lap1d(n):
    """1-D Laplacian tridiag(-1, 2, -1) (the reference's solver test matrix,
    solver/src/lib.rs:1197-1221) as CSR and as a numpy apply."""
    rp = np.zeros(n + 1, np.uint64)
    cnt = np.full(n, 3)
    cnt[0] = cnt[-1] = 2
    rp[1:] = np.cumsum(cnt)
    rows = np.repeat(np.arange(n), cnt)
    off = np.arange(int(rp[-1])) - rp[rows].astype(np.int64)
    ci = np.where(rows == 0, rows + off, rows - 1 + off).astype(np.uint32)
    v = np.where(ci == rows, 2.0, -1.0)

    def apply(x, y):
        y[:] = 2.0 * x
        y[1:] -= x[:-1]
        y[:-1] -= x[1:]
    return rp, ci, v, apply


def test_cg_operator_host_callback_vs_oracle(fg, fr, ctx):
    """The closure runs on the host: same iteration count as the oracle's
    operator CG with the same closure, solution within 1e-12, and within 1e-8
    of CSR CG (operator_cg.rs:73-97); ||Ax - b|| < 1e-6 (lib.rs:1197-1221)."""
    n = 50
    rp, ci, v, apply = _lap1d(n)
    b = np.ones(n)
    x = np.zeros(n)
    calls = []

    def counted(xx, yy):
        calls.append(1)
        apply(xx, yy)
    res = fg.solve_cg_operator(n, n, counted, b, x, fg.SolverConfig(rtol=1e-10, max_iter=1000))
    def fr_apply(xx):
        y = np.empty_like(xx)
        apply(xx, y)
        return y
    xo, ro = fr.solve_cg_operator(fr._Op(n, fn=fr_apply), b, np.zeros(n), rtol=1e-10, max_iter=1000)
    assert res.converged and res.iterations == ro.iterations
    assert len(calls) >= res.iterations + 1  # r0 plus one apply per iteration
    assert np.abs(x - xo).max() < 1e-12 * np.abs(xo).max()
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    xc = np.zeros(n)
    fg.solve_cg_operator(n, n, A, b, xc, fg.SolverConfig(rtol=1e-10, max_iter=1000))
    assert np.abs(x - xc).max() < 1e-8
    y = np.empty(n)
    apply(x, y)
    assert np.linalg.norm(y - b) < 1e-6
    # in exact arithmetic CG on this symmetric problem takes <= n/2 steps (b is even)
    assert res.iterations <= n // 2 + 1


def test_callback_device_pointers_bitwise(fg, ctx):
    """A device-pointer closure that applies a CSR matrix on the context's
    stream gives CG and PCG-Jacobi the iterates of the CSR operator itself,
    bit for bit (the same SpMV kernel and the same vector kernels)."""
    n = 3001
    rp, ci, v, _ = _lap1d(n)
    v[v > 0] += 0.01 * (np.arange(n) % 5)  # a varying diagonal
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    import femgpu
    applies = []

    def dev_apply(xp, yp):
        applies.append(1)
        femgpu._check(femgpu.lib.femgpu_op_apply(A.h, C.c_void_p(xp), C.c_void_p(yp), 0, femgpu.DEVICE))
    op = fg.CallbackOperator(n, dev_apply, mem=fg.DEVICE, diag=A.diagonal(), ctx=ctx)
    b = np.sin(np.arange(n) * 0.01) + 0.3
    cfg = fg.SolverConfig(rtol=1e-11, max_iter=20000)
    x1, x2 = np.zeros(n), np.zeros(n)
    r1 = fg.solve_cg_operator(n, n, A, b, x1, cfg)
    r2 = fg.solve_cg_operator(n, n, op, b, x2, cfg)
    assert r1 == r2 and np.array_equal(x1, x2)
    # PCG-Jacobi: the callback operator runs the separate-kernel loop, the CSR
    # one the fused loop; the two are bitwise the same iteration
    x1, x2 = np.zeros(n), np.zeros(n)
    r1 = fg.solve_pcg_jacobi(A, b, x1, cfg)
    r2 = fg.solve_pcg_jacobi(op, b, x2, cfg)
    assert r1 == r2 and np.array_equal(x1, x2)
    # MatFreeOperator::apply accumulates
    xx = np.random.default_rng(0).standard_normal(n)
    y0 = np.random.default_rng(1).standard_normal(n)
    y1 = y0.copy()
    op.apply(xx, y1)
    assert np.abs(y1 - (y0 + A.matvec(xx))).max() < 1e-14 * np.abs(y1).max()
    assert len(applies) > 0


def test_callback_errors(fg, ctx):
    n = 20
    _, _, _, apply = _lap1d(n)

    def bad(x, y):
        raise ValueError("operator failed")
    with pytest.raises(ValueError, match="operator failed"):
        fg.solve_cg_operator(n, n, bad, np.ones(n), np.zeros(n))
    op = fg.CallbackOperator(n, apply, ctx=ctx)  # no diagonal
    with pytest.raises(fg.FemGpuError):
        fg.solve_pcg_jacobi(op, np.ones(n), np.zeros(n))
    with pytest.raises(fg.DimensionMismatch):
        fg.solve_cg_operator(n, n, apply, np.ones(n + 1), np.zeros(n))


def test_residual_history_pcg_jacobi(fg, fr, ctx, capfd, monkeypatch):
    """verbose = true (PrintLevel::Iterations): the history holds the initial
    residual and the relative residual of every iteration, ending at
    final_residual; the fused CSR loop and the separate-kernel loop record the
    same history; one line per iteration is printed, as par_solve_pcg_jacobi
    logs them; Summary prints one line and records nothing new."""
    rp, ci, v, rhs, _, _ = fr.poisson_p1(48)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    hists = []
    for fused in ("1", "0"):
        monkeypatch.setenv("FEMGPU_PCG_FUSED", fused)
        x = np.zeros(len(rhs))
        res = fg.solve_pcg_jacobi(A, rhs, x, fg.SolverConfig(rtol=1e-10, max_iter=5000, verbose=True))
        h = fg.residual_history(ctx)
        assert res.converged and len(h) == res.iterations + 1
        assert h[0] == pytest.approx(1.0, rel=1e-14) and h[-1] == res.final_residual
        assert np.all(h[:-1] >= 1e-10) and h[-1] < 1e-10
        hists.append(h)
    assert np.array_equal(hists[0], hists[1])
    # the oracle's PCG-Jacobi reaches the same count
    ref = fr.CsrRef(rp, ci, v)
    _, ro = fr.pcg_jacobi(fr._Op(len(rhs), csr=ref), ref.diagonal(), rhs, np.zeros(len(rhs)), rtol=1e-10,
                          max_iter=5000)
    assert abs(ro.iterations - (len(hists[0]) - 1)) <= 2
    out = capfd.readouterr().err
    lines = [l for l in out.splitlines() if l.startswith("femgpu pcg_jacobi iter ")]
    assert len(lines) == 2 * (len(hists[0]) - 1)
    assert lines[0].endswith(f"residual = {hists[0][1]:.3e}")
    monkeypatch.setenv("FEMGPU_PCG_FUSED", "1")
    x = np.zeros(len(rhs))
    res = fg.solve_pcg_jacobi(A, rhs, x, fg.SolverConfig(rtol=1e-10, max_iter=5000, print_level=1))
    out = capfd.readouterr().err
    assert out.count("femgpu pcg_jacobi: converged after") == 1 and "iter 1:" not in out


def test_residual_history_cg_and_chebyshev(fg, ctx):
    """The history of operator CG holds absolute residuals (final_residual is
    absolute there, lib.rs:262-268) and PCG-Chebyshev's relative ones; a solve
    stopped at max_iter records max_iter + 1 entries."""
    n = 200
    rp, ci, v, _ = _lap1d(n)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    b = np.ones(n)
    x = np.zeros(n)
    res = fg.solve_cg_operator(n, n, A, b, x, fg.SolverConfig(rtol=1e-10, max_iter=1000, print_level=2))
    h = fg.residual_history(ctx)
    assert len(h) == res.iterations + 1 and h[0] == pytest.approx(np.sqrt(n)) and h[-1] == res.final_residual
    op = fg.HexPAOperator(2, 3, 6, form=fg.DIFFUSION, kappa="sinprod", dirichlet=True, ctx=ctx)
    bb = op.rhs_unit()
    xx = np.zeros_like(bb)
    res, _ = fg.solve_pcg_chebyshev(op, bb, xx, fg.SolverConfig(rtol=1e-30, max_iter=9, verbose=True))
    h = fg.residual_history(ctx)
    assert not res.converged and len(h) == 10 and h[-1] == res.final_residual
    assert h[0] == pytest.approx(1.0, rel=1e-14) and h[-1] < h[0]


def _tet_mesh(n):
    g = np.linspace(0, 1, n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).ravel()
    vid = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    conn = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for pm in [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]:
                    cur = [i, j, k]
                    path = [vid(*cur)]
                    for ax in pm:
                        cur[ax] += 1
                        path.append(vid(*cur))
                    conn += path
    return coords, np.array(conn, np.uint32)


def _kappa(xq):
    # a coefficient of the physical point, as a ScalarCoeff closure gives it
    return 1.0 + xq[..., 0] + 0.5 * xq[..., 1] ** 2


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2)])
def test_simplex_coefficient_per_point(fg, fr, ctx, dim, order):
    """PADiffusionOperator<S, K: ScalarCoeff> / PAMassOperator with the
    coefficient evaluated at every quadrature point (partial.rs:116-117,
    137-147, 186-190): the device PA apply, the device assembly (bit for bit)
    and the lumped mass against the oracle with the same point values; the
    quadrature points themselves equal the oracle's.  A constant array gives
    the constant coefficient's operator."""
    if dim == 2:
        coords, conn = fr.unit_square_tri(12)
    else:
        coords, conn = _tet_mesh(3)
    nd, ed = fr.simplex_dofs(dim, order, len(coords) // dim, conn)
    qo = 2 * order + 1
    xq = fg.simplex_qp_coords(dim, coords, conn, qo)
    assert np.array_equal(xq, fr.simplex_qp_coords(dim, coords, conn, qo))
    kq = _kappa(xq)
    x = np.random.default_rng(dim * 10 + order).standard_normal(nd)
    for Op, kind, kw in ((fg.PADiffusionOperator, fr.FR_DIFFUSION, "kappa"), (fg.PAMassOperator, fr.FR_MASS, "rho")):
        op = Op(dim, order, coords, conn, ed, nd, quad_order=qo, ctx=ctx, **{kw: _kappa})
        ref = fr.simplex_pa_apply(dim, order, coords, conn, ed, kind, 0.0, qo, x, coef_qp=kq)
        assert np.abs(op.matvec(x) - ref).max() < 1e-12 * np.abs(ref).max()
        rp, ci, v = fr.simplex_assemble_csr(dim, order, coords, conn, ed, nd, kind, 0.0, qo, coef_qp=kq)
        assert np.abs(op.diagonal() - fr.CsrRef(rp, ci, v).diagonal()).max() < 1e-12 * np.abs(v).max()
        A = fg.assemble_bilinear(dim, order, coords, conn, ed, nd, kind, kq, qo, ctx=ctx)
        rpg, cig = A.pattern()
        assert np.array_equal(rpg, rp) and np.array_equal(cig, ci)
        if order == 1:  # same basis values on both sides: bit for bit
            assert np.array_equal(A.values(), v)
        else:  # the P2 / P3 tables come from two restatements
            assert np.abs(A.values() - v).max() < 1e-13 * np.abs(v).max()
    K1 = fg.PADiffusionOperator(dim, order, coords, conn, ed, nd, kappa=2.5, quad_order=qo, ctx=ctx)
    Kq = fg.PADiffusionOperator(dim, order, coords, conn, ed, nd, kappa=np.full(kq.shape, 2.5), quad_order=qo,
                                ctx=ctx)
    assert np.abs(K1.matvec(x) - Kq.matvec(x)).max() < 1e-13 * np.abs(K1.matvec(x)).max()
    lm = fg.LumpedMassOperator(dim, order, coords, conn, ed, nd, rho=kq, quad_order=qo, ctx=ctx)
    ref = fr.simplex_lumped_mass(dim, order, coords, conn, ed, nd, 0.0, qo, rho_qp=kq)
    assert np.abs(lm.diag - ref).max() < 1e-13 * np.abs(ref).max()
    if order == 1:  # P1 patch: v^T K u = a.b int(1 + x), 1^T M 1 = int(1 + x)
        c = np.asarray(coords).reshape(-1, dim)
        a, b = np.arange(1.0, dim + 1.0), np.linspace(-1.0, 0.5, dim)
        K = fg.PADiffusionOperator(dim, 1, coords, conn, ed, nd, kappa=lambda p: 1.0 + p[..., 0],
                                   quad_order=qo, ctx=ctx)
        assert abs((c @ b) @ K.matvec(c @ a) - 1.5 * (a @ b)) < 1e-12
        M = fg.PAMassOperator(dim, 1, coords, conn, ed, nd, rho=lambda p: 1.0 + p[..., 0], quad_order=qo, ctx=ctx)
        assert abs(M.matvec(np.ones(nd)).sum() - 1.5) < 1e-13


def test_ir_continued_recovers_from_inner_breakdown(fg, ctx):
    """Continued refinement whose FP32 PCG breaks down (p^T A p = 0: the FP32
    operator's closure returns zeros for the fifth search direction) restarts
    the FP32 recursion from the replaced FP64 residual and still reaches 1e-10;
    the reported inner count is the iterations that ran (comparable with the
    run without the breakdown), not the iteration budget."""
    n = 300
    rp, ci, v, apply = _lap1d(n)
    A64 = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    b = np.ones(n)
    cfg = fg.SolverConfig(rtol=1e-10, max_iter=20000)
    counts = {}
    for sabotage in (False, True):
        calls = []

        def apply32(xx, yy):
            calls.append(1)
            if sabotage and len(calls) == 5:
                yy[:] = 0.0
                return
            t = np.empty(n)
            apply(xx.astype(np.float64), t)
            yy[:] = t.astype(np.float32)
        A32 = fg.CallbackOperator(n, apply32, dtype=fg.F32, diag=np.full(n, 2.0), ctx=ctx)
        x = np.zeros(n)
        # degree 1: the preconditioner applies no A, so the k-th closure call is
        # the k-th iteration's A p
        res, inner = fg.solve_ir_mixed(A64, A32, b, x, cfg, degree=1, lmax=2.0, mode=1, update_interval=0)
        assert res.converged, (sabotage, res, inner)
        y = np.empty(n)
        apply(x, y)
        assert np.linalg.norm(y - b) / np.linalg.norm(b) < 1e-10
        counts[sabotage] = inner
        A32.close()
    print(f"inner iterations: {counts[False]} without, {counts[True]} with the breakdown")
    assert counts[True] < 3 * counts[False] + 50
    assert counts[True] < cfg.max_iter // 4
