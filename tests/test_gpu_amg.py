"""Smoothed-aggregation AMG (AmgSolver / solve_amg_cg, amg/src/lib.rs:125-207;
the reference's hierarchy lives in the absent linger crate, so these tests pin
the algebra: Galerkin coarse operators, zero prolongator rows for eliminated
DOFs, a symmetric V-cycle, and PCG results against direct and oracle solves)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

pytestmark = pytest.mark.gpu


def _csr(rp, ci, v, shape):
    return sp.csr_matrix((v, ci.astype(np.int64), rp.astype(np.int64)), shape=shape)


def _lap2d_interior(m):
    """5-point Laplacian on the m x m interior grid (SPD, Dirichlet eliminated)."""
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(m, m))
    A = (sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m))).tocsr()
    A.sort_indices()
    return A


def test_amg_hierarchy_galerkin(fg, fr, ctx):
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(64)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    amg = fg.AmgSolver(A)
    L = amg.n_levels()
    assert L >= 3
    bnd = fr.unit_square_boundary(64)
    sizes = []
    for l in range(L - 1):
        Al = _csr(*amg.level(l, "A"))
        P = _csr(*amg.level(l, "P"))
        Ac = _csr(*amg.level(l + 1, "A"))
        sizes.append(Al.shape[0])
        ref = (P.T @ Al @ P).tocsr()
        assert abs(Ac - ref).max() <= 1e-12 * abs(ref).max()
        assert abs(Ac - Ac.T).max() <= 1e-12 * abs(Ac).max()
        assert P.shape[1] < P.shape[0]
        if l == 0:
            assert not P[bnd].toarray().any()  # eliminated (identity) rows stay out of the aggregates
    assert np.array_equal(_csr(*amg.level(0, "A")).toarray()[:5], _csr(rp, ci, v, (len(rhs),) * 2).toarray()[:5])
    print("level sizes", sizes)


def test_amg_vcycle_symmetric_and_pcg(fg, ctx):
    """On an SPD matrix the V-cycle is symmetric (v.Vu = u.Vv), and AMG-PCG
    solves to 1e-10 in far fewer iterations than PCG-Jacobi, matching a
    direct solve."""
    M = _lap2d_interior(100)
    n = M.shape[0]
    A = fg.CsrMatrix(M.indptr.astype(np.uint64), M.indices.astype(np.uint32), M.data, ctx=ctx)
    amg = fg.AmgSolver(A)
    rng = np.random.default_rng(0)
    u, w = rng.standard_normal(n), rng.standard_normal(n)
    Vu, Vw = amg.vcycle(u), amg.vcycle(w)
    assert abs(w @ Vu - u @ Vw) <= 1e-10 * abs(w @ Vu)
    assert u @ Vu > 0 and w @ Vw > 0
    b = rng.standard_normal(n)
    x = np.zeros(n)
    res = amg.solve(b, x, fg.SolverConfig(rtol=1e-10, max_iter=500))
    xj = np.zeros(n)
    rj = fg.solve_pcg_jacobi(A, b, xj, fg.SolverConfig(rtol=1e-10, max_iter=5000))
    print(f"AMG-PCG {res.iterations} it, PCG-Jacobi {rj.iterations} it")
    assert res.converged and res.iterations * 5 < rj.iterations
    xe = spl.spsolve(M.tocsc(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-8


def test_amg_p1_poisson_256(fg, fr, ctx):
    """256^2 P1 Poisson (row-zeroed Dirichlet): AMG-PCG reaches 1e-10
    and agrees with the oracle's PCG-Jacobi solution."""
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(256)
    N = len(rhs)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    x = np.zeros(N)
    res = fg.solve_amg_cg(A, rhs, x, fg.SolverConfig(rtol=1e-10, max_iter=500))
    assert res.converged and res.final_residual < 1e-10
    ref = fr.CsrRef(rp, ci, v)
    xo, ro = fr.pcg_jacobi(fr._Op(N, csr=ref), ref.diagonal(), rhs, np.zeros(N), rtol=1e-12, max_iter=5000)
    print(f"AMG-PCG {res.iterations} it, oracle PCG-Jacobi {ro.iterations} it")
    assert res.iterations * 5 < ro.iterations
    assert np.abs(x - xo).max() < 1e-8 * np.abs(xo).max()


def test_amg_errors(fg, ctx):
    op = fg.HexPAOperator(2, 3, 3, ctx=ctx)
    with pytest.raises(fg.FemGpuError):
        fg.AmgSolver(op)


def _p1_l2_error(fr, u, coords, conn):
    """L2 error of a P1 field against sin(pi x) sin(pi y) with the 5th-order
    triangle rule (amg/tests/poisson_amg.rs:38-61)."""
    pts, wts = fr.tri_rule(5)
    X = coords.reshape(-1, 2)
    T = conn.reshape(-1, 3).astype(np.int64)
    x0, x1, x2 = X[T[:, 0]], X[T[:, 1]], X[T[:, 2]]
    det = np.abs((x1[:, 0] - x0[:, 0]) * (x2[:, 1] - x0[:, 1]) - (x2[:, 0] - x0[:, 0]) * (x1[:, 1] - x0[:, 1]))
    err = 0.0
    for (xi, eta), w in zip(pts, wts):
        uh = (1 - xi - eta) * u[T[:, 0]] + xi * u[T[:, 1]] + eta * u[T[:, 2]]
        xp = x0 + xi * (x1 - x0) + eta * (x2 - x0)
        d = uh - np.sin(np.pi * xp[:, 0]) * np.sin(np.pi * xp[:, 1])
        err += float(np.sum(w * det * d * d))
    return float(np.sqrt(err))


def test_amg_reference_poisson(fg, fr, ctx):
    """The reference's AMG integration tests (amg/tests/poisson_amg.rs:67-115),
    rtol 1e-8, atol 1e-10, max_iter 500: on 16x16 P1 Poisson AMG-CG converges
    with L2 error < 6e-3; one hierarchy with >= 2 levels solves twice; on 64x64
    AMG-CG needs fewer iterations than plain CG."""
    cfg = fg.SolverConfig(rtol=1e-8, atol=1e-10, max_iter=500)
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(16)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    x = np.zeros(len(rhs))
    res = fg.solve_amg_cg(A, rhs, x, cfg)
    assert res.converged
    assert _p1_l2_error(fr, x, coords, conn) < 6e-3
    amg = fg.AmgSolver.setup(A)
    assert amg.n_levels() >= 2
    for _ in range(2):
        x = np.zeros(len(rhs))
        assert amg.solve(rhs, x, cfg).converged
        assert _p1_l2_error(fr, x, coords, conn) < 6e-3
    amg.close()

    rp, ci, v, rhs, coords, conn = fr.poisson_p1(64)
    N = len(rhs)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    x = np.zeros(N)
    res = fg.solve_amg_cg(A, rhs, x, cfg)
    _, rc = fr.solve_cg_operator(fr._Op(N, csr=fr.CsrRef(rp, ci, v)), rhs, np.zeros(N), rtol=1e-8, atol=1e-10,
                                 max_iter=500)
    print(f"64x64: CG {rc.iterations} it, AMG-CG {res.iterations} it")
    assert res.converged and res.iterations < rc.iterations


def _lap1d(fg, ctx, n):
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(n, n)).tocsr()
    T.sort_indices()
    return T, fg.CsrMatrix(T.indptr.astype(np.uint64), T.indices.astype(np.uint32), T.data, ctx=ctx)


def test_amg_reference_unit_tests(fg, ctx):
    """The reference's in-crate AMG tests (amg/src/lib.rs:232-340) on the 1-D
    Laplacian with b = 1 and the default SolverConfig: V-cycle AMG-CG under 30
    iterations (n = 100); one hierarchy of >= 2 levels reused three times
    (n = 80); Chebyshev(degree 3, ratio 3) smoothing under 50 iterations
    (n = 100); F-cycle under 30 iterations (n = 100); Chebyshev + F-cycle
    converges (n = 80).  Each solution is checked against a direct solve."""
    cfg = fg.SolverConfig()

    def check(M, x, res, limit):
        xe = spl.spsolve(M.tocsc(), np.ones(M.shape[0]))
        assert res.converged and res.iterations < limit, res.iterations
        assert np.linalg.norm(x - xe) <= 1e-6 * np.linalg.norm(xe)

    M, A = _lap1d(fg, ctx, 100)
    x = np.zeros(100)
    check(M, x, fg.solve_amg_cg(A, np.ones(100), x, cfg), 30)
    x = np.zeros(100)
    check(M, x, fg.solve_amg_cg(A, np.ones(100), x, cfg, smoother_degree=3, smoother_ratio=3.0), 50)
    s = fg.AmgSolver.setup(A).with_cycle("F")
    x = np.zeros(100)
    check(M, x, s.solve(np.ones(100), x, cfg), 30)
    s.close()

    M, A = _lap1d(fg, ctx, 80)
    s = fg.AmgSolver.setup(A)
    assert s.n_levels() >= 2
    for _ in range(3):
        x = np.zeros(80)
        check(M, x, s.solve(np.ones(80), x, cfg), 1000)
    s.close()
    s = fg.AmgSolver.setup(A, smoother_degree=3, smoother_ratio=3.0).with_cycle(fg.CYCLE_F)
    x = np.zeros(80)
    check(M, x, s.solve(np.ones(80), x, cfg), 1000)
    s.close()


def test_amg_cycles(fg, ctx):
    """W-cycle: symmetric like the V-cycle and contracts at least as well per
    cycle; F-cycle and W-cycle PCG agree with a direct solve; an unknown cycle
    kind is refused."""
    M = _lap2d_interior(120)
    n = M.shape[0]
    A = fg.CsrMatrix(M.indptr.astype(np.uint64), M.indices.astype(np.uint32), M.data, ctx=ctx)
    s = fg.AmgSolver(A)
    assert s.n_levels() >= 4
    rng = np.random.default_rng(3)
    u, w = rng.standard_normal(n), rng.standard_normal(n)
    xe = spl.spsolve(M.tocsc(), u)
    contraction = {}
    for kind in ("V", "W", "F"):
        s.with_cycle(kind)
        Vu = s.vcycle(u)
        contraction[kind] = np.linalg.norm(M @ (xe - Vu)) / np.linalg.norm(u)  # residual after one cycle from 0
        if kind != "F":
            Vw = s.vcycle(w)
            assert abs(w @ Vu - u @ Vw) <= 1e-10 * abs(w @ Vu)
        x = np.zeros(n)
        res = s.solve(u, x, fg.SolverConfig(rtol=1e-10, max_iter=300))
        assert res.converged
        assert np.linalg.norm(x - xe) <= 1e-8 * np.linalg.norm(xe)
        print(kind, res.iterations, "it, one-cycle residual", contraction[kind])
    assert contraction["W"] <= contraction["V"] * 1.05
    with pytest.raises(fg.FemGpuError):
        s.with_cycle(7)
    with pytest.raises(fg.FemGpuError):
        s.with_cycle("X")
    s.close()


def test_amg_graph_replays(fg, ctx, monkeypatch):
    """The cycle replayed from its HIP graph equals the directly launched
    cycle bit for bit on every replay (with other GPU work and host syncs in
    between), for V, W and F cycles, and AMG-PCG takes the same iterates."""
    M = _lap2d_interior(90)
    n = M.shape[0]
    A = fg.CsrMatrix(M.indptr.astype(np.uint64), M.indices.astype(np.uint32), M.data, ctx=ctx)
    s = fg.AmgSolver(A)
    rng = np.random.default_rng(5)
    b = fg.DeviceArray(ctx, n, np.float64, rng.standard_normal(n))
    x = fg.DeviceArray(ctx, n, np.float64)
    other = fg.DeviceArray(ctx, n, np.float64, rng.standard_normal(n))
    for kind in ("V", "W", "F"):
        s.with_cycle(kind)
        monkeypatch.setenv("FEMGPU_GRAPHS", "0")
        s.vcycle(b, x)
        ref = x.download()
        monkeypatch.delenv("FEMGPU_GRAPHS")
        for rep in range(4):
            x.upload(np.full(n, 7.0))
            s.vcycle(b, x)  # captured on the first call, replayed after
            assert np.array_equal(x.download(), ref), (kind, rep)
            A.spmv(other, x)  # unrelated work between replays
            ctx.synchronize()
    s.with_cycle("V")
    xs = []
    for env in ("0", None):
        if env:
            monkeypatch.setenv("FEMGPU_GRAPHS", env)
        else:
            monkeypatch.delenv("FEMGPU_GRAPHS", raising=False)
        xh = np.zeros(n)
        r = s.solve(b.download(), xh, fg.SolverConfig(rtol=1e-10, max_iter=300))
        xs.append((r.iterations, xh))
    assert xs[0][0] == xs[1][0] and np.array_equal(xs[0][1], xs[1][1])
    s.close()


def test_amg_single_level_and_tiny(fg, ctx):
    """A matrix at or below the coarsest size is one level solved by the dense
    inverse: one cycle is the exact solve and PCG stops after one iteration;
    a 1 x 1 matrix works too."""
    for n in (1, 3, 40):
        T = sp.diags([-1.0, 2.5, -1.0], [-1, 0, 1], shape=(n, n)).tocsr()
        T.sort_indices()
        A = fg.CsrMatrix(T.indptr.astype(np.uint64), T.indices.astype(np.uint32), T.data, ctx=ctx)
        s = fg.AmgSolver(A)
        assert s.n_levels() == 1
        b = np.linspace(1.0, 2.0, n)
        xe = np.linalg.solve(T.toarray(), b)
        z = s.vcycle(b)
        assert np.abs(z - xe).max() <= 1e-12 * np.abs(xe).max()
        x = np.zeros(n)
        r = s.solve(b, x, fg.SolverConfig(rtol=1e-10))
        assert r.converged and r.iterations <= 1
        assert np.abs(x - xe).max() <= 1e-10 * np.abs(xe).max()
        s.close()


def test_amg_staged_levels_bitwise(fg, fr, ctx, monkeypatch):
    """Levels from 2^19 rows run the level SpMV with each workgroup's entries
    staged in LDS, in the same row order. An AMG-PCG solve on the 768^2 P1
    Poisson matrix (591,361 rows) with staging on by default equals the one with
    staging forced off (FEMGPU_CSR_STAGE=0), bit for bit."""
    rp, ci, v, rhs, _, _ = fr.poisson_p1(768)
    out = []
    for stage in ("0", None):
        if stage is None:
            monkeypatch.delenv("FEMGPU_CSR_STAGE", raising=False)
        else:
            monkeypatch.setenv("FEMGPU_CSR_STAGE", stage)
        A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
        amg = fg.AmgSolver(A)
        x = np.zeros(len(rhs))
        res = amg.solve(rhs, x, fg.SolverConfig(rtol=1e-8, max_iter=200))
        assert res.converged
        out.append((res.iterations, x))
        amg.close()
        A.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("lanes", [None, "32"])
def test_amg_long_rows(fg, fr, ctx, monkeypatch, lanes):
    """AMG set-up takes any f64 CSR, whatever SpMV layout the matrix got:
    the Q1 hex diffusion + mass matrix of a perturbed 10^3 mesh (27 nonzeros
    per interior row, so the SpMV runs several lanes per row; FEMGPU_CSR_LANES
    forces 32) solves to 1e-10 and agrees with a direct solve."""
    if lanes:
        monkeypatch.setenv("FEMGPU_CSR_LANES", lanes)
    n = (10, 10, 10)
    c = fr.hex_vertices(*n, 0.2, 4)
    d = fr.hex_desc(1, 2, n, kind=fr.FR_DIFF_MASS, kappa=fr.coef_const(1.0), rho=1.0, dirichlet=False)
    rp, ci, v = fr.hex_assemble_csr(d, c)
    N = len(rp) - 1
    assert v.size / N > 16
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    b = np.random.default_rng(3).standard_normal(N)
    x = np.zeros(N)
    res = fg.solve_amg_cg(A, b, x, fg.SolverConfig(rtol=1e-10, max_iter=500))
    assert res.converged and res.final_residual < 1e-10
    xd = spl.spsolve(_csr(rp, ci, v, (N, N)).tocsc(), b)
    assert np.abs(x - xd).max() < 1e-8 * np.abs(xd).max()
