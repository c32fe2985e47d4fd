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


def test_amg_config1(fg, fr, ctx):
    """Config 1 (256^2 P1 Poisson, row-zeroed Dirichlet): AMG-PCG reaches 1e-10
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
