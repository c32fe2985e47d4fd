"""Pins the oracle's CSR, config-1 assembly and solvers against the
reference's own known-answer tests and closed forms."""
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def lap1d(fr, n):
    rp = [0]; ci = []; v = []
    for i in range(n):
        if i > 0: ci.append(i - 1); v.append(-1.0)
        ci.append(i); v.append(2.0)
        if i < n - 1: ci.append(i + 1); v.append(-1.0)
        rp.append(len(ci))
    return fr.CsrRef(np.array(rp), np.array(ci), np.array(v))


def test_csr_known_answers(fr):
    A = lap1d(fr, 3)
    assert A.spmv(np.array([1.0, 2.0, 3.0])).tolist() == [0.0, 0.0, 4.0]  # csr.rs:589-598
    assert A.diagonal().tolist() == [2.0, 2.0, 2.0]  # csr.rs:601-605
    v, rhs = fr.dirichlet_row_zeroing(A.row_ptr, A.col_idx, A.values, [0], np.array([1.0, 2, 3]),
                                      np.array([5.0]))
    B = fr.CsrRef(A.row_ptr, A.col_idx, v).dense()
    assert B[0].tolist() == [1.0, 0.0, 0.0] and rhs[0] == 5.0  # csr.rs:608-615


def test_spmv_4way_order_bitwise(fr):
    """The f64 row sum groups four products before adding (csr.rs:38-64)."""
    rng = np.random.default_rng(0)
    n = 50
    rows = [np.sort(rng.choice(n, size=rng.integers(1, 12), replace=False)) for _ in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    ci = np.concatenate(rows).astype(np.uint32)
    v = rng.standard_normal(len(ci)) * 10.0 ** rng.integers(-8, 8, len(ci))
    x = rng.standard_normal(n)
    y = fr.CsrRef(rp, ci, v).spmv(x)
    for r in range(n):
        s, e = int(rp[r]), int(rp[r + 1])
        acc = 0.0
        k = s
        while k + 4 <= e:
            t = v[k] * x[ci[k]] + v[k + 1] * x[ci[k + 1]]
            t = t + v[k + 2] * x[ci[k + 2]]
            t = t + v[k + 3] * x[ci[k + 3]]
            acc = acc + t
            k += 4
        while k < e:
            acc = acc + v[k] * x[ci[k]]
            k += 1
        assert y[r] == acc


def test_p1_poisson_256_matrix(fr):
    n = 256
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(n)
    assert len(rp) - 1 == 66049 and len(v) == 460289
    # before BCs: symmetric, zero row sums, interior rows are the 5-point stencil
    coords, conn = fr.unit_square_tri(16)
    N = 17 * 17
    rp, ci, v = fr.simplex_assemble_csr(2, 1, coords, conn, conn, N, fr.FR_DIFFUSION, 1.0, 3)
    A = fr.CsrRef(rp, ci, v).dense()
    assert np.abs(A - A.T).max() < 1e-14
    assert np.abs(A.sum(1)).max() < 1e-13
    i = 8 * 17 + 8
    row = A[i]
    assert abs(row[i] - 4) < 1e-14
    for j in (i - 1, i + 1, i - 17, i + 17):
        assert abs(row[j] + 1) < 1e-14
    assert np.abs(np.delete(row, [i, i - 1, i + 1, i - 17, i + 17])).max() < 1e-14
    # 1^T M 1 = area (mass.rs:66-82)
    rp, ci, v = fr.simplex_assemble_csr(2, 1, coords, conn, conn, N, fr.FR_MASS, 1.0, 3)
    assert abs(v.sum() - 1.0) < 1e-13


def _l2_error(fr, n, u, coords, conn):
    pts, wts = fr.tri_rule(5)
    err = 0.0
    X = coords.reshape(-1, 2)
    for e in range(len(conn) // 3):
        nd = conn[3 * e:3 * e + 3]
        x0, x1, x2 = X[nd[0]], X[nd[1]], X[nd[2]]
        J = np.column_stack([x1 - x0, x2 - x0])
        det = abs(np.linalg.det(J))
        for (xi, eta), w in zip(pts, wts):
            phi = np.array([1 - xi - eta, xi, eta])
            xp = x0 + J @ np.array([xi, eta])
            d = phi @ u[nd] - math.sin(math.pi * xp[0]) * math.sin(math.pi * xp[1])
            err += w * det * d * d
    return math.sqrt(err)


def test_ex1_poisson_small(fr):
    """examples/mfem_ex1_poisson.rs:184-237: n = 8 converges, residual < 1e-8,
    L2 error < 2.2e-2; observed order > 1.5."""
    errs = []
    for n in (8, 16):
        rp, ci, v, rhs, coords, conn = fr.poisson_p1(n)
        A = fr.CsrRef(rp, ci, v)
        x, res = fr.pcg_jacobi(fr._Op(len(rhs), csr=A), A.diagonal(), rhs, np.zeros(len(rhs)),
                               rtol=1e-10, max_iter=5000)
        assert res.converged and res.final_residual < 1e-8
        errs.append(_l2_error(fr, n, x, coords, conn))
    assert errs[0] < 2.2e-2
    assert math.log2(errs[0] / errs[1]) > 1.5
    z = np.load(os.path.join(GOLD, "csr_fixtures.npz"))
    assert np.array_equal(z["p8_v"], fr.poisson_p1(8)[2])


def test_cg_and_pcg_known_answers(fr):
    A = lap1d(fr, 50)
    b = np.ones(50)
    x, res = fr.solve_cg_operator(fr._Op(50, csr=A), b, np.zeros(50))
    assert res.converged and np.linalg.norm(A.spmv(x) - b) < 1e-6  # lib.rs:1197-1210
    assert res.iterations <= 25 + 1  # exact arithmetic: n/2 by symmetry of b = 1
    exact = np.array([(i + 1) * (50 - i) / 2 for i in range(50)])  # (-u'' = 1 discrete)
    assert np.abs(x - exact).max() < 1e-8 * exact.max()
    x, res = fr.pcg_jacobi(fr._Op(50, csr=A), A.diagonal(), b, np.zeros(50))
    assert res.converged and res.iterations < 60  # lib.rs:1212-1221
    # failure paths
    x, res = fr.solve_cg_operator(fr._Op(50, csr=A), b, np.zeros(50), max_iter=3)
    assert res.status == 3 and not res.converged
    Z = fr.CsrRef(np.arange(51), np.arange(50), np.zeros(50))
    x, res = fr.solve_cg_operator(fr._Op(50, csr=Z), b, np.zeros(50))
    assert res.status == 4  # breakdown


def test_operator_cg_matches_csr_cg(fr):
    """crates/solver/tests/operator_cg.rs:73-97: operator CG vs CSR CG < 1e-8."""
    rp, ci, v, rhs, _, _ = fr.poisson_p1(16)
    A = fr.CsrRef(rp, ci, v)
    x1, _ = fr.solve_cg_operator(fr._Op(len(rhs), csr=A), rhs, np.zeros(len(rhs)), rtol=1e-12)
    x2, _ = fr.solve_cg_operator(fr._Op(len(rhs), fn=A.spmv), rhs, np.zeros(len(rhs)), rtol=1e-12)
    assert np.abs(x1 - x2).max() < 1e-8


def test_spectral_radius_and_chebyshev(fr):
    A = lap1d(fr, 50)
    rho = fr.estimate_spectral_radius(A, 30)
    assert 3.5 < rho < 4.1  # amg/src/lib.rs:278-286
    lam_max = 2 - 2 * math.cos(50 * math.pi / 51)
    assert rho <= lam_max + 1e-12  # power iteration from below
    # v0 = 1/sqrt(n) is symmetric, so it has no component on the (antisymmetric)
    # top eigenvector k = 50: the iteration converges to k = 49
    lam49 = 2 - 2 * math.cos(49 * math.pi / 51)
    assert abs(fr.estimate_spectral_radius(A, 3000) - lam49) < 1e-6
    rho = fr.estimate_spectral_radius(A, 20)
    x = fr.chebyshev_smooth_ref(A, np.zeros(50), np.ones(50), rho * 1.1 / 30, rho * 1.1, 20)
    r = np.ones(50) - A.spmv(x)
    assert np.linalg.norm(r) / math.sqrt(50) < 0.9  # amg/src/lib.rs:254-275
    z = np.load(os.path.join(GOLD, "csr_fixtures.npz"))
    assert np.array_equal(x, z["cheb_x20"])


def test_chebyshev_reference_recurrence_by_hand(fr):
    """Two steps of amg/src/lib.rs:53-103 written out on a 2x2 matrix."""
    A = fr.CsrRef(np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([4.0, 1.0, 1.0, 3.0]))
    b = np.array([1.0, 2.0])
    lmin, lmax = 0.5, 5.0
    alpha, delta = (lmax + lmin) / 2, (lmax - lmin) / 2
    dinv = np.array([0.25, 1 / 3])
    x = np.zeros(2)
    r = b - A.dense() @ x
    d = (2 / alpha) * dinv * r
    x = x + d
    r = b - A.dense() @ x
    rho = 1 / (1 - (delta / (2 * alpha)) ** 2 * 1.0)
    d = rho * ((2 / alpha) * dinv * r + (rho - 1) * d)
    x = x + d
    assert np.allclose(fr.chebyshev_smooth_ref(A, np.zeros(2), b, lmin, lmax, 2), x, rtol=1e-15)


def test_textbook_chebyshev_polynomial(fr):
    """On A = diag(lam) (D = I) the preconditioner is z = p(A) r with
    1 - lam p(lam) = T_m((theta - lam)/delta) / T_m(theta/delta)."""
    lam = np.linspace(0.05, 2.0, 40)
    n = len(lam)
    op = fr._Op(n, fn=lambda v: lam * v)
    lmin, lmax, m = 0.1, 2.0, 4
    z = fr.chebyshev_jacobi_apply(op, np.ones(n), m, lmin, lmax, np.ones(n))
    theta, delta = (lmax + lmin) / 2, (lmax - lmin) / 2
    T = lambda t: np.cosh(m * np.arccosh(t)) if np.all(np.abs(t) >= 1) else np.cos(m * np.arccos(np.clip(t, -1, 1)))

    def Tm(t):
        out = np.empty_like(t)
        big = np.abs(t) > 1
        out[~big] = np.cos(m * np.arccos(t[~big]))
        out[big] = np.sign(t[big]) ** m * np.cosh(m * np.arccosh(np.abs(t[big])))
        return out
    expect = Tm((theta - lam) / delta) / Tm(np.array([theta / delta]))[0]
    assert np.allclose(1 - lam * z, expect, atol=1e-13)
    assert np.all(z > 0)  # positive preconditioner below lmax


def test_pcg_chebyshev_hex_matches_direct(fr):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    n, p = 6, 2
    d = fr.hex_desc(p, p + 1, n, kind=fr.FR_DIFFUSION, kappa=fr.coef_sinprod(), dirichlet=True)
    c = fr.hex_vertices(n, n, n, 0.1, 2)
    Dq, Mq = fr.hex_qdata(d, c)
    op = fr._Op(fr.hex_ndofs(d), hexop=(d, Dq, Mq, 2))
    diag = fr.hex_diagonal(d, Dq, Mq)
    b = fr.hex_rhs_unit(d, c)
    lmax = 1.1 * fr.estimate_lmax_jacobi(op, diag, 30)
    x, res = fr.pcg_chebyshev(op, diag, 4, lmax / 30, lmax, b, np.zeros(len(b)), rtol=1e-13)
    assert res.converged
    rp, ci, v = fr.hex_assemble_csr(d, c)
    A = sp.csr_matrix((v, ci.astype(np.int64), rp.astype(np.int64)), shape=(len(b), len(b)))
    xe = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-11
    xj, rj = fr.pcg_jacobi(op, diag, b, np.zeros(len(b)), rtol=1e-13, max_iter=5000)
    assert rj.converged and rj.iterations > res.iterations  # the polynomial preconditioner helps
