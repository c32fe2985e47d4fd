"""Pins the hex Q_p oracle (no reference counterpart: the reference has no hex
PA, partial.rs:278-288) three ways -- reference-style element loop
(partial.rs:151-208 generalised, assembler.rs:96-123 geometry), sum
factorisation on stored qdata, and brute-force assembled CSR -- and against
what the mathematics fixes: K 1 = 0, symmetry, 1^T M 1 = |Omega| = 1 on the
perturbed mesh, exact integration of polynomials on an affine mesh, and the
discrete maximum of the Dirichlet-masked operator."""
import numpy as np
import pytest

CASES = [(1, 2), (2, 3), (3, 4), (4, 5), (2, 4)]


def _setup(fr, p, q, n=(2, 3, 2), amp=0.2, kind=None, kappa=None, dirichlet=False, rho=1.3):
    kind = fr.FR_DIFF_MASS if kind is None else kind
    kappa = fr.coef_sinprod() if kappa is None else kappa
    d = fr.hex_desc(p, q, n, kind=kind, kappa=kappa, rho=rho, dirichlet=dirichlet)
    c = fr.hex_vertices(*((n,) * 3 if np.isscalar(n) else n), amp, 7)
    return d, c


@pytest.mark.parametrize("p,q", CASES)
@pytest.mark.parametrize("dirichlet", [False, True])
def test_three_way(fr, p, q, dirichlet):
    d, c = _setup(fr, p, q, dirichlet=dirichlet)
    N = fr.hex_ndofs(d)
    x = np.random.default_rng(p).standard_normal(N)
    y1 = fr.hex_apply_refloop(d, c, x)
    Dq, Mq = fr.hex_qdata(d, c)
    y2 = fr.hex_apply_sumfac(d, Dq, Mq, x, nthreads=2)
    A = fr.CsrRef(*fr.hex_assemble_csr(d, c))
    y3 = A.spmv(x)
    s = np.abs(y1).max()
    assert np.abs(y1 - y2).max() < 1e-13 * s
    assert np.abs(y1 - y3).max() < 1e-13 * s
    Ad = A.dense()
    assert np.abs(Ad - Ad.T).max() < 1e-13 * np.abs(Ad).max()
    assert np.allclose(fr.hex_diagonal(d, Dq, Mq), np.diag(Ad), rtol=1e-13, atol=1e-15)
    # accumulate semantics (MatFreeOperator::apply adds into y)
    y0 = np.random.default_rng(9).standard_normal(N)
    assert np.allclose(fr.hex_apply_sumfac(d, Dq, Mq, x, y=y0), y0 + y2, rtol=0, atol=1e-13 * s)


@pytest.mark.parametrize("p,q", CASES)
def test_invariants(fr, p, q):
    dk, c = _setup(fr, p, q, kind=fr.FR_DIFFUSION)
    N = fr.hex_ndofs(dk)
    Dq, Mq = fr.hex_qdata(dk, c)
    assert np.abs(fr.hex_apply_sumfac(dk, Dq, Mq, np.ones(N))).max() < 1e-12  # K 1 = 0
    dm, _ = _setup(fr, p, q, kind=fr.FR_MASS, rho=1.0)
    Dq, Mq = fr.hex_qdata(dm, c)
    vol = fr.hex_apply_sumfac(dm, Dq, Mq, np.ones(N)).sum()
    assert abs(vol - 1.0) < 1e-13  # perturbed interior, planar boundary: volume 1
    assert abs(fr.hex_rhs_unit(dm, c).sum() - 1.0) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_affine_polynomial_exactness(fr, p):
    """On the unperturbed (affine) mesh with q = p+1, u^T M v = int u v and
    u^T K v = int grad u . grad v exactly for u, v in Q_p."""
    n = (2, 2, 2)
    q = p + 1
    dm = fr.hex_desc(p, q, n, kind=fr.FR_MASS, rho=1.0)
    dk = fr.hex_desc(p, q, n, kind=fr.FR_DIFFUSION)
    c = fr.hex_vertices(2, 2, 2, 0.0, 0)
    Nd = 2 * p + 1
    g = np.arange(Nd) / (Nd - 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    X, Y, Z = X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(), Z.transpose(2, 1, 0).ravel()
    a = min(p, 2)
    u = X ** a * Y  # in Q_p for p >= 1
    v = Y ** min(p, 2) + Z
    Dm, Mm = fr.hex_qdata(dm, c)
    Dk, Mk = fr.hex_qdata(dk, c)
    uMv = u @ fr.hex_apply_sumfac(dm, Dm, Mm, v)
    b2 = min(p, 2)
    # int_0^1 x^a dx * int y (y^b2) dy + int x^a dx int y dy int z dz
    exact_m = (1 / (a + 1)) * (1 / (b2 + 2)) + (1 / (a + 1)) * 0.5 * 0.5
    assert abs(uMv - exact_m) < 1e-13
    uKv = u @ fr.hex_apply_sumfac(dk, Dk, Mk, v)
    # grad u = (a x^(a-1) y, x^a, 0); grad v = (0, b2 y^(b2-1), 1)
    exact_k = (1 / (a + 1)) * b2 * (1 / b2) if b2 > 0 else 0.0  # int x^a * b2 y^(b2-1)
    assert abs(uKv - exact_k) < 1e-12


def test_golden_fixtures(fr):
    """Committed fixtures (tests/golden/hex_*.npz, written by
    tests/golden/make_golden.py) reproduce exactly."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "hex_fixtures.npz")
    z = np.load(path)
    for p in (1, 2, 3, 4):
        for amp in (0.0, 0.2):
            key = f"q{p}_{'pert' if amp else 'affine'}"
            d, c = _setup(fr, p, p + 1, n=(2, 2, 2), amp=amp)
            assert np.array_equal(c, z[key + "_coords"])
            y = fr.hex_apply_refloop(d, c, z[key + "_x"])
            assert np.allclose(y, z[key + "_y"], rtol=0, atol=1e-13 * np.abs(y).max())


def test_vertices_perturbation(fr):
    c = fr.hex_vertices(4, 4, 4, 0.2, 3).reshape(-1, 3)
    base = fr.hex_vertices(4, 4, 4, 0.0, 0).reshape(-1, 3)
    dev = np.abs(c - base).max(1)
    assert dev.max() <= 0.2 * 0.25 + 1e-15 and dev.max() > 0.1 * 0.25
    idx = np.arange(125).reshape(5, 5, 5)
    bnd = np.zeros((5, 5, 5), bool)
    bnd[0], bnd[-1], bnd[:, 0], bnd[:, -1], bnd[:, :, 0], bnd[:, :, -1] = (True,) * 6
    assert np.all(dev[idx[bnd]] == 0)


def test_grf_statistics(fr):
    """kappa = exp(g): g has zero mean and variance sigma^2, covariance
    sigma^2 exp(-r^2 / (2 l^2)) (checked on samples)."""
    coef = fr.coef_grf(5, n_features=512, sigma=1.5, length=0.1, lo=1e-30, hi=1e30)
    rng = np.random.default_rng(0)
    gs, gs2 = [], []
    for s in range(40):
        coef.seed = s
        x = rng.random(3)
        y = x + np.array([0.1, 0.0, 0.0])
        gs.append(np.log(fr.coef_eval(coef, x)))
        gs2.append(np.log(fr.coef_eval(coef, y)))
    gs, gs2 = np.array(gs), np.array(gs2)
    assert abs(gs.mean()) < 0.8 and 1.0 < gs.std() < 2.2
    corr = np.mean(gs * gs2) / 1.5 ** 2
    assert 0.2 < corr < 1.0  # exp(-0.5) = 0.61 in expectation
    coef = fr.coef_grf(1)
    for _ in range(20):
        k = fr.coef_eval(coef, rng.random(3))
        assert 1e-2 <= k <= 1e2


@pytest.mark.parametrize("p,q", [(1, 2), (2, 3), (3, 4), (4, 5), (2, 4)])
def test_linear_patch_perturbed(fr, p, q):
    """Patch test on the perturbed (non-affine) mesh, fixed by the mathematics
    alone.  For linear u = a.x + c and v = b.x + d (exactly in Q_p on a
    trilinear mesh) and kappa = rho = 1:
      v^T K u = int grad u . grad v = a.b |Omega| = a.b (the boundary is the
        planar unit cube; the integrand a.b |J| has degree <= 2 per reference
        direction, exact for q >= 2),
      (K u)_i = int a . grad phi_i = 0 at every interior DOF,
      v^T M u = int (a.x + c)(b.x + d) (degree <= 4 per direction, q >= 3).
    A transposed metric (J^-T J^-1 for J^-1 J^-T, the defect of the reference's
    FemCeed, reed/context.rs:282-284) or a wrong |J| fails the first two."""
    import hexgeom
    n = (3, 2, 3)
    c = fr.hex_vertices(*n, 0.2, 13)
    X = hexgeom.dof_coords(*n, p, c)
    rng = np.random.default_rng(100 + p)
    a, b = rng.standard_normal(3), rng.standard_normal(3)
    c0, d0 = rng.standard_normal(2)
    u, v = X @ a + c0, X @ b + d0
    dk = fr.hex_desc(p, q, n, kind=fr.FR_DIFFUSION, kappa=fr.coef_const(1.0))
    Dk, Mk = fr.hex_qdata(dk, c)
    Ku = fr.hex_apply_sumfac(dk, Dk, Mk, u)
    assert abs(v @ Ku - a @ b) < 1e-12 * (1 + abs(a @ b))
    inner = ~hexgeom.boundary_mask(*n, p)
    assert np.abs(Ku[inner]).max() < 1e-12
    assert np.allclose(fr.hex_apply_refloop(dk, c, u), Ku, rtol=0, atol=1e-13)
    if q >= 3:
        dm = fr.hex_desc(p, q, n, kind=fr.FR_MASS, rho=1.0)
        Dm, Mm = fr.hex_qdata(dm, c)
        exact = hexgeom.int_linear_product(a, c0, b, d0)
        assert abs(v @ fr.hex_apply_sumfac(dm, Dm, Mm, u) - exact) < 1e-13 * (1 + abs(exact))
        dd = fr.hex_desc(p, q, n, kind=fr.FR_DIFF_MASS, kappa=fr.coef_const(1.0), rho=1.0)
        assert abs(v @ fr.hex_apply_refloop(dd, c, u) - (a @ b + exact)) < 1e-12 * (1 + abs(exact))


def test_patch_test_detects_transposed_metric(fr):
    """The patch test above is sensitive to the metric's orientation: with the
    qdata's J^-1 J^-T replaced by J^-T J^-1 (built here from the same
    Jacobians), v^T K u moves away from a.b on the perturbed mesh."""
    import hexgeom
    p, q, n = 2, 3, (2, 2, 2)
    c = fr.hex_vertices(*n, 0.2, 13)
    X = hexgeom.dof_coords(*n, p, c)
    a, b = np.array([1.0, 0.3, -0.7]), np.array([0.2, -1.1, 0.5])
    u, v = X @ a, X @ b
    # transposed metric on every Gauss point, from the trilinear Jacobian
    xg, wg = fr.gauss_legendre(q)
    V = c.reshape(n[2] + 1, n[1] + 1, n[0] + 1, 3)
    Nx = n[0] * p + 1
    Ny = n[1] * p + 1
    B, G = fr.lagrange_1d(p, xg)
    B, G = B.reshape(q, p + 1), G.reshape(q, p + 1)
    # trilinear map x = sum_v X_v prod_d (1 + s_d xi_d) / 8, J_ij = dx_i / dxi_j
    total_ok, total_bad = 0.0, 0.0
    for ez in range(n[2]):
        for ey in range(n[1]):
            for ex in range(n[0]):
                idx = np.array([[[ (ex * p + i) + Nx * ((ey * p + j) + Ny * (ez * p + k))
                                   for i in range(p + 1)] for j in range(p + 1)] for k in range(p + 1)])
                ue, ve = u[idx], v[idx]  # [k][j][i]
                for cz in range(q):
                    for cy in range(q):
                        for cx in range(q):
                            xi = np.array([xg[cx], xg[cy], xg[cz]])
                            J = np.zeros((3, 3))
                            for aa in (0, 1):
                                for bb in (0, 1):
                                    for cc in (0, 1):
                                        w = [(1 + (2 * aa - 1) * xi[0]), (1 + (2 * bb - 1) * xi[1]),
                                             (1 + (2 * cc - 1) * xi[2])]
                                        dw = np.array([(2 * aa - 1) * w[1] * w[2], (2 * bb - 1) * w[0] * w[2],
                                                       (2 * cc - 1) * w[0] * w[1]]) / 8.0
                                        J += np.outer(V[ez + cc, ey + bb, ex + aa], dw)
                            gu = np.array([np.einsum("k,j,i,kji", B[cz], B[cy], G[cx], ue),
                                           np.einsum("k,j,i,kji", B[cz], G[cy], B[cx], ue),
                                           np.einsum("k,j,i,kji", G[cz], B[cy], B[cx], ue)])
                            gv = np.array([np.einsum("k,j,i,kji", B[cz], B[cy], G[cx], ve),
                                           np.einsum("k,j,i,kji", B[cz], G[cy], B[cx], ve),
                                           np.einsum("k,j,i,kji", G[cz], B[cy], B[cx], ve)])
                            Ji = np.linalg.inv(J)
                            wq = wg[cx] * wg[cy] * wg[cz] * abs(np.linalg.det(J))
                            total_ok += wq * gv @ (Ji @ Ji.T) @ gu
                            total_bad += wq * gv @ (Ji.T @ Ji) @ gu
    assert abs(total_ok - a @ b) < 1e-12
    assert abs(total_bad - a @ b) > 1e-4


@pytest.mark.parametrize("dirichlet", [False, True])
def test_refloop_threaded_matches_serial(fr, dirichlet):
    """The parity-coloured threaded reference loop (the CPU baseline's
    threaded form) is the serial loop's operator up to summation order, on
    odd and even element counts."""
    d, c = _setup(fr, 2, 3, n=(3, 4, 5), dirichlet=dirichlet)
    x = np.random.default_rng(4).standard_normal(fr.hex_ndofs(d))
    y1 = fr.hex_apply_refloop(d, c, x)
    for nt in (1, 3):
        y2 = fr.hex_apply_refloop(d, c, x, nthreads=nt)
        assert np.abs(y1 - y2).max() < 1e-13 * np.abs(y1).max()
