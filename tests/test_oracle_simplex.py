"""Pins the simplex PA oracle (partial.rs:85-274) the way the reference's own
tests do (partial.rs:323-412), tightened: PA == assembled to 1e-13, K 1 = 0,
lumped mass sums to the area, linearity; plus P2/P3 DOF-numbering
conformity (continuity across shared edges)."""
import numpy as np
import pytest


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
                    cur = [i, j, k]; path = [vid(*cur)]
                    for ax in pm:
                        cur[ax] += 1; path.append(vid(*cur))
                    conn += path
    return coords, np.array(conn, np.uint32)


def _mesh(fr, dim, order):
    if dim == 2:
        coords, conn = fr.unit_square_tri(4)
    else:
        coords, conn = _tet_mesh(2)
    nd, ed = fr.simplex_dofs(dim, order, len(coords) // dim, conn)
    return coords, conn, nd, ed


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_pa_matches_assembled(fr, dim, order):
    coords, conn, nd, ed = _mesh(fr, dim, order)
    qo = 2 * order + 1
    x = np.random.default_rng(order).standard_normal(nd)
    for kind, coef in ((fr.FR_MASS, 1.0), (fr.FR_DIFFUSION, 2.0)):
        y = fr.simplex_pa_apply(dim, order, coords, conn, ed, kind, coef, qo, x)
        rp, ci, v = fr.simplex_assemble_csr(dim, order, coords, conn, ed, nd, kind, coef, qo)
        A = fr.CsrRef(rp, ci, v)
        assert np.abs(y - A.spmv(x)).max() < 1e-13 * np.abs(y).max()
        Ad = A.dense()
        assert np.abs(Ad - Ad.T).max() < 1e-13 * np.abs(Ad).max()
    K1 = fr.simplex_pa_apply(dim, order, coords, conn, ed, fr.FR_DIFFUSION, 1.0, qo, np.ones(nd))
    assert np.abs(K1).max() < 1e-12
    M1 = fr.simplex_pa_apply(dim, order, coords, conn, ed, fr.FR_MASS, 1.0, qo, np.ones(nd))
    # unit square / unit cube; the reference's 12-point triangle rule (order 7)
    # carries a ~6e-12 weight-sum error (quadrature.rs:444-448)
    tol = 1e-10 if (dim == 2 and qo >= 6) else 1e-12
    assert abs(M1.sum() - 1.0) < tol


def test_reference_pa_tests(fr):
    """partial.rs:323-412 on unit_square_tri(4) / (3), P1."""
    coords, conn = fr.unit_square_tri(4)
    n = 25
    x = np.ones(n)
    ym = fr.simplex_pa_apply(2, 1, coords, conn, conn, fr.FR_MASS, 1.0, 3, x)
    rp, ci, v = fr.simplex_assemble_csr(2, 1, coords, conn, conn, n, fr.FR_MASS, 1.0, 3)
    assert np.linalg.norm(ym - fr.CsrRef(rp, ci, v).spmv(x)) < 1e-11
    lm = fr.simplex_lumped_mass(2, 1, coords, conn, conn, n, 1.0, 3)
    assert abs(lm.sum() - 1.0) < 1e-11
    coords, conn = fr.unit_square_tri(3)
    xs = np.arange(16) / 16
    y1 = fr.simplex_pa_apply(2, 1, coords, conn, conn, fr.FR_MASS, 1.0, 3, xs)
    y2 = fr.simplex_pa_apply(2, 1, coords, conn, conn, fr.FR_MASS, 1.0, 3, xs, y=y1)
    assert np.linalg.norm(y2 - 2 * y1) < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [2, 3])
def test_dof_numbering_conforming(fr, dim, order):
    """A global polynomial of degree <= order interpolated through the DOF
    numbering is reproduced inside every element (so shared DOFs agree)."""
    coords, conn, nd, ed = _mesh(fr, dim, order)
    nl = len(ed) // (len(conn) // (dim + 1))
    X = coords.reshape(-1, dim)
    f = lambda p: 1 + p[0] ** order - 2 * p[1] * p[0] ** (order - 1) + (p[-1] ** 2 if order >= 2 else 0)
    vals = np.full(nd, np.nan)
    from test_oracle_basics import _simplex_nodes
    ref_nodes = _simplex_nodes(dim, order)
    for e in range(len(conn) // (dim + 1)):
        v = X[conn[e * (dim + 1):(e + 1) * (dim + 1)]]
        J = (v[1:] - v[0]).T
        for l in range(nl):
            p = v[0] + J @ ref_nodes[l]
            g = ed[e * nl + l]
            if np.isnan(vals[g]):
                vals[g] = f(p)
            else:
                assert abs(vals[g] - f(p)) < 1e-12  # same physical point -> same DOF
    assert not np.isnan(vals).any()


def _kappa(xq):
    # a coefficient of the physical point, as a ScalarCoeff closure would give it
    return 1.0 + xq[..., 0] + 0.5 * xq[..., 1] ** 2


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_qp_coefficient(fr, dim, order):
    """Coefficients given per quadrature point (the ScalarCoeff of
    partial.rs:137-147, 186-190 evaluated at x_q):
    - a constant array equals the constant coefficient;
    - PA apply equals the assembled matrix with the same point values;
    - P1 linear patch: with kappa = 1 + x, v^T K u = a.b int(1 + x) = 1.5 a.b
      over the unit square / cube (grad u = a, exact for any rule);
    - 1^T M 1 = int rho and the lumped mass sums to it (rho = 1 + x)."""
    coords, conn, nd, ed = _mesh(fr, dim, order)
    qo = 2 * order + 1
    xq = fr.simplex_qp_coords(dim, coords, conn, qo)
    ne, nq = xq.shape[:2]
    x = np.random.default_rng(order).standard_normal(nd)
    for kind in (fr.FR_DIFFUSION, fr.FR_MASS):
        y_c = fr.simplex_pa_apply(dim, order, coords, conn, ed, kind, 2.5, qo, x)
        y_q = fr.simplex_pa_apply(dim, order, coords, conn, ed, kind, 0.0, qo, x, coef_qp=np.full((ne, nq), 2.5))
        assert np.abs(y_c - y_q).max() < 1e-14 * np.abs(y_c).max()
        kq = _kappa(xq)
        y = fr.simplex_pa_apply(dim, order, coords, conn, ed, kind, 0.0, qo, x, coef_qp=kq)
        rp, ci, v = fr.simplex_assemble_csr(dim, order, coords, conn, ed, nd, kind, 0.0, qo, coef_qp=kq)
        ya = fr.CsrRef(rp, ci, v).spmv(x)
        assert np.abs(y - ya).max() < 1e-13 * np.abs(y).max()
    # the physical points: sum_q w |J| x_q integrates x exactly
    c = np.asarray(coords).reshape(-1, dim)
    if order == 1:
        a, b = np.arange(1.0, dim + 1.0), np.linspace(-1.0, 0.5, dim)
        u, w = c @ a, c @ b
        kq = 1.0 + xq[..., 0]
        rp, ci, v = fr.simplex_assemble_csr(dim, 1, coords, conn, ed, nd, fr.FR_DIFFUSION, 0.0, qo, coef_qp=kq)
        assert abs(w @ fr.CsrRef(rp, ci, v).spmv(u) - 1.5 * (a @ b)) < 1e-12
        rp, ci, v = fr.simplex_assemble_csr(dim, 1, coords, conn, ed, nd, fr.FR_MASS, 0.0, qo, coef_qp=kq)
        assert abs(fr.CsrRef(rp, ci, v).spmv(np.ones(nd)).sum() - 1.5) < 1e-13
        lm = fr.simplex_lumped_mass(dim, 1, coords, conn, ed, nd, 0.0, qo, rho_qp=kq)
        assert abs(lm.sum() - 1.5) < 1e-13
