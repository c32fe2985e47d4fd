"""GPU assembly of simplex bilinear / linear forms (Assembler::assemble_bilinear
/ assemble_linear, assembly/src/assembler.rs:147-333, then CooMatrix::into_csr,
linalg/src/coo.rs:76-) against the oracle's restatement: the same pattern, and
values summed in the same element order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


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


def test_assemble_p1_poisson_256_bitwise(fg, fr, ctx):
    """Config-1 system at 64^2: the P1 stiffness pattern and values equal the
    oracle's bit for bit, the 3-point source load within rounding of sin,
    and after row zeroing PCG-Jacobi gives the oracle's iterations."""
    n = 64
    coords, conn = fr.unit_square_tri(n)
    N = (n + 1) ** 2
    rp, ci, v = fr.simplex_assemble_csr(2, 1, coords, conn, conn, N, fr.FR_DIFFUSION, 1.0, 3)
    A = fg.assemble_bilinear(2, 1, coords, conn, conn, N, fg.DIFFUSION, 1.0, 3, ctx=ctx)
    rpg, cig = A.pattern()
    assert A.nnz() == len(ci)
    assert np.array_equal(rpg, rp) and np.array_equal(cig, ci)
    assert np.array_equal(A.values(), v)
    b = fg.assemble_linear(2, 1, coords, conn, conn, N, fg.SOURCE_SINPROD, (2.0 * np.pi) * np.pi, 3, ctx=ctx)
    bo = fr.p1_source_rhs(n, coords, conn, 3)
    assert _rel(b, bo) < 1e-14
    bnd = fr.unit_square_boundary(n)
    fg.apply_dirichlet(A, b, bnd, np.zeros(len(bnd)))
    rp2, ci2, v2, rhs2, _, _ = fr.poisson_p1(n)
    assert np.array_equal(A.values(), v2)
    assert _rel(b, rhs2) < 1e-14
    x = np.zeros(N)
    res = fg.solve_pcg_jacobi(A, b, x, fg.SolverConfig(rtol=1e-10, max_iter=5000))
    ref = fr.CsrRef(rp2, ci2, v2)
    xo, ro = fr.pcg_jacobi(fr._Op(N, csr=ref), ref.diagonal(), rhs2, np.zeros(N), rtol=1e-10, max_iter=5000)
    assert res.converged and abs(res.iterations - ro.iterations) <= 1
    assert _rel(x, xo) < 1e-10


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
@pytest.mark.parametrize("form", [1, 2])
def test_assemble_simplex(fg, fr, ctx, dim, order, form):
    """P1-P3 triangles and tetrahedra, mass and diffusion with coefficient 1.7:
    same pattern as the oracle, values within 1e-13 (the reference-element
    tables come from two restatements of the bases), the device matrix equals
    the PA operator's action, and the constant-source load sums to the domain
    measure times the constant."""
    if dim == 2:
        coords, conn = fr.unit_square_tri(5)
    else:
        coords, conn = _tet_mesh(2)
    nn = len(coords) // dim
    nd, ed = fr.simplex_dofs(dim, order, nn, conn)
    qo = 2 * order + 1
    rp, ci, v = fr.simplex_assemble_csr(dim, order, coords, conn, ed, nd, form, 1.7, qo)
    A = fg.assemble_bilinear(dim, order, coords, conn, ed, nd, form, 1.7, qo, ctx=ctx)
    rpg, cig = A.pattern()
    assert np.array_equal(rpg, rp) and np.array_equal(cig, ci)
    assert _rel(A.values(), v) < 1e-13
    cls = fg.PAMassOperator if form == 1 else fg.PADiffusionOperator
    op = cls(dim, order, coords, conn, ed, nd, 1.7, qo, ctx=ctx)
    x = np.random.default_rng(dim * 10 + order).standard_normal(nd)
    assert _rel(A.matvec(x), op.matvec(x)) < 1e-12
    b = fg.assemble_linear(dim, order, coords, conn, ed, nd, fg.SOURCE_CONST, 2.5, qo, ctx=ctx)
    # sum_i phi_i = 1, so the load sums to 2.5 times the domain measure, up to
    # the weight-sum error of the reference's high-order rules (~1e-11)
    assert abs(b.sum() - 2.5) < 1e-10
    if form == 1:  # mass with coefficient 1.7: M 1 / 1.7 = the load of f = 1
        assert _rel(A.matvec(np.ones(nd)) * (2.5 / 1.7), b) < 1e-12


def test_assemble_errors(fg, fr, ctx):
    coords, conn = fr.unit_square_tri(2)
    bad = conn.copy()
    bad[0] = 10 ** 6
    with pytest.raises(fg.FemGpuError):
        fg.assemble_bilinear(2, 1, coords, bad, bad, 9, fg.DIFFUSION, ctx=ctx)
    with pytest.raises(fg.FemGpuError):
        fg.assemble_bilinear(2, 1, coords, conn, conn, 9, fg.DIFF_MASS, ctx=ctx)


def test_matrix_market_of_device_matrix(fg, fr, ctx, tmp_path):
    """A device-assembled CsrMatrix written as Matrix Market and read back
    (io/src/matrix_market.rs round trip) is the same matrix, bit for bit."""
    from femgpu import io as fio
    coords, conn = fr.unit_square_tri(8)
    A = fg.assemble_bilinear(2, 1, coords, conn, conn, 81, fg.DIFFUSION, 1.0, 3, ctx=ctx)
    p = tmp_path / "k.mtx"
    fio.write_matrix_market(p, A)
    rp, ci, v, shape = fio.read_matrix_market(p)
    rpg, cig = A.pattern()
    assert shape == (81, 81)
    assert np.array_equal(rp, rpg) and np.array_equal(ci, cig) and np.array_equal(v, A.values())
