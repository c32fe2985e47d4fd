"""GPU parity: hex sum-factorised PA and simplex PA through the C ABI against
the oracle.  Bar: 1e-12 relative (max-norm) in FP64, 1e-5 in FP32."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HEX_CASES = [(1, 2), (2, 3), (3, 4), (4, 5), (1, 3), (2, 4), (3, 5), (4, 6), (5, 6)]


def _hex(fg, fr, ctx, p, q, n, form, kappa="sinprod", rho=1.3, dirichlet=False, amp=0.2, seed=3,
         dtype=None):
    dtype = fg.F64 if dtype is None else dtype
    nx, ny, nz = n
    c = fg.hex_vertices(nx, ny, nz, amp, seed)
    assert np.array_equal(c, fr.hex_vertices(nx, ny, nz, amp, seed))  # same generator
    if kappa == "sinprod":
        kg, kr = "sinprod", fr.coef_sinprod()
    elif isinstance(kappa, dict):
        kg, kr = kappa, fr.coef_grf(kappa["seed"])
    else:
        kg, kr = float(kappa), fr.coef_const(kappa)
    op = fg.HexPAOperator(p, q, n, form=form, kappa=kg, rho=rho, dirichlet=dirichlet, dtype=dtype,
                          vertices=c, ctx=ctx)
    d = fr.hex_desc(p, q, n, kind=form, kappa=kr, rho=rho, dirichlet=dirichlet)
    Dq, Mq = fr.hex_qdata(d, c)
    return op, d, c, Dq, Mq


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _iters_close(i, j):
    """Iteration counts of two solves whose operators differ only in rounding
    (1e-13): the stopping test can fall a few iterations apart when the
    residual curve is flat near the tolerance, so the bar is 3 iterations or
    2%; the solutions themselves are compared to 1e-8."""
    return abs(i - j) <= max(3, 0.02 * max(i, j))


@pytest.fixture(params=["default", "general"])
def qdata(request, monkeypatch):
    """Runs a test with the operators' default quadrature data (the geometry
    formed on the fly on these perturbed meshes) and with the stored general
    form (FEMGPU_HEX_QDATA=general)."""
    if request.param == "general":
        monkeypatch.setenv("FEMGPU_HEX_QDATA", "general")
    return request.param


@pytest.mark.parametrize("p,q", HEX_CASES)
@pytest.mark.parametrize("form", [1, 2, 3])
def test_hex_apply_f64(fg, fr, ctx, qdata, p, q, form):
    n = (3, 2, 4)
    op, d, c, Dq, Mq = _hex(fg, fr, ctx, p, q, n, form)
    assert op.qdata_mode() == ("geometry" if qdata == "default" and form != 1 else "general")
    N = op.n_dofs()
    assert N == fr.hex_ndofs(d)
    x = np.random.default_rng(p * 10 + q).standard_normal(N)
    y = np.zeros(N)
    op.apply_zero(x, y)
    ref = fr.hex_apply_sumfac(d, Dq, Mq, x)
    assert _rel(y, ref) < 1e-12
    # MatFreeOperator::apply accumulates
    y0 = np.random.default_rng(5).standard_normal(N)
    y1 = y0.copy()
    op.apply(x, y1)
    assert np.abs(y1 - (y0 + ref)).max() < 1e-12 * np.abs(ref).max() + 1e-12 * np.abs(y0).max()


# meshes spanning several workgroup tiles in x and y (with partial edge tiles)
# and several z chunks, so every halo / spill / fix-up path runs
MULTI_TILE = [(1, 2, (17, 9, 5)), (2, 3, (16, 9, 7)), (2, 3, (15, 5, 9)), (3, 4, (9, 6, 5)),
              (4, 5, (11, 5, 4)), (5, 6, (8, 3, 3)), (1, 3, (9, 17, 4))]


@pytest.mark.parametrize("p,q,n", MULTI_TILE)
@pytest.mark.parametrize("form", [1, 2, 3])
@pytest.mark.parametrize("dirichlet", [False, True])
def test_hex_apply_multi_tile(fg, fr, ctx, qdata, p, q, n, form, dirichlet):
    op, d, c, Dq, Mq = _hex(fg, fr, ctx, p, q, n, form, dirichlet=dirichlet)
    N = op.n_dofs()
    x = np.random.default_rng(p + 7 * form).standard_normal(N)
    ref = fr.hex_apply_sumfac(d, Dq, Mq, x)
    y = np.full(N, np.nan)
    op.apply_zero(x, y)
    assert _rel(y, ref) < 1e-12
    y0 = np.random.default_rng(3).standard_normal(N)
    y1 = y0.copy()
    op.apply(x, y1)
    assert np.abs(y1 - (y0 + ref)).max() < 1e-12 * (np.abs(ref).max() + np.abs(y0).max())


@pytest.mark.parametrize("p,q", [(2, 3), (3, 4), (4, 5)])
def test_hex_dirichlet_diag_rhs(fg, fr, ctx, qdata, p, q):
    n = (4, 3, 3)
    op, d, c, Dq, Mq = _hex(fg, fr, ctx, p, q, n, 3, dirichlet=True)
    x = np.random.default_rng(1).standard_normal(op.n_dofs())
    assert _rel(op.matvec(x), fr.hex_apply_sumfac(d, Dq, Mq, x)) < 1e-12
    assert _rel(op.diagonal(), fr.hex_diagonal(d, Dq, Mq)) < 1e-12
    assert _rel(op.rhs_unit(), fr.hex_rhs_unit(d, c)) < 1e-12
    # reference element loop (partial.rs:151-208 generalised) agrees too
    assert _rel(op.matvec(x), fr.hex_apply_refloop(d, c, x)) < 1e-12


def test_hex_grf_and_qp_array(fg, fr, ctx, qdata):
    n = (3, 3, 3)
    op, d, c, Dq, Mq = _hex(fg, fr, ctx, 3, 4, n, 2, kappa=dict(kind="grf", seed=11))
    x = np.random.default_rng(2).standard_normal(op.n_dofs())
    assert _rel(op.matvec(x), fr.hex_apply_sumfac(d, Dq, Mq, x)) < 1e-12
    # QP_ARRAY: per-point kappa values (e, qp) from the oracle's sin-product
    E, Q3 = 27, 64
    kq = np.random.default_rng(4).random((E, Q3)) + 0.5
    op2 = fg.HexPAOperator(3, 4, n, form=2, kappa=kq.ravel(), vertices=c, ctx=ctx)
    dk = fr.hex_desc(3, 4, n, kind=2, kappa=fr.coef_const(1.0))
    D1, M1 = fr.hex_qdata(dk, c)
    D1 = D1 * kq[:, :, None]
    assert _rel(op2.matvec(x), fr.hex_apply_sumfac(dk, D1, M1, x)) < 1e-12


@pytest.mark.parametrize("p,q", [(2, 3), (3, 4), (4, 5)])
def test_hex_f32(fg, fr, ctx, qdata, p, q):
    n = (4, 4, 3)
    op, d, c, Dq, Mq = _hex(fg, fr, ctx, p, q, n, 3, dtype=fg.F32)
    x = np.random.default_rng(7).standard_normal(op.n_dofs()).astype(np.float32)
    y = np.zeros_like(x)
    op.apply_zero(x, y)
    ref64 = fr.hex_apply_sumfac(d, Dq, Mq, x.astype(np.float64))
    assert _rel(y.astype(np.float64), ref64) < 1e-5
    ref32 = fr.hex_apply_sumfac_f32(d, Dq, Mq, x)
    assert _rel(y, ref32) < 1e-5


def test_hex_full_size_properties(fg, ctx):
    """Q2 on 64^3 (2.1M DOFs, the config-2 operator): K 1 = 0, 1^T M 1 = 1,
    symmetry x^T A y = y^T A x, linearity."""
    n = 64
    K = fg.HexPAOperator(2, 3, n, form=fg.DIFFUSION, kappa="sinprod", ctx=ctx)
    M = fg.HexPAOperator(2, 3, n, form=fg.MASS, rho=1.0, ctx=ctx)
    N = K.n_dofs()
    assert N == 129 ** 3
    ones = np.ones(N)
    assert np.abs(K.matvec(ones)).max() < 1e-12
    assert abs(M.matvec(ones).sum() - 1.0) < 1e-12
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(N), rng.standard_normal(N)
    Kx, Ky = K.matvec(x), K.matvec(y)
    assert abs(x @ Ky - y @ Kx) < 1e-12 * abs(x @ Ky)
    assert _rel(K.matvec(2.0 * x - 3.0 * y), 2.0 * Kx - 3.0 * Ky) < 1e-12


# ---------------------------------------------------------------- simplex
def _tri_mesh(fr, n, order):
    coords, conn = fr.unit_square_tri(n)
    nd, ed = fr.simplex_dofs(2, order, len(coords) // 2, conn)
    return coords, conn, nd, ed


def _tet_mesh(n):
    """Structured unit cube split into 6 tets per cube (Kuhn split)."""
    g = np.linspace(0, 1, n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).ravel()
    vid = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    conn = []
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for pm in perms:
                    cur = [i, j, k]
                    path = [vid(*cur)]
                    for ax in pm:
                        cur[ax] += 1
                        path.append(vid(*cur))
                    conn += path
    return coords, np.array(conn, np.uint32)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("kind", [1, 2])
def test_simplex_pa(fg, fr, ctx, dim, order, kind):
    if dim == 2:
        coords, conn, nd, ed = _tri_mesh(fr, 6, order)
    else:
        coords, conn = _tet_mesh(3)
        nd, ed = fr.simplex_dofs(3, order, len(coords) // 3, conn)
    qo = 2 * order + 1
    cls = fg.PAMassOperator if kind == 1 else fg.PADiffusionOperator
    op = cls(dim, order, coords, conn, ed, nd, 1.7, qo, ctx=ctx)
    x = np.random.default_rng(order).standard_normal(nd)
    ref = fr.simplex_pa_apply(dim, order, coords, conn, ed, kind, 1.7, qo, x)
    assert _rel(op.matvec(x), ref) < 1e-12
    # matches the assembled matrix (the reference's own PA test, partial.rs:323-366)
    rp, ci, v = fr.simplex_assemble_csr(dim, order, coords, conn, ed, nd, kind, 1.7, qo)
    assert _rel(op.matvec(x), fr.CsrRef(rp, ci, v).spmv(x)) < 1e-12
    assert _rel(op.diagonal(), fr.CsrRef(rp, ci, v).diagonal()) < 1e-12


@pytest.mark.parametrize("renumber", [False, True])
def test_simplex_pa_window_and_direct(fg, fr, ctx, renumber):
    """k_pa_simplex sums a workgroup's element values in an LDS window when its
    DOF indices span at most 4096 entries, else adds them to y directly. A 100^2
    P1 triangle mesh has 20,000 elements: 79 workgroups of 256, the last one
    ragged. Its natural numbering puts every workgroup in the window. Randomly
    renumbered DOFs (10,201 of them) put none there. Both forms, zeroing and
    accumulating applies."""
    coords, conn, nd, ed = _tri_mesh(fr, 100, 1)
    if renumber:
        ed = np.random.default_rng(3).permutation(nd).astype(np.uint32)[ed]
    rng = np.random.default_rng(4)
    for kind, cls in ((1, fg.PAMassOperator), (2, fg.PADiffusionOperator)):
        op = cls(2, 1, coords, conn, ed, nd, 1.7, 3, ctx=ctx)
        x = rng.standard_normal(nd)
        y0 = rng.standard_normal(nd)
        ref = fr.simplex_pa_apply(2, 1, coords, conn, ed, kind, 1.7, 3, x)
        assert _rel(op.matvec(x), ref) < 1e-12
        y = y0.copy()
        op.apply(x, y)
        assert _rel(y, fr.simplex_pa_apply(2, 1, coords, conn, ed, kind, 1.7, 3, x, y0)) < 1e-12


def test_lumped_mass(fg, fr, ctx):
    coords, conn, nd, ed = _tri_mesh(fr, 4, 1)
    lm = fg.LumpedMassOperator(2, 1, coords, conn, ed, nd, 1.0, 3, ctx=ctx)
    assert abs(lm.diag.sum() - 1.0) < 1e-11  # partial.rs:370-378
    ref = fr.simplex_lumped_mass(2, 1, coords, conn, ed, nd, 1.0, 3)
    assert _rel(lm.diag, ref) < 1e-14
    x = np.random.default_rng(0).standard_normal(nd)
    y = np.zeros(nd)
    lm.apply_inverse(x, y)
    assert _rel(y, x / ref) < 1e-14


def test_essential_bc_operator(fg, fr, ctx):
    """The matrix-free essential-BC operator (examples/src/maxwell.rs:142-197)
    around a P2 triangle PA diffusion operator: apply = A applied to x with the
    constrained entries zeroed, identity rows there (also accumulating);
    diagonal 1 there; CG on it solves the free block of the assembled matrix
    with zero constrained values."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    coords, conn, nd, ed = _tri_mesh(fr, 8, 2)
    K = fg.PADiffusionOperator(2, 2, coords, conn, ed, nd, 1.0, 5, ctx=ctx)
    rng = np.random.default_rng(5)
    cons = np.unique(rng.choice(nd, nd // 8, replace=False)).astype(np.uint32)
    Kb = fg.EssentialBcOperator(K, np.concatenate([cons, cons[:3]]))  # repeats are ignored
    rp, ci, v = fr.simplex_assemble_csr(2, 2, coords, conn, ed, nd, 2, 1.0, 5)
    A = sp.csr_matrix((v, ci.astype(np.int64), rp.astype(np.int64)), shape=(nd, nd))
    x = rng.standard_normal(nd)
    xm = x.copy(); xm[cons] = 0.0
    want = A @ xm; want[cons] = x[cons]
    assert _rel(Kb.matvec(x), want) < 1e-12
    y0 = rng.standard_normal(nd)
    y = y0.copy()
    Kb.apply(x, y)
    assert _rel(y, y0 + want) < 1e-12
    dg = A.diagonal().copy(); dg[cons] = 1.0
    assert _rel(Kb.diagonal(), dg) < 1e-12
    b = rng.standard_normal(nd)
    Kb.mask_rhs(b)
    assert not np.any(b[cons])
    sol = np.zeros(nd)
    res = fg.solve_cg_operator(nd, nd, Kb, b, sol, fg.SolverConfig(rtol=1e-12, max_iter=5000))
    assert res.converged
    free = np.setdiff1d(np.arange(nd), cons)
    xf = spl.spsolve(A[free][:, free].tocsc(), b[free])
    assert np.abs(sol[free] - xf).max() < 1e-9 * np.abs(xf).max()
    assert not np.any(sol[cons])
    with pytest.raises(fg.FemGpuError):
        fg.EssentialBcOperator(K, [nd])


def test_hex_cartesian_qdata(fg, fr, ctx, monkeypatch):
    """On an axis-aligned box mesh (uniform, or graded per axis) the operator
    stores one quadrature value per point (two with mass); its applies and
    diagonal equal those of the general six-value form (FEMGPU_HEX_QDATA=general)
    and the oracle's; a perturbed mesh forms its geometry on the fly, or keeps
    the stored general form when asked to."""
    rng = np.random.default_rng(4)
    n = 5
    # graded per axis: x depends on i only, y on j, z on k
    ax = [np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, n)]) for _ in range(3)]
    ax = [a / a[-1] for a in ax]
    V = np.array([[ax[0][i], ax[1][j], ax[2][k]] for k in range(n + 1) for j in range(n + 1)
                  for i in range(n + 1)])
    for p, form, kappa in ((2, fg.DIFFUSION, "sinprod"), (3, fg.DIFF_MASS, fg.grf(3)), (1, fg.DIFFUSION, 2.5)):
        q = p + 1
        for verts in (None, V):
            K = fg.HexPAOperator(p, q, n, form=form, kappa=kappa, rho=1.3, dirichlet=True, vertices=verts, ctx=ctx)
            assert K.cartesian and K.qdata_per_point() == (1 if form == fg.DIFFUSION else 2)
            monkeypatch.setenv("FEMGPU_HEX_QDATA", "general")
            G = fg.HexPAOperator(p, q, n, form=form, kappa=kappa, rho=1.3, dirichlet=True, vertices=verts, ctx=ctx)
            monkeypatch.delenv("FEMGPU_HEX_QDATA")
            assert not G.cartesian and G.qdata_per_point() == (6 if form == fg.DIFFUSION else 7)
            x = rng.standard_normal(K.n_dofs())
            yk, yg = K.matvec(x), G.matvec(x)
            assert np.abs(yk - yg).max() <= 1e-13 * np.abs(yg).max()
            dk, dg = K.diagonal(), G.diagonal()
            assert np.abs(dk - dg).max() <= 1e-13 * np.abs(dg).max()
            K32 = K.convert(fg.F32)
            assert K32.cartesian
            y32 = K32.matvec(x.astype(np.float32)).astype(np.float64)
            assert np.abs(y32 - yg).max() <= 1e-5 * np.abs(yg).max()
            K.close(); G.close(); K32.close()
    P = fg.HexPAOperator(2, 3, n, form=fg.DIFFUSION, kappa=1.0, dirichlet=True,
                         vertices=fg.hex_vertices(n, n, n, amp=0.1, seed=2), ctx=ctx)
    assert not P.cartesian and P.qdata_mode() == "geometry" and P.qdata_per_point() == 1
    monkeypatch.setenv("FEMGPU_HEX_QDATA", "general")
    P = fg.HexPAOperator(2, 3, n, form=fg.DIFFUSION, kappa=1.0, dirichlet=True,
                         vertices=fg.hex_vertices(n, n, n, amp=0.1, seed=2), ctx=ctx)
    monkeypatch.delenv("FEMGPU_HEX_QDATA")
    assert P.qdata_mode() == "general" and P.qdata_per_point() == 6
    M = fg.HexPAOperator(2, 3, n, form=fg.MASS, rho=1.0, ctx=ctx)
    assert not M.cartesian and M.qdata_per_point() == 1


@pytest.mark.parametrize("shape", [(5, 5, 5), (16, 2, 3), (17, 3, 2), (33, 4, 3), (40, 2, 2)])
def test_hex_q3_quad_per_element(fg, fr, ctx, monkeypatch, shape):
    """Q3 diffusion with 4^3 points on an axis-aligned mesh runs one quad of lanes
    per element (z-layers exchanged by DPP inside the quad, 16 elements per
    wavefront); its applies, diagonal and solves equal the workgroup kernel's
    (FEMGPU_HEX_KERNEL=lds) on meshes whose rows are shorter or longer than 16
    elements, with wavefronts that cross rows and z-planes and a ragged last
    workgroup, FP64 and FP32, with and without essential DOFs."""
    rng = np.random.default_rng(13)
    for dirichlet in (True, False):
        kw = dict(form=fg.DIFFUSION, kappa="sinprod", dirichlet=dirichlet, ctx=ctx)
        K = fg.HexPAOperator(3, 4, shape, **kw)
        assert K.element_kernel() == "quad-per-element" and K.qdata_per_point() == 1
        monkeypatch.setenv("FEMGPU_HEX_KERNEL", "lds")
        L = fg.HexPAOperator(3, 4, shape, **kw)
        monkeypatch.delenv("FEMGPU_HEX_KERNEL")
        assert L.element_kernel() == "workgroup"
        x = rng.standard_normal(K.n_dofs())
        yk, yl = K.matvec(x), L.matvec(x)
        assert np.abs(yk - yl).max() <= 1e-13 * np.abs(yl).max()
        d = fr.hex_desc(3, 4, shape, kind=fr.FR_DIFFUSION, kappa=fr.coef_sinprod(), dirichlet=dirichlet)
        Dq, Mq = fr.hex_qdata(d, fr.hex_vertices(*shape, 0.0, 0))
        assert _rel(yk, fr.hex_apply_sumfac(d, Dq, Mq, x)) < 1e-12
        y0 = rng.standard_normal(K.n_dofs())
        ya, yb = y0.copy(), y0.copy()
        K.apply(x, ya)
        L.apply(x, yb)
        assert np.abs(ya - yb).max() <= 1e-13 * np.abs(yb).max()
        dk, dl = K.diagonal(), L.diagonal()
        assert np.abs(dk - dl).max() <= 1e-13 * np.abs(dl).max()
        K32, L32 = K.convert(fg.F32), L.convert(fg.F32)
        y32 = K32.matvec(x.astype(np.float32)).astype(np.float64)
        assert np.abs(y32 - yl).max() <= 1e-5 * np.abs(yl).max()
        if dirichlet:
            b = K.rhs_unit()
            assert np.array_equal(b, L.rhs_unit())
            xk, xl = np.zeros(K.n_dofs()), np.zeros(K.n_dofs())
            rk, _ = fg.solve_pcg_chebyshev(K, b, xk, fg.SolverConfig(rtol=1e-10, max_iter=2000))
            rl, _ = fg.solve_pcg_chebyshev(L, b, xl, fg.SolverConfig(rtol=1e-10, max_iter=2000))
            assert rk.converged and rl.converged and _iters_close(rk.iterations, rl.iterations)
            assert np.abs(xk - xl).max() <= 1e-8 * np.abs(xl).max()
        for op in (K, L, K32, L32):
            op.close()


@pytest.mark.parametrize("shape", [(5, 5, 5), (70, 3, 2), (64, 2, 3), (64, 5, 3), (128, 3, 2), (192, 3, 2),
                                   (256, 2, 2)])
def test_hex_q2_thread_per_element(fg, fr, ctx, monkeypatch, shape):
    """Q2 diffusion (3^3 points) runs one thread per element with 64-element
    element-vector groups (and, when nx is a multiple of 64, y-faces merged
    inside 256-element blocks); its applies, diagonal and solves equal the
    workgroup kernel's (FEMGPU_HEX_KERNEL=lds) on meshes whose rows are
    shorter or longer than a wavefront, with blocks that cross z-planes,
    box-shaped and perturbed, FP64 and FP32."""
    rng = np.random.default_rng(11)
    nx, ny, nz = shape
    for verts in (None, fg.hex_vertices(nx, ny, nz, amp=0.15, seed=3)):
        for dirichlet in (True, False):
            kw = dict(form=fg.DIFFUSION, kappa="sinprod", dirichlet=dirichlet, vertices=verts, ctx=ctx)
            K = fg.HexPAOperator(2, 3, shape, **kw)
            monkeypatch.setenv("FEMGPU_HEX_KERNEL", "lds")
            L = fg.HexPAOperator(2, 3, shape, **kw)
            monkeypatch.delenv("FEMGPU_HEX_KERNEL")
            x = rng.standard_normal(K.n_dofs())
            yk, yl = K.matvec(x), L.matvec(x)
            assert np.abs(yk - yl).max() <= 1e-13 * np.abs(yl).max()
            dk, dl = K.diagonal(), L.diagonal()
            assert np.abs(dk - dl).max() <= 1e-13 * np.abs(dl).max()
            K32, L32 = K.convert(fg.F32), L.convert(fg.F32)
            y32, l32 = K32.matvec(x.astype(np.float32)), L32.matvec(x.astype(np.float32))
            assert np.abs(y32.astype(np.float64) - yl).max() <= 1e-5 * np.abs(yl).max()
            assert np.abs(l32.astype(np.float64) - yl).max() <= 1e-5 * np.abs(yl).max()
            if dirichlet:
                b = K.rhs_unit()
                assert np.array_equal(b, L.rhs_unit())
                xk, xl = np.zeros(K.n_dofs()), np.zeros(K.n_dofs())
                rk, _ = fg.solve_pcg_chebyshev(K, b, xk, fg.SolverConfig(rtol=1e-10, max_iter=2000))
                rl, _ = fg.solve_pcg_chebyshev(L, b, xl, fg.SolverConfig(rtol=1e-10, max_iter=2000))
                assert rk.converged and rl.converged and _iters_close(rk.iterations, rl.iterations)
                assert np.abs(xk - xl).max() <= 1e-8 * np.abs(xl).max()
            for op in (K, L, K32, L32):
                op.close()


def test_hex_cartesian_mirrored_and_anisotropic(fg, ctx, monkeypatch):
    """Cartesian qdata on a box mesh whose x axis runs backwards (negative
    spacing) and whose spacings differ by 100x between axes equals the general
    form, for the workgroup and the thread-per-element kernels."""
    rng = np.random.default_rng(21)
    n = (64, 3, 2)
    ax = [np.linspace(1.0, 0.0, n[0] + 1), np.linspace(0.0, 0.01, n[1] + 1), np.linspace(0.0, 1.0, n[2] + 1)]
    V = np.array([[ax[0][i], ax[1][j], ax[2][k]] for k in range(n[2] + 1) for j in range(n[1] + 1)
                  for i in range(n[0] + 1)])
    for p in (2, 3):
        kw = dict(form=fg.DIFFUSION, kappa="sinprod", dirichlet=True, vertices=V, ctx=ctx)
        K = fg.HexPAOperator(p, p + 1, n, **kw)
        assert K.cartesian
        monkeypatch.setenv("FEMGPU_HEX_QDATA", "general")
        G = fg.HexPAOperator(p, p + 1, n, **kw)
        monkeypatch.delenv("FEMGPU_HEX_QDATA")
        assert not G.cartesian
        x = rng.standard_normal(K.n_dofs())
        yk, yg = K.matvec(x), G.matvec(x)
        assert np.abs(yk - yg).max() <= 1e-12 * np.abs(yg).max()
        assert np.abs(K.diagonal() - G.diagonal()).max() <= 1e-12 * np.abs(G.diagonal()).max()
        K.close(); G.close()
