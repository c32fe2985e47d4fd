"""GPU parity of the hex PA apply with the geometry formed on the fly (the
default on meshes that are not axis-aligned boxes: w kappa stored per Gauss
point, the Jacobian rebuilt from the element's 8 vertices by every apply;
FEMGPU_HEX_QDATA=geometry asks for it explicitly) on perturbed
meshes: against the oracle's stored-qdata apply and the reference element loop
(partial.rs:151-208 generalised to hex), the linear patch test (which pins
|J| and J^-1 J^-T independently of any shared derivation), the diagonal,
essential rows, the FP32 copy and a solve.
Bars: 1e-12 relative (max-norm) in FP64, 1e-5 in FP32."""
import numpy as np
import pytest

import hexgeom

pytestmark = pytest.mark.gpu

CASES = [(1, 2), (2, 3), (3, 4), (4, 5), (5, 6), (2, 4), (3, 5), (4, 6)]


def _rel(a, b):
    a = np.asarray(a, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture
def geo(monkeypatch):
    monkeypatch.setenv("FEMGPU_HEX_QDATA", "geometry")


def _pair(fg, fr, ctx, p, q, n, form, kappa="sinprod", rho=1.3, dirichlet=False, amp=0.2, seed=3,
          dtype=None):
    c = fg.hex_vertices(*n, amp, seed)
    if kappa == "sinprod":
        kg, kr = "sinprod", fr.coef_sinprod()
    elif isinstance(kappa, dict):
        kg, kr = kappa, fr.coef_grf(kappa["seed"])
    else:
        kg, kr = float(kappa), fr.coef_const(kappa)
    op = fg.HexPAOperator(p, q, n, form=form, kappa=kg, rho=rho, dirichlet=dirichlet,
                          dtype=fg.F64 if dtype is None else dtype, vertices=c, ctx=ctx)
    d = fr.hex_desc(p, q, n, kind=form, kappa=kr, rho=rho, dirichlet=dirichlet)
    Dq, Mq = fr.hex_qdata(d, c)
    return op, d, c, Dq, Mq


@pytest.mark.parametrize("p,q", CASES)
@pytest.mark.parametrize("form", [2, 3])
@pytest.mark.parametrize("n,dirichlet", [((3, 2, 4), False), ((11, 5, 4), True)])
def test_geometry_apply_f64(fg, fr, ctx, geo, p, q, form, n, dirichlet):
    op, d, c, Dq, Mq = _pair(fg, fr, ctx, p, q, n, form, dirichlet=dirichlet)
    assert op.qdata_mode() == "geometry" and op.qdata_per_point() == 1
    N = op.n_dofs()
    x = np.random.default_rng(p * 10 + q + form).standard_normal(N)
    ref = fr.hex_apply_sumfac(d, Dq, Mq, x)
    y = np.full(N, np.nan)
    op.apply_zero(x, y)
    assert _rel(y, ref) < 1e-12
    # MatFreeOperator::apply accumulates
    y0 = np.random.default_rng(5).standard_normal(N)
    y1 = y0.copy()
    op.apply(x, y1)
    assert np.abs(y1 - (y0 + ref)).max() < 1e-12 * (np.abs(ref).max() + np.abs(y0).max())


@pytest.mark.parametrize("p,q", [(2, 3), (3, 4), (4, 5)])
def test_geometry_grf_refloop_diag_rhs(fg, fr, ctx, geo, p, q):
    """GRF kappa with essential rows: apply against the stored-qdata oracle and
    the reference element loop; the diagonal (built from stored qdata at
    set-up) and the load vector."""
    op, d, c, Dq, Mq = _pair(fg, fr, ctx, p, q, (4, 3, 3), 3, kappa=dict(kind="grf", seed=11),
                             dirichlet=True)
    assert op.qdata_mode() == "geometry"
    x = np.random.default_rng(1).standard_normal(op.n_dofs())
    assert _rel(op.matvec(x), fr.hex_apply_sumfac(d, Dq, Mq, x)) < 1e-12
    assert _rel(op.matvec(x), fr.hex_apply_refloop(d, c, x)) < 1e-12
    assert _rel(op.diagonal(), fr.hex_diagonal(d, Dq, Mq)) < 1e-12
    assert _rel(op.rhs_unit(), fr.hex_rhs_unit(d, c)) < 1e-12


def test_geometry_qp_array(fg, fr, ctx, geo):
    """Per-point kappa values (QP_ARRAY) with the geometry on the fly."""
    n = (3, 3, 3)
    c = fg.hex_vertices(*n, 0.2, 3)
    kq = np.random.default_rng(4).random((27, 64)) + 0.5
    op = fg.HexPAOperator(3, 4, n, form=2, kappa=kq.ravel(), vertices=c, ctx=ctx)
    assert op.qdata_mode() == "geometry"
    dk = fr.hex_desc(3, 4, n, kind=2, kappa=fr.coef_const(1.0))
    D1, M1 = fr.hex_qdata(dk, c)
    x = np.random.default_rng(2).standard_normal(op.n_dofs())
    assert _rel(op.matvec(x), fr.hex_apply_sumfac(dk, D1 * kq[:, :, None], M1, x)) < 1e-12


@pytest.mark.parametrize("p,q", [(2, 3), (3, 4), (4, 5)])
def test_geometry_f32_and_convert(fg, fr, ctx, monkeypatch, p, q):
    """The FP32 operator with the geometry on the fly, created directly and as
    the converted copy of an FP64 one (which keeps its source's mode even when
    the environment asks for stored qdata by then)."""
    monkeypatch.setenv("FEMGPU_HEX_QDATA", "geometry")
    op32, d, c, Dq, Mq = _pair(fg, fr, ctx, p, q, (4, 4, 3), 3, dtype=fg.F32)
    op64, _, _, _, _ = _pair(fg, fr, ctx, p, q, (4, 4, 3), 3)
    monkeypatch.setenv("FEMGPU_HEX_QDATA", "general")
    conv = op64.convert(fg.F32)
    assert op32.qdata_mode() == "geometry" and conv.qdata_mode() == "geometry"
    x = np.random.default_rng(7).standard_normal(op32.n_dofs()).astype(np.float32)
    ref64 = fr.hex_apply_sumfac(d, Dq, Mq, x.astype(np.float64))
    for op in (op32, conv):
        y = np.zeros_like(x)
        op.apply_zero(x, y)
        assert _rel(y, ref64) < 1e-5


@pytest.mark.parametrize("p,q", [(1, 2), (2, 3), (3, 4), (4, 5), (5, 6)])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_geometry_patch_test(fg, ctx, geo, p, q, dtype):
    """Linear u = a.x + c, v = b.x + d on the perturbed mesh (kappa = rho = 1):
    v^T K u = a.b, (K u)_i = 0 at interior DOFs, v^T (K + M) u = a.b + int uv
    -- exact in the mathematics, so the Jacobian the kernel rebuilds is pinned
    on its own."""
    n = (5, 3, 4)
    c = fg.hex_vertices(*n, 0.2, 21)
    X = hexgeom.dof_coords(*n, p, c)
    rng = np.random.default_rng(7 * p + q)
    a, b = rng.standard_normal(3), rng.standard_normal(3)
    c0, d0 = rng.standard_normal(2)
    u, v = X @ a + c0, X @ b + d0
    dt = fg.F64 if dtype == "f64" else fg.F32
    npt = np.float64 if dtype == "f64" else np.float32
    tol = 1e-12 if dtype == "f64" else 1e-5
    K = fg.HexPAOperator(p, q, n, form=fg.DIFFUSION, kappa=1.0, vertices=c, dtype=dt, ctx=ctx)
    assert K.qdata_mode() == "geometry"
    Ku = K.matvec(u.astype(npt)).astype(np.float64)
    assert abs(v @ Ku - a @ b) < tol * (1 + abs(a @ b) + np.abs(v).max() * np.abs(Ku).sum())
    inner = ~hexgeom.boundary_mask(*n, p)
    assert np.abs(Ku[inner]).max() < tol * max(np.abs(Ku).max(), 1.0)
    if q >= 3:
        KM = fg.HexPAOperator(p, q, n, form=fg.DIFF_MASS, kappa=1.0, rho=1.0, vertices=c, dtype=dt, ctx=ctx)
        assert KM.qdata_mode() == "geometry"
        exact = hexgeom.int_linear_product(a, c0, b, d0)
        KMu = KM.matvec(u.astype(npt)).astype(np.float64)
        assert abs(v @ KMu - (a @ b + exact)) < tol * (1 + abs(exact) + np.abs(v).max() * np.abs(KMu).sum())


def test_geometry_box_mesh_keeps_cartesian(fg, ctx, geo):
    """On an axis-aligned box mesh the one-value Cartesian qdata is cheaper,
    and the request for the geometry on the fly does not replace it; the
    mass-only form keeps its one stored value per point."""
    K = fg.HexPAOperator(3, 4, (4, 4, 4), form=fg.DIFFUSION, kappa=1.0, ctx=ctx)
    assert K.qdata_mode() == "cartesian"
    M = fg.HexPAOperator(3, 4, (4, 4, 4), form=fg.MASS, rho=1.0,
                         vertices=fg.hex_vertices(4, 4, 4, 0.2, 1), ctx=ctx)
    assert M.qdata_mode() == "general" and M.qdata_per_point() == 1


def test_geometry_solve_matches_stored(fg, ctx, monkeypatch):
    """PCG + Chebyshev-Jacobi(4) on a Q4 perturbed diffusion + mass system
    (GRF kappa): the geometry-on-the-fly operator's solution equals the
    stored-qdata operator's to 1e-10."""
    n = (6, 6, 6)
    c = fg.hex_vertices(*n, 0.2, 5)
    monkeypatch.setenv("FEMGPU_HEX_QDATA", "general")
    A = fg.HexPAOperator(4, 5, n, form=fg.DIFF_MASS, kappa=fg.grf(7), rho=1.0, dirichlet=True,
                         vertices=c, ctx=ctx)
    monkeypatch.setenv("FEMGPU_HEX_QDATA", "geometry")
    G = fg.HexPAOperator(4, 5, n, form=fg.DIFF_MASS, kappa=fg.grf(7), rho=1.0, dirichlet=True,
                         vertices=c, ctx=ctx)
    assert A.qdata_mode() == "general" and G.qdata_mode() == "geometry"
    b = A.rhs_random(11)
    cfg = fg.SolverConfig(rtol=1e-12, max_iter=5000)
    xa = np.zeros(A.n_dofs())
    xg = np.zeros(A.n_dofs())
    ra, _ = fg.solve_pcg_chebyshev(A, b, xa, cfg)
    rg, _ = fg.solve_pcg_chebyshev(G, b, xg, cfg)
    assert ra.converged and rg.converged
    assert abs(ra.iterations - rg.iterations) <= 2
    assert _rel(xg, xa) < 1e-10
