"""The allocation-free PCG-Jacobi entry point (femgpu_solve_pcg_jacobi_noalloc,
no counterpart in the reference) against femgpu_solve_pcg_jacobi on the same
device vectors: bit-identical iterates on two calls in a row, the second with
a smaller system (the shared scratch array keeps the first call's contents),
in FP64 and FP32; sizes beyond the scratch array and operators whose apply
needs a workspace are refused with an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(fr, n):
    # the config-1 system at n x n: P1 Poisson on the unit square, row-zeroed Dirichlet rows
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(n)
    return rp, ci, v, rhs


def _pair(fg, ctx, A, b, cfg, dtype):
    N = len(b)
    bd = fg.DeviceArray(ctx, N, dtype, b.astype(dtype))
    x1 = fg.DeviceArray(ctx, N, dtype)
    x2 = fg.DeviceArray(ctx, N, dtype)
    r1 = fg.solve_pcg_jacobi(A, bd, x1, cfg)
    r2 = fg.solve_pcg_jacobi_noalloc(A, bd, x2, cfg)
    return r1, x1.download(), r2, x2.download()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_pcg_jacobi(fg, fr, ctx, dtype):
    dt = fg.F64 if dtype == np.float64 else fg.F32
    cfg = fg.SolverConfig(rtol=1e-10 if dt == fg.F64 else 1e-5, max_iter=5000)
    for n in (96, 40):  # the second call is smaller
        rp, ci, v, b = _system(fr, n)
        A = fg.CsrMatrix(rp, ci, v.astype(dtype), dtype=dt, ctx=ctx)
        r1, x1, r2, x2 = _pair(fg, ctx, A, b, cfg, dtype)
        print(f"n = {n}: solve_pcg_jacobi {r1}, noalloc {r2}")
        assert r1 == r2 and r1.converged
        assert np.array_equal(x1, x2)
        # a stop at max_iter returns the same partial iterate
        r1, x1, r2, x2 = _pair(fg, ctx, A, b, fg.SolverConfig(rtol=1e-10, max_iter=7), dtype)
        assert r1 == r2 and not r1.converged and r1.iterations == 7
        assert np.array_equal(x1, x2)
        A.close()


def test_noalloc_refuses(fg, fr, ctx):
    # beyond the scratch array: a diagonal system of 838,849 unknowns
    n = 838849
    A = fg.CsrMatrix(np.arange(n + 1, dtype=np.uint64), np.arange(n, dtype=np.uint32), np.full(n, 2.0), ctx=ctx)
    b = fg.DeviceArray(ctx, n, np.float64, np.ones(n))
    x = fg.DeviceArray(ctx, n, np.float64)
    with pytest.raises(fg.FemGpuError, match="scratch"):
        fg.solve_pcg_jacobi_noalloc(A, b, x)
    A.close()
    # the largest size it takes solves this diagonal system exactly
    n = 838848
    A = fg.CsrMatrix(np.arange(n + 1, dtype=np.uint64), np.arange(n, dtype=np.uint32), np.full(n, 2.0), ctx=ctx)
    b = fg.DeviceArray(ctx, n, np.float64, np.ones(n))
    x = fg.DeviceArray(ctx, n, np.float64)
    res = fg.solve_pcg_jacobi_noalloc(A, b, x)
    assert res.converged and np.array_equal(x.download(), np.full(n, 0.5))
    A.close()
    # an operator whose apply needs workspace (hex PA) is refused
    op = fg.HexPAOperator(2, 3, 3, form=fg.DIFFUSION, ctx=ctx)
    b = fg.DeviceArray(ctx, op.n_dofs(), np.float64, np.ones(op.n_dofs()))
    x = fg.DeviceArray(ctx, op.n_dofs(), np.float64)
    with pytest.raises(fg.FemGpuError, match="CSR"):
        fg.solve_pcg_jacobi_noalloc(op, b, x)
