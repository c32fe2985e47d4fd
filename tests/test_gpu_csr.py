"""GPU parity: CsrMatrix SpMV / spmv_add / diagonal through the C ABI against
the oracle's restatement of linalg/src/csr.rs.  FP64 row sums follow the
reference's 4-way blocked order (csr.rs:38-64) with FMA contraction off, so
the short-row kernel is compared bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lap1d(n):
    rp = [0]; ci = []; v = []
    for i in range(n):
        if i > 0: ci.append(i - 1); v.append(-1.0)
        ci.append(i); v.append(2.0)
        if i < n - 1: ci.append(i + 1); v.append(-1.0)
        rp.append(len(ci))
    return np.array(rp, np.uint64), np.array(ci, np.uint32), np.array(v)


def test_spmv_tridiag_known_answer(fg, ctx):
    # csr.rs:589-598: [2 -1 0; -1 2 -1; 0 -1 2] [1 2 3] = [0 0 4]
    A = fg.CsrMatrix(*_lap1d(3), ctx=ctx)
    y = np.zeros(3)
    A.spmv(np.array([1.0, 2.0, 3.0]), y)
    assert y.tolist() == [0.0, 0.0, 4.0]
    assert A.diagonal().tolist() == [2.0, 2.0, 2.0]  # csr.rs:601-605


@pytest.mark.parametrize("stage", ["0", "1"])
def test_spmv_p1_poisson_256_bitwise(fg, fr, ctx, monkeypatch, stage):
    """Config-1 matrix, bitwise, with the row kernel's LDS staging off (its default
    below 2^19 rows) and forced on."""
    rp, ci, v, rhs, _, _ = fr.poisson_p1(256)
    assert len(v) == 460289
    monkeypatch.setenv("FEMGPU_CSR_STAGE", stage)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    ref = fr.CsrRef(rp, ci, v)
    x = np.random.default_rng(0).standard_normal(len(rhs))
    y = np.zeros_like(x)
    A.spmv(x, y)
    assert np.array_equal(y, ref.spmv(x))
    y2 = np.random.default_rng(1).standard_normal(len(rhs))
    y3 = y2.copy()
    A.spmv_add(0.7, x, -1.3, y3)
    assert np.array_equal(y3, ref.spmv_add(0.7, x, -1.3, y2))
    assert np.array_equal(A.diagonal(), ref.diagonal())


def test_spmv_large_default_staged(fg, fr, ctx, monkeypatch):
    """P1 Poisson on 768^2 quads: 591,361 rows, above the 2^19 rows from which the
    row kernel stages its entries in LDS by default. Bitwise."""
    monkeypatch.delenv("FEMGPU_CSR_STAGE", raising=False)
    rp, ci, v, rhs, _, _ = fr.poisson_p1(768)
    assert len(rhs) == 769 * 769 >= 1 << 19
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    x = np.random.default_rng(5).standard_normal(len(rhs))
    y = np.zeros_like(x)
    A.spmv(x, y)
    assert np.array_equal(y, fr.CsrRef(rp, ci, v).spmv(x))


def test_spmv_f32(fg, fr, ctx):
    rp, ci, v, rhs, _, _ = fr.poisson_p1(64)
    A = fg.CsrMatrix(rp, ci, v.astype(np.float32), dtype=fg.F32, ctx=ctx)
    x = np.random.default_rng(2).standard_normal(len(rhs)).astype(np.float32)
    y = np.zeros_like(x)
    A.spmv(x, y)
    ref = fr.csr_spmv_f32(rp, ci, v.astype(np.float32), x)
    assert np.array_equal(y, ref)
    y64 = fr.CsrRef(rp, ci, v).spmv(x.astype(np.float64))
    assert np.abs(y - y64).max() <= 1e-5 * np.abs(y64).max()


@pytest.mark.parametrize("f32", [False, True])
def test_spmv_staged_and_unstaged_blocks(fg, fr, ctx, monkeypatch, f32):
    """With staging on (forced here; by default from 2^19 rows), the row kernel
    copies a 256-row block's entries to LDS when they number at most 2048, else
    reads them from memory: 2,000 rows averaging under 16 entries (so the row
    kernel runs). Rows 300-399 hold 41 entries each, so block 1 holds 4,468
    entries and is not staged. Every 7th row is empty, and the last block is
    ragged. Bitwise against the oracle in both precisions, for y = A x and
    y = a A x + b y."""
    monkeypatch.setenv("FEMGPU_CSR_STAGE", "1")
    n = 2000
    rng = np.random.default_rng(11)
    rp = [0]; ci = []
    for i in range(n):
        k = 0 if i % 7 == 0 else 41 if 300 <= i < 400 else 1 + i % 6
        ci += sorted(rng.choice(n, k, replace=False).tolist())
        rp.append(len(ci))
    rp = np.array(rp, np.uint64); ci = np.array(ci, np.uint32)
    v = rng.standard_normal(len(ci))
    x = rng.standard_normal(n); y0 = rng.standard_normal(n)
    if f32:
        v32, x32, y32 = v.astype(np.float32), x.astype(np.float32), y0.astype(np.float32)
        A = fg.CsrMatrix(rp, ci, v32, dtype=fg.F32, ctx=ctx)
        y = np.zeros(n, np.float32)
        A.spmv(x32, y)
        assert np.array_equal(y, fr.csr_spmv_f32(rp, ci, v32, x32))
        return
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    ref = fr.CsrRef(rp, ci, v)
    y = np.zeros(n)
    A.spmv(x, y)
    assert np.array_equal(y, ref.spmv(x))
    y3 = y0.copy()
    A.spmv_add(0.7, x, -1.3, y3)
    assert np.array_equal(y3, ref.spmv_add(0.7, x, -1.3, y0))


def test_spmv_long_rows(fg, fr, ctx):
    """Assembled hex Q3 matrix (343 nnz per interior row): sub-wavefront kernel."""
    d = fr.hex_desc(3, 4, (3, 3, 3), kind=fr.FR_DIFF_MASS, kappa=fr.coef_sinprod())
    c = fr.hex_vertices(3, 3, 3, 0.2, 1)
    rp, ci, v = fr.hex_assemble_csr(d, c)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    x = np.random.default_rng(3).standard_normal(len(rp) - 1)
    y = np.zeros_like(x)
    A.spmv(x, y)
    ref = fr.CsrRef(rp, ci, v).spmv(x)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_spmv_device_arrays_and_errors(fg, fr, ctx):
    rp, ci, v = _lap1d(1000)
    A = fg.CsrMatrix(rp, ci, v, ctx=ctx)
    x = np.linspace(0, 1, 1000)
    dx = fg.DeviceArray(ctx, 1000, np.float64, x)
    dy = fg.DeviceArray(ctx, 1000, np.float64)
    A.spmv(dx, dy)
    assert np.array_equal(dy.download(), fr.CsrRef(rp, ci, v).spmv(x))
    with pytest.raises(fg.DimensionMismatch):
        A.spmv(np.zeros(999), np.zeros(1000))
    with pytest.raises(fg.FemGpuError):
        fg.CsrMatrix(rp, np.full_like(ci, 5000), v, ctx=ctx)  # column out of range
    # empty matrix
    E = fg.CsrMatrix(np.zeros(1, np.uint64), np.zeros(0, np.uint32), np.zeros(0), ctx=ctx)
    assert E.n_dofs() == 0


def test_device_array_zero_filled(fg, ctx):
    """femgpu_malloc returns zero-filled memory, also when the runtime hands
    back a block that held data (a fresh DeviceArray is a valid zero initial
    guess)."""
    n = 1 << 20
    for _ in range(3):
        a = fg.DeviceArray(ctx, n, np.float64, np.full(n, 3.5))
        a.free()
        b = fg.DeviceArray(ctx, n, np.float64)
        assert not np.any(b.download())
        b.free()
