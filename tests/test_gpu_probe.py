"""The contraction probe (femgpu_sumfac_probe): the vector-unit, even-odd and
MFMA engines compute the same batched sweep Out = [B u | G u], checked against
a numpy evaluation of the same recurrence with the oracle's 1-D tables
(Gauss-Legendre points, equispaced Lagrange nodes).  The MFMA engine's
agreement also pins the assumed operand and result lane layouts of
v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32."""
import numpy as np
import pytest

import femgpu as fg

pytestmark = pytest.mark.gpu


def tables(p, q):
    x, _ = np.polynomial.legendre.leggauss(q)
    nodes = np.linspace(-1.0, 1.0, p + 1)
    B = np.zeros((q, p + 1)); G = np.zeros_like(B)
    for i in range(p + 1):
        c = np.zeros(p + 1); c[i] = 1.0
        poly = np.polynomial.Polynomial.fit(nodes, c, p, domain=[-1, 1], window=[-1, 1])
        B[:, i] = poly(x)
        G[:, i] = poly.deriv()(x)
    return B, G


def reference(u, B, G, reps):
    """numpy restatement of the probe's recurrence on every line."""
    q, n1 = B.shape
    u = u.copy()
    for _ in range(reps):
        out = np.concatenate([u @ B.T, u @ G.T], axis=1)
        u[:, :n1] += 2.0 ** -10 * (out[:, :n1] + out[:, q:q + n1])
    return out


@pytest.mark.parametrize("p,q", [(3, 4), (4, 5)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_probe_engines_agree(p, q, dtype):
    ctx = fg.default_context()
    nblocks, reps = 6, 7
    lines = nblocks * 256
    rng = np.random.default_rng(3)
    u = rng.standard_normal((lines, p + 1))
    x = fg.DeviceArray(ctx, lines * (p + 1), dtype).upload(u.ravel().astype(dtype))
    B, G = tables(p, q)
    want = reference(u.astype(dtype).astype(np.float64), B, G, reps)
    tol = 1e-12 if dtype == np.float64 else 2e-5
    fdt = fg.F64 if dtype == np.float64 else fg.F32
    for engine in fg.SUMFAC_ENGINES:
        y = fg.DeviceArray(ctx, lines * 2 * q, dtype)
        ms = fg.sumfac_probe(p, q, fdt, engine, nblocks, reps, x, y, launches=2, ctx=ctx)
        got = y.download().reshape(lines, 2 * q).astype(np.float64)
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"p={p} q={q} {np.dtype(dtype).name} {engine}: {ms:.4f} ms, rel err {err:.2e}")
        assert err < tol, (engine, err)


def test_probe_rejects_bad_arguments():
    ctx = fg.default_context()
    x = fg.DeviceArray(ctx, 256 * 4, np.float64)
    y = fg.DeviceArray(ctx, 256 * 8, np.float64)
    with pytest.raises(ValueError):
        fg.sumfac_probe(3, 4, fg.F64, "mfma", 2, 1, x, y, ctx=ctx)
    with pytest.raises(fg.FemGpuError):
        fg.sumfac_probe(2, 3, fg.F64, "valu", 1, 1, fg.DeviceArray(ctx, 256 * 3), fg.DeviceArray(ctx, 256 * 6),
                        ctx=ctx)
