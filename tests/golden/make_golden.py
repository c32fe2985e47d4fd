"""Writes the small golden fixtures in tests/golden/ from the CPU oracle.

Run from the repository root: ``python tests/golden/make_golden.py``.
The fixtures are data only (inputs and expected outputs):
  hex_fixtures.npz   2x2x2 hex Q1..Q4 (affine and perturbed): coords, x, y = A x
                     for diffusion+mass with kappa = 1 + 0.5 sin sin sin, rho = 1.3
  csr_fixtures.npz   the 3x3 tridiagonal matrix of linalg/src/csr.rs:575-585 and
                     its known answers; the n = 8 P1 Poisson system of
                     examples/mfem_ex1_poisson.rs (A, b after row zeroing, PCG
                     solution); the 1-D Laplacian n = 50 Chebyshev run of
                     amg/src/lib.rs:254-286 (x after 20 reference-mode steps)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import femref as fr  # noqa: E402


def hex_fixtures():
    out = {}
    for p in (1, 2, 3, 4):
        for amp in (0.0, 0.2):
            key = f"q{p}_{'pert' if amp else 'affine'}"
            d = fr.hex_desc(p, p + 1, (2, 2, 2), kind=fr.FR_DIFF_MASS, kappa=fr.coef_sinprod(), rho=1.3)
            c = fr.hex_vertices(2, 2, 2, amp, 7)
            x = np.random.default_rng(100 + p).standard_normal(fr.hex_ndofs(d))
            out[key + "_coords"] = c
            out[key + "_x"] = x
            out[key + "_y"] = fr.hex_apply_refloop(d, c, x)
    np.savez_compressed(os.path.join(HERE, "hex_fixtures.npz"), **out)


def lap1d(n):
    rp = [0]; ci = []; v = []
    for i in range(n):
        if i > 0: ci.append(i - 1); v.append(-1.0)
        ci.append(i); v.append(2.0)
        if i < n - 1: ci.append(i + 1); v.append(-1.0)
        rp.append(len(ci))
    return np.array(rp, np.uint64), np.array(ci, np.uint32), np.array(v)


def csr_fixtures():
    out = {}
    rp, ci, v = lap1d(3)
    out["tri_rp"], out["tri_ci"], out["tri_v"] = rp, ci, v
    out["tri_x"] = np.array([1.0, 2.0, 3.0])
    out["tri_y"] = fr.CsrRef(rp, ci, v).spmv(out["tri_x"])
    rp, ci, v, rhs, coords, conn = fr.poisson_p1(8)
    A = fr.CsrRef(rp, ci, v)
    op = fr._Op(len(rhs), csr=A)
    x, res = fr.pcg_jacobi(op, A.diagonal(), rhs, np.zeros(len(rhs)), rtol=1e-10, max_iter=5000)
    out["p8_rp"], out["p8_ci"], out["p8_v"], out["p8_b"], out["p8_x"] = rp, ci, v, rhs, x
    out["p8_iters"] = np.array([res.iterations])
    rp, ci, v = lap1d(50)
    A = fr.CsrRef(rp, ci, v)
    rho = fr.estimate_spectral_radius(A, 20)
    lmax = 1.1 * rho; lmin = lmax / 30
    out["cheb_rho20"] = np.array([rho])
    out["cheb_rho30"] = np.array([fr.estimate_spectral_radius(A, 30)])
    out["cheb_x20"] = fr.chebyshev_smooth_ref(A, np.zeros(50), np.ones(50), lmin, lmax, 20)
    np.savez_compressed(os.path.join(HERE, "csr_fixtures.npz"), **out)


if __name__ == "__main__":
    hex_fixtures()
    csr_fixtures()
    print("written")
