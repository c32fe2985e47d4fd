#!/usr/bin/env python3
"""Benchmark of the fem-rs solve path on one MI355X.

Default workload (q4-128-perturbed, the large case):
  3-D diffusion + mass (rho = 1) on the unit cube, 128^3 hex elements with
  interior vertices perturbed by up to 0.2 h (seed 5), H1 Q4 (513^3 =
  135,005,697 DOFs), kappa = exp(GRF) clipped to [1e-2, 1e2] (seed 7), random
  nodal RHS (seed 11), homogeneous Dirichlet BCs, FP64 PA (5^3 Gauss points),
  PCG + Chebyshev-Jacobi(4).

One step = one fixed chunk of --chunk plain FP64 PCG + Chebyshev-Jacobi(4)
iterations (40 by default for this workload) from x0 = 0, with b, x and the
operator resident in HBM.  value = N_dofs * PCG iterations / time over the
timed steps (GDOF/s).  A workload with chunk 0 (the Q2 one) times whole solves.

The line also carries the PA apply rate and its roofline fraction
(algorithmic bytes 8 (2N + s E q^3) per apply, s = quadrature-data values per
Gauss point: 6 diffusion / 7 diffusion + mass on a general mesh, 1 / 2 on an
axis-aligned box mesh; over the HIP-event-timed mean duration of the
apply's two kernels), the HBM traffic per apply from the committed counter
summary for that workload (profiles/*_hbm_traffic.md), the bandwidth of a
1 GiB device-to-device copy timed in the same process, and a CPU baseline
(the oracle on the host cores, bounded samples: the reference-algorithm
element loop and the sum-factorised apply on 1 core and on all cores, and
the PCG-Chebyshev solve rate on all cores).

"legs" runs whole solves once on the same operator and right-hand side:
plain FP64 PCG-Chebyshev to the workload's rtol, then continued (restarted
from that solution) to --tight-rtol; the FP32 PCG-Chebyshev with FP64
refinement (continued form) to rtol and on to --tight-rtol; and FP64
refinement around the FP32 flexible CG whose preconditioner is
Chebyshev-Jacobi smoothing plus the learned U-Net's correction, the network
running on a coarsened copy of the DOF grid (trained first, on pairs from
PCG-Chebyshev iterates of solves with a known solution; harvest and training
times are reported beside, not inside, the solve time), to rtol and on to
--tight-rtol.  Errors are against an FP64 reference solution: the plain solve
continued to --ref-rtol (restarted from the tight solution, with a recomputed
residual).

--workload selects the other configurations:
  p1-256-csr        2-D Poisson on the unit square, 256^2 quads split into two
                    triangles each, P1 (66,049 DOFs), CSR and load vector
                    assembled on the GPU (assemble_bilinear / assemble_linear,
                    row zeroing by apply_dirichlet), PCG + Jacobi to 1e-10
  q2-64-sinprod     3-D diffusion, 64^3 hex Q2 (2,146,689 DOFs), kappa = 1 +
                    0.5 sin sin sin, RHS int phi_i, PCG + Chebyshev-Jacobi(4)
                    to 1e-8; one step = one whole solve
  q3-96-grf         Q3, 96^3, kappa = exp(GRF) clipped to [1e-2, 1e2], random
                    nodal RHS, rtol 1e-10 (24,137,569 DOFs)
  q3-96-learned     the learned preconditioner on the Q3 96^3 GRF family: the
                    U-Net is trained on --train-fields fields (kappa seeds
                    0, 1, ..., pairs from PCG-Chebyshev iterates) and compared
                    with PCG + Chebyshev-Jacobi(4) on --test-fields held-out
                    fields (seeds 256, 257, ...), FP64, to 1e-8; value = N x
                    iterations / time of the learned solves
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fem-rs_amd"))

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak


def sumfac_fma_per_element(n1, q, diffusion, mass, cartesian):
    """Multiply-adds of one element in the sum-factorised apply, as the element
    kernels perform them: X and Y^T / X^T contractions on the B (and G) paths,
    the Z contraction with the pointwise product and its transpose."""
    paths_x = (2 if diffusion else 1)
    paths_y = (3 if diffusion else 1)
    x = n1 * n1 * q * n1 * paths_x                  # X: N1^2 lines, Q outputs, N1 inputs
    y = q * n1 * q * n1 * paths_y                   # Y: (qx, k) lanes, Q outputs, N1 inputs
    z_col = (3 * n1 if diffusion else 0) + (n1 if mass else 0)            # gradients / value at a point
    z_col += ((3 if cartesian else 9) if diffusion else 0) + (1 if mass else 0)  # pointwise product
    z_col += (3 * n1 if diffusion else 0) + (n1 if mass else 0)           # transpose accumulation
    z = q * q * q * z_col                           # Z: Q^2 columns, Q points each
    return x + y + z + y + x                        # Y^T and X^T mirror Y and X


# Floating-point operations per Gauss point of the geometry formed on the fly,
# counting a multiply-add as 2: the Jacobian's two zeta-dependent columns (6
# multiply-adds), the rows of det(J) J^-1 as cross products (9 multiplies + 9
# multiply-adds), det (1 + 2), the refined reciprocal of |det| (1 + 4) and
# its product with w kappa (1), h = R^T g (3 + 6), v = f R h (3 + 6 + 3), and
# the lane's set-up from the 8 vertices (about 126 per column of Q points,
# spread over its points).  The stored form's pointwise product is 9
# multiply-adds (18).
def geometry_flops_per_point(q):
    return 12 + 27 + 5 + 9 + 1 + 15 + 18 + 126.0 / q


def sumfac_flops_even_odd(n1, q, mass):
    """Flops (multiply-add = 2) of one element's diffusion (+ mass)
    contractions in the even-odd form of the geometry kernel (eo_split,
    eo_contract, eo_contract2): per line, the mirrored sums and differences of
    its input, the even and odd half-products, and the recombination."""
    def split(ni):
        return ni // 2                                   # additions (and as many subtractions)
    def eo(r, ni, sig):                                  # (multiply-adds, additions) of M u
        hr, oddr, hc, oddc = r // 2, r % 2, ni // 2, ni % 2
        ev = hr + (oddr if sig == 1 else 0)
        od = hr + (oddr if sig == -1 else 0)
        return ev * (hc + oddc) + od * hc, 2 * hr
    def eo2(r, ni):                                      # M1 u1 + M2 u2, symmetries +1 / -1
        hr, oddr, hc, oddc = r // 2, r % 2, ni // 2, ni % 2
        return (hr + oddr) * (hc + oddc) + (hr + oddr) * hc + hr * hc + hr * (hc + oddc), 2 * hr
    def tot(parts, nsplit_ni):
        fma = sum(p[0] for p in parts)
        add = sum(p[1] for p in parts) + sum(2 * split(ni) for ni in nsplit_ni)
        return 2 * fma + add
    x = n1 * n1 * tot([eo(q, n1, 1), eo(q, n1, -1)], [n1])
    y = q * n1 * tot([eo(q, n1, 1), eo(q, n1, -1), eo(q, n1, 1)], [n1, n1])
    zf = q * q * tot([eo(q, n1, 1), eo(q, n1, 1), eo(q, n1, -1)] + ([eo(q, n1, 1)] if mass else []), [n1] * 3)
    zr = q * q * tot([eo(n1, q, 1), eo(n1, q, 1), eo2(n1, q) if mass else eo(n1, q, -1)], [q] * (4 if mass else 3))
    yt = q * n1 * tot([eo(n1, q, 1), eo2(n1, q)], [q] * 3)
    xt = n1 * n1 * tot([eo2(n1, q)], [q] * 2)
    return float(x + y + zf + zr + yt + xt)


def apply_flops_per_element(n1, q, mass, mode):
    """Flops of one element's diffusion (+ mass) apply in the given
    quadrature-data mode ("general", "cartesian" or "geometry"; the geometry
    kernel contracts in the even-odd form)."""
    if mode == "geometry":
        return sumfac_flops_even_odd(n1, q, mass) + q ** 3 * (geometry_flops_per_point(q) + (2.0 if mass else 0.0))
    return 2.0 * sumfac_fma_per_element(n1, q, True, mass, mode == "cartesian")

WORKLOADS = {
    "q3-96-learned": dict(p=3, n=96, form="learned", rtol=1e-8, coarsen=16,
                          desc="learned preconditioner, 3-D diffusion {n}^3 hex Q{p} ({N} DOFs), kappa = exp(GRF) "
                               "per field, random nodal RHS, Dirichlet, FP64 flexible CG (Chebyshev-Jacobi(4) + "
                               "U-Net 16/32/64 BF16) vs PCG + Chebyshev-Jacobi(4) to 1e-8 on held-out fields"),
    "p1-256-csr": dict(p=1, n=256, form="csr", rtol=1e-10,
                       desc="2-D Poisson -lap u = 2 pi^2 sin(pi x) sin(pi y) on the unit square, {n}^2 quads "
                            "split into 2 triangles, P1 ({N} DOFs, nnz {nnz}), assembled CSR, homogeneous "
                            "Dirichlet by row zeroing, FP64 PCG + Jacobi to 1e-10"),
    "q2-64-sinprod": dict(p=2, n=64, form="diffusion", kappa="sinprod", amp=0.0, rhs="unit",
                          rtol=1e-8, mass_apply=True, chunk=0,
                          desc="3-D diffusion, {n}^3 hex Q{p} ({N} DOFs), kappa = 1 + 0.5 sin sin sin, "
                               "Dirichlet, FP64 PA ({q}^3 Gauss points), PCG + Chebyshev-Jacobi(4) to 1e-8"),
    "q3-96-grf": dict(p=3, n=96, form="diffusion", kappa=("grf", 7), amp=0.0, rhs=("random", 11),
                      rtol=1e-10, mass_apply=False, chunk=40,
                      desc="3-D diffusion, {n}^3 hex Q{p} ({N} DOFs), kappa = exp(GRF, l = 0.1, sigma = 1.5) "
                           "in [1e-2, 1e2] (seed 7), random nodal RHS (seed 11), Dirichlet, FP64 PA "
                           "({q}^3 Gauss points), PCG + Chebyshev-Jacobi(4) to 1e-10"),
    "q4-128-perturbed": dict(p=4, n=128, form="diff_mass", kappa=("grf", 7), amp=0.2, rhs=("random", 11),
                             learned_degree=2, learned_truncation=4,
                             rtol=1e-8, mass_apply=False, chunk=40,
                             desc="3-D diffusion + mass (rho = 1), {n}^3 hex Q{p} ({N} DOFs), vertices "
                                  "perturbed <= 0.2 h (seed 5), kappa = exp(GRF) in [1e-2, 1e2] (seed 7), "
                                  "random nodal RHS (seed 11), Dirichlet, FP64 PA ({q}^3 Gauss points), "
                                  "PCG + Chebyshev-Jacobi(4) to 1e-8"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="q4-128-perturbed", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override elements per direction")
    ap.add_argument("--p", type=int, default=0, help="override the order")
    ap.add_argument("--chunk", type=int, default=-1,
                    help="PCG iterations per step (0: one whole solve per step; default: the workload's)")
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (counter calibration)")
    ap.add_argument("--no-legs", action="store_true", help="skip the whole-solve comparison")
    ap.add_argument("--tight-rtol", type=float, default=1e-12,
                    help="legs: tolerance of the continued solves (error target 1e-10)")
    ap.add_argument("--ref-rtol", type=float, default=1e-13, help="legs: tolerance of the FP64 reference")
    ap.add_argument("--train-steps", type=int, default=1200)
    ap.add_argument("--truncation", type=int, default=0,
                    help="flexible CG: directions kept (0: the workload's)")
    ap.add_argument("--learned-degree", type=int, default=0,
                    help="learned legs: Chebyshev-Jacobi degree of the smoothing inside the flexible-CG "
                         "preconditioner (0: the workload's; the network is trained on degree-4 smoothed pairs)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--learned-only", action="store_true",
                    help="run only the learned-preconditioner leg (no reference solve: no errors or ratios)")
    ap.add_argument("--train-fields", type=int, default=256, help="q3-96-learned: training fields (seeds 0, 1, ...)")
    ap.add_argument("--test-fields", type=int, default=32, help="q3-96-learned: held-out fields (seeds 256, ...)")
    ap.add_argument("--epochs", type=int, default=1, help="q3-96-learned: passes over the training fields")
    ap.add_argument("--steps-per-field", type=int, default=4, help="q3-96-learned: training steps per field")
    ap.add_argument("--unet-passes", type=int, default=0,
                    help="time only the U-Net: this many forward + backward passes on a --unet-grid^3 grid")
    ap.add_argument("--unet-grid", type=int, default=129)
    ap.add_argument("--sumfac-probe", action="store_true",
                    help="time only the batched sum-factorisation sweep on the vector unit and on MFMA "
                         "(femgpu_sumfac_probe) at the Q3 and Q4 shapes, FP64 and FP32")
    ap.add_argument("--probe-blocks", type=int, default=4096)
    ap.add_argument("--probe-reps", type=int, default=64)
    ap.add_argument("--coarsen", type=int, default=0, help="learned: network grid coarsening (0: the workload's, else the smallest "
                    "divisor leaving at most 64 intervals per direction)")
    return ap.parse_args()


def unit_square_mesh(n):
    """The config-1 mesh, as the caller builds it: the unit square with n x n
    quads, each split into (v0, v1, v3) and (v1, v2, v3) (mesh/src/simplex.rs:408-413);
    P1 DOFs are the vertices.  -> (coords f64[(n+1)^2, 2], conn u32[2 n^2, 3],
    boundary vertices)."""
    m = n + 1
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    xy = np.stack([i.ravel() / n, j.ravel() / n], axis=1)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    n0 = (jj * m + ii).ravel(); n1 = n0 + 1; n3 = n0 + m; n2 = n3 + 1
    conn = np.empty((2 * n0.size, 3), np.uint32)
    conn[0::2] = np.stack([n0, n1, n3], 1)
    conn[1::2] = np.stack([n1, n2, n3], 1)
    bnd = np.where((i.ravel() == 0) | (i.ravel() == n) | (j.ravel() == 0) | (j.ravel() == n))[0]
    return xy, conn, bnd.astype(np.uint32)


def run_csr(a, w):
    """Config-1 workload: the system assembled on the GPU (P1 stiffness, 3-point
    source load, row zeroing), then PCG + Jacobi on the GPU."""
    import femgpu as fg
    n = a.n or w["n"]
    xy, conn, bnd = unit_square_mesh(n)
    ctx = fg.default_context()
    N = len(xy)

    def assemble():
        A = fg.assemble_bilinear(2, 1, xy, conn, conn, N, fg.DIFFUSION, 1.0, 3, ctx=ctx)
        b = fg.DeviceArray(ctx, N, np.float64)
        fg.assemble_linear(2, 1, xy, conn, conn, N, fg.SOURCE_SINPROD, (2.0 * np.pi) * np.pi, 3, out=b, ctx=ctx)
        fg.apply_dirichlet(A, b, bnd, np.zeros(len(bnd)))
        return A, b

    assemble()  # warm-up (code objects, allocator)
    ctx.synchronize()
    t0 = time.perf_counter()
    A, b = assemble()
    ctx.synchronize()
    t_asm = time.perf_counter() - t0
    nnz = A.nnz()
    x = fg.DeviceArray(ctx, N, np.float64)
    x0 = fg.DeviceArray(ctx, N, np.float64)
    cfg = fg.SolverConfig(rtol=w["rtol"], max_iter=5000)
    xv = fg.DeviceArray(ctx, N, np.float64).upload(np.random.default_rng(0).standard_normal(N))
    yv = fg.DeviceArray(ctx, N, np.float64)
    A.time_apply(xv, yv, 5)
    ms_spmv, _ = A.time_apply(xv, yv, a.apply_reps)
    spmv_bytes = 12.0 * nnz + 8.0 * (N + 1) + 16.0 * N

    def step():
        x.copy_from(x0)
        return fg.solve_pcg_jacobi(A, b, x, cfg)

    for _ in range(a.warmup):
        step()
    times, iters = [], []
    t_all = time.perf_counter()
    for _ in range(a.steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        res = step()
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
        iters.append(res.iterations)
        if not res.converged:
            print(f"bench.py: solve did not converge ({res})", file=sys.stderr)
            sys.exit(1)
    total = time.perf_counter() - t_all
    xs = x.download()
    err = float(np.abs(xs - np.sin(np.pi * xy[:, 0]) * np.sin(np.pi * xy[:, 1])).max())
    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import femref as fr
        t0 = time.perf_counter()
        rp, ci, v, rhs, _, _ = fr.poisson_p1(n)
        t_cpu_asm = time.perf_counter() - t0
        ref = fr.CsrRef(rp, ci, v)
        t0 = time.perf_counter()
        _, ro = fr.pcg_jacobi(fr._Op(N, csr=ref), ref.diagonal(), rhs, np.zeros(N), rtol=w["rtol"], max_iter=5000)
        dt = time.perf_counter() - t0
        cpu = {"value": N * ro.iterations / dt / 1e9, "unit": "GDOF/s", "cores": 1, "kind": "port",
               "sample": f"oracle PCG-Jacobi (par_solver.rs semantics) with the reference's 4-way CSR row sum, "
                         f"full solve, {ro.iterations} iterations ({dt:.2f} s); the oracle's assembly of the "
                         f"same system (element loop, COO sort, row zeroing) took {1e3 * t_cpu_asm:.1f} ms",
               "assembly_ms": round(1e3 * t_cpu_asm, 2)}
    t_med = float(np.median(times))
    # the stronger preconditioner of the path's next row: smoothed-aggregation AMG
    ctx.synchronize(); t0 = time.perf_counter()
    amg = fg.AmgSolver(A)
    ctx.synchronize(); t_setup = time.perf_counter() - t0
    ta = []
    for k in range(1 + a.steps):
        x.copy_from(x0)
        ctx.synchronize(); t0 = time.perf_counter()
        ra = amg.solve(b, x, cfg)
        ctx.synchronize()
        if k:
            ta.append(time.perf_counter() - t0)
    amg_leg = {"method": "PCG + smoothed-aggregation AMG V-cycle (Chebyshev-Jacobi(2) smoothing)",
               "levels": amg.n_levels(), "setup_ms": round(1e3 * t_setup, 2),
               "solve_time_ms": round(1e3 * float(np.median(ta)), 3), "iterations": ra.iterations,
               "rel_residual": ra.final_residual,
               "speedup_vs_pcg_jacobi": round(t_med / float(np.median(ta)), 2)}
    amg.close()
    # simplex PA applies (PADiffusionOperator / PAMassOperator, partial.rs:65-209) on a
    # finer mesh of the same kind: one thread per element, per-element qdata, atomics into y
    ns = 2048
    xy2, conn2, bnd2 = unit_square_mesh(ns)
    N2, E2 = len(xy2), len(conn2)
    xs2 = fg.DeviceArray(ctx, N2, np.float64).upload(np.random.default_rng(1).standard_normal(N2))
    ys2 = fg.DeviceArray(ctx, N2, np.float64)
    simplex = {"mesh": f"{ns}^2 unit square, P1 triangles ({E2} elements, {N2} DOFs)"}
    for name, Op, kw, s_e in (("diffusion", fg.PADiffusionOperator, {"kappa": 1.0}, 3),
                              ("mass", fg.PAMassOperator, {"rho": 1.0}, 1)):
        op = Op(2, 1, xy2, conn2, conn2, N2, quad_order=3, ctx=ctx, **kw)
        op.time_apply(xs2, ys2, 5)
        ms, _ = op.time_apply(xs2, ys2, a.apply_reps)
        byts = 16.0 * N2 + (8.0 * s_e + 12.0) * E2  # x, y; qdata values and 3 DOF indices per element
        simplex[name] = {"ms": round(ms, 5), "gdofs": round(N2 / (ms * 1e-3) / 1e9, 3),
                         "algorithmic_bytes": byts, "gb_s": round(byts / (ms * 1e-3) / 1e9, 1),
                         "frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        op.close()
    # CSR SpMV (csr.rs:140-193), the reference Chebyshev smoother and the spectral
    # radius estimate (amg/src/lib.rs:53-123) at a bandwidth-bound size: the
    # 2048^2 P1 stiffness matrix, assembled on the GPU
    Ab = fg.assemble_bilinear(2, 1, xy2, conn2, conn2, N2, fg.DIFFUSION, 1.0, 3, ctx=ctx)
    Ab.apply_dirichlet_row_zeroing(bnd2, np.zeros(len(bnd2)))  # as for the 2-D P1 Poisson system: A 1 != 0, rho > 0
    nnz2 = Ab.nnz()
    Ab.time_apply(xs2, ys2, 5)
    ms_sp, _ = Ab.time_apply(xs2, ys2, a.apply_reps)
    sp_bytes = 12.0 * nnz2 + 8.0 * (N2 + 1) + 16.0 * N2
    ones = fg.DeviceArray(ctx, N2, np.float64, np.ones(N2))
    xc = fg.DeviceArray(ctx, N2, np.float64)
    fg.estimate_spectral_radius(Ab, 2)
    ctx.synchronize(); t0 = time.perf_counter()
    rho = fg.estimate_spectral_radius(Ab, 20)
    ctx.synchronize(); t_rho = (time.perf_counter() - t0) / 20
    lmax = 1.1 * rho
    fg.chebyshev_smooth(Ab, xc, ones, lmax / 30.0, lmax, 2)
    ctx.synchronize(); t0 = time.perf_counter()
    fg.chebyshev_smooth(Ab, xc, ones, lmax / 30.0, lmax, 20)
    ctx.synchronize(); t_cheb = (time.perf_counter() - t0) / 20
    cheb_bytes = sp_bytes + 64.0 * N2  # A x, then b, r, dinv, d, x read and r, d, x written
    csr_scale = {"matrix": f"2048^2 P1 stiffness assembled on the GPU, Dirichlet rows zeroed ({N2} rows, "
                           f"{nnz2} nonzeros)",
                 "spmv": {"ms": round(ms_sp, 5), "algorithmic_bytes": sp_bytes,
                          "gb_s": round(sp_bytes / (ms_sp * 1e-3) / 1e9, 1),
                          "frac": round(sp_bytes / (ms_sp * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
                 "chebyshev_smooth_sweep": {"ms": round(1e3 * t_cheb, 4), "algorithmic_bytes": cheb_bytes,
                                            "gb_s": round(cheb_bytes / t_cheb / 1e9, 1),
                                            "frac": round(cheb_bytes / t_cheb / 1e9 / HBM_PEAK_GBS, 4)},
                 "spectral_radius_iteration_ms": round(1e3 * t_rho, 4), "spectral_radius": rho}
    # AMG-PCG against PCG-Jacobi on the same matrix, right-hand side 1 on the free rows
    b2h = np.ones(N2)
    b2h[bnd2] = 0.0
    b2 = fg.DeviceArray(ctx, N2, np.float64, b2h)
    x2 = fg.DeviceArray(ctx, N2, np.float64)
    z2 = fg.DeviceArray(ctx, N2, np.float64)
    cfg2 = fg.SolverConfig(rtol=1e-8, max_iter=20000)
    ctx.synchronize(); t0 = time.perf_counter()
    amg2 = fg.AmgSolver(Ab)
    ctx.synchronize(); t_setup2 = time.perf_counter() - t0
    ta2 = []
    for k in range(3):
        x2.copy_from(z2)
        ctx.synchronize(); t0 = time.perf_counter()
        ra2 = amg2.solve(b2, x2, cfg2)
        ctx.synchronize()
        if k:
            ta2.append(time.perf_counter() - t0)
    levels2 = amg2.n_levels()
    amg2.close()
    x2.copy_from(z2)
    ctx.synchronize(); t0 = time.perf_counter()
    rj2 = fg.solve_pcg_jacobi(Ab, b2, x2, cfg2)
    ctx.synchronize(); t_j2 = time.perf_counter() - t0
    csr_scale["amg_pcg_vs_pcg_jacobi"] = {
        "rhs": "1 on the free rows, 0 on the Dirichlet rows; rtol 1e-8",
        "amg_levels": levels2,
        "amg_setup_ms": round(1e3 * t_setup2, 1), "amg_solve_ms": round(1e3 * float(np.median(ta2)), 2),
        "amg_iterations": ra2.iterations, "amg_converged": bool(ra2.converged),
        "pcg_jacobi_solve_ms": round(1e3 * t_j2, 1), "pcg_jacobi_iterations": rj2.iterations,
        "pcg_jacobi_converged": bool(rj2.converged),
        "speedup_solve": round(t_j2 / float(np.median(ta2)), 1)}
    del b2, x2, z2
    Ab.close()
    del xs2, ys2, ones, xc
    out = {
        "metric": "solve time to 1e-8 rel.res + PA apply GDOF/s; plain FP64 vs mixed+learned",
        "value": round(N * sum(iters) / total / 1e9, 4), "unit": "GDOF/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["desc"].format(n=n, N=N, nnz=nnz), "name": a.workload, "n_dofs": N, "nnz": nnz,
                   "global_batch": 1, "parallelism": "single GPU"},
        "assembly": {"gpu_ms": round(1e3 * t_asm, 2),
                     "what": "P1 stiffness (element matrices, stable radix sort of the COO triplets, run sums), "
                             "source load, row zeroing; mesh arrays from the host"},
        "solve": {"iterations": int(iters[-1]), "solve_time_ms": round(1e3 * t_med, 4),
                  "us_per_iteration": round(1e6 * t_med / max(iters[-1], 1), 2),
                  "rel_residual": res.final_residual, "max_nodal_error_vs_sin_sin": err},
        "spmv": {"ms": ms_spmv, "bytes": spmv_bytes, "gb_s": round(spmv_bytes / (ms_spmv * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "achieved": round(spmv_bytes / (ms_spmv * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_bytes / (ms_spmv * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "CSR SpMV k_spmv_row<double> (launch-latency-bound at this size)",
                     "algorithmic_bytes": spmv_bytes},
        "legs": {"amg_pcg": amg_leg},
        "pa_simplex": simplex,
        "csr_at_scale": csr_scale,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def device_copy_gbs(fg, ctx, mib=1024, reps=5):
    """Bandwidth of the runtime's device-to-device copy of `mib` MiB (read +
    write bytes over HIP-event time): the practical streaming ceiling that the
    roofline fraction can be read against."""
    nc = mib * (1 << 20) // 8
    src = fg.DeviceArray(ctx, nc, np.float64)
    dst = fg.DeviceArray(ctx, nc, np.float64)
    dst.copy_from(src)
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(reps):
        dst.copy_from(src)
    ms = ctx.timer_stop() / reps
    del src, dst
    return 2.0 * nc * 8 / (ms * 1e-3) / 1e9


def run_learned(a, w):
    """The learned preconditioner on the Q3 96^3 GRF family.  The
    U-Net (placed on the DOF grid coarsened by --coarsen, default as for the
    other workloads) is trained on --train-fields fields (kappa seeds 0, 1, ...;
    each field in turn: its operator is built, pairs are harvested from
    PCG-Chebyshev iterates of a solve with a known solution, --steps-per-field
    energy-loss steps are taken, and the field is freed) and compared with
    PCG + Chebyshev-Jacobi(4) on --test-fields held-out fields (seeds 256,
    257, ...), FP64 flexible CG to 1e-8 judged on the recomputed residual;
    value = N x iterations / time of the learned solves."""
    import femgpu as fg
    from femgpu import learned
    n = a.n or w["n"]
    p = a.p or w["p"]
    ctx = fg.default_context()
    say = lambda m: print(f"bench.py: {m}", file=sys.stderr, flush=True)

    def field(seed):
        K = fg.HexPAOperator(p, p + 1, n, form=fg.DIFFUSION, kappa=fg.grf(seed), dirichlet=True, ctx=ctx)
        b = fg.DeviceArray(ctx, K.n_dofs(), np.float64)
        K.rhs_random(1000 + seed, b)
        x = fg.DeviceArray(ctx, K.n_dofs(), np.float64)
        _, lmax = fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=1e-8, max_iter=1))
        return K, b, lmax

    K0, _, _ = field(0)
    cf = a.coarsen or w.get("coarsen") or coarsening(K0)
    net = fg.UNet.for_operator(K0, seed=11, coarsen=cf)
    K0.close()
    rng = np.random.default_rng(0)
    losses = []
    steps = 0
    t_harvest = t_train = 0.0
    t_all = time.perf_counter()
    for epoch in range(a.epochs):
        for sd in range(a.train_fields):
            t0 = time.perf_counter()
            K, b, lmax = field(sd)
            pairs = learned.harvest_pairs_known(K, 5000 + 1000 * epoch + sd, iters=(2, 6, 15, 40), lmax=lmax)
            ctx.synchronize()
            t1 = time.perf_counter()
            t_harvest += t1 - t0
            net.aux_from(K)
            for _ in range(a.steps_per_field):
                r_, e_ = pairs[rng.integers(len(pairs))]
                total = a.epochs * a.train_fields * a.steps_per_field
                losses.append(net.train_step_energy(K, r_, e_, 2e-3 * 0.1 ** (steps / max(total, 1))))
                steps += 1
            ctx.synchronize()
            t_train += time.perf_counter() - t1
            for r_, e_ in pairs:
                r_.free(); e_.free()
            K.close()
            del b
            if sd % 32 == 31:
                say(f"epoch {epoch}: {sd + 1} fields, {steps} steps, loss {np.mean(losses[-64:]):.3f}, "
                    f"{time.perf_counter() - t_all:.0f} s")
    cfg = fg.SolverConfig(rtol=w["rtol"], max_iter=100000)
    rows, it_l, t_l_sum, it_p, t_p_sum = [], 0, 0.0, 0, 0.0
    N = 0
    for j in range(a.test_fields):
        sd = 256 + j
        K, b, lmax = field(sd)
        N = K.n_dofs()
        x = fg.DeviceArray(ctx, N, np.float64)
        fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=1e-8, max_iter=3), lmax=lmax)  # warm
        x0 = fg.DeviceArray(ctx, N, np.float64)
        x.copy_from(x0)
        ctx.synchronize(); t0 = time.perf_counter()
        rp, _ = fg.solve_pcg_chebyshev(K, b, x, cfg, lmax=lmax)
        ctx.synchronize(); tp = time.perf_counter() - t0
        net.aux_from(K)
        x.copy_from(x0)
        ctx.synchronize(); t0 = time.perf_counter()
        rl, info = fg.solve_fcg_learned(K, net, b, x, cfg, lmax=lmax, truncation=a.truncation,
                                        degree=a.learned_degree)
        ctx.synchronize(); tl = time.perf_counter() - t0
        r = fg.DeviceArray(ctx, N, np.float64)
        K.apply_zero(x, r)
        bs = b.download()
        relres = float(np.linalg.norm(bs - r.download()) / np.linalg.norm(bs))
        rows.append({"seed": sd, "pcg_chebyshev_iterations": rp.iterations, "pcg_chebyshev_ms": round(1e3 * tp, 2),
                     "learned_fcg_iterations": rl.iterations, "learned_fcg_ms": round(1e3 * tl, 2),
                     "net_corrections": info.net_corrections, "fell_back": info.fell_back,
                     "learned_rel_residual_fp64": relres,
                     "iteration_ratio": round(rp.iterations / max(rl.iterations, 1), 3),
                     "time_ratio": round(tp / tl, 3)})
        say(f"held-out field {sd}: PCG-Chebyshev {rp.iterations} it {1e3 * tp:.0f} ms, "
            f"learned FCG {rl.iterations} it {1e3 * tl:.0f} ms")
        it_l += rl.iterations
        t_l_sum += tl
        it_p += rp.iterations
        t_p_sum += tp
        K.close()
        del r, x, x0, b
    out = {
        "metric": "solve time to 1e-8 rel.res + PA apply GDOF/s; plain FP64 vs mixed+learned",
        "value": round(N * it_l / t_l_sum / 1e9, 4), "unit": "GDOF/s", "n_gpus": 1,
        "steps": a.test_fields, "warmup": 0, "ms_per_step": round(1e3 * t_l_sum / a.test_fields, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["desc"].format(n=n, p=p, N=N), "name": a.workload, "n_dofs": N,
                   "train_fields": a.train_fields, "test_fields": a.test_fields, "global_batch": 1,
                   "parallelism": "single GPU"},
        "train": {"epochs": a.epochs, "steps": steps, "steps_per_field": a.steps_per_field,
                  "coarsening": cf, "network_grid": list(net.shape), "harvest_s": round(t_harvest, 1),
                  "train_s": round(t_train, 1), "truncation": a.truncation,
                  "pairs_from": "PCG-Chebyshev iterates (2, 6, 15, 40) of a solve with b = A x*, x* a seeded "
                                "uniform field, per training field",
                  "energy_loss_first_last": [round(float(np.mean(losses[:64])), 4),
                                             round(float(np.mean(losses[-64:])), 4)]},
        "held_out_summary": {"pcg_chebyshev_iterations": it_p, "learned_fcg_iterations": it_l,
                             "pcg_chebyshev_s": round(t_p_sum, 3), "learned_fcg_s": round(t_l_sum, 3),
                             "iteration_ratio": round(it_p / max(it_l, 1), 3),
                             "time_ratio": round(t_p_sum / t_l_sum, 3)},
        "held_out": rows,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def make_operator(fg, ctx, w, n, p):
    form = fg.DIFFUSION if w["form"] == "diffusion" else fg.DIFF_MASS
    kappa = w["kappa"] if isinstance(w["kappa"], str) else fg.grf(w["kappa"][1])
    verts = fg.hex_vertices(n, n, n, amp=w["amp"], seed=5) if w["amp"] > 0 else None
    return fg.HexPAOperator(p, p + 1, n, form=form, kappa=kappa, rho=1.0, dirichlet=True, vertices=verts,
                            ctx=ctx)


def committed_traffic(name, n, p, qdata_per_point):
    """HBM bytes per apply from the committed counter summary of this
    workload (its line 'traffic_bytes_per_apply: B'), or None; the summary's
    'qdata_per_point: s' line must match the operator being measured."""
    path = os.path.join(ROOT, "profiles", f"pa_hex_{name}_hbm_traffic.md")
    w = WORKLOADS[name]
    if not os.path.exists(path) or (n, p) != (w["n"], w["p"]):
        return None
    text = open(path).read()
    m = re.search(r"traffic_bytes_per_apply:\s*([0-9]+)", text)
    sq = re.search(r"qdata_per_point:\s*([0-9]+)", text)
    if not m or not sq or int(sq.group(1)) != qdata_per_point:
        return None
    return float(m.group(1))


def cpu_info():
    """Host CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def best_of(fn, reps=5):
    fn()  # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return min(ts)


def cpu_baseline(w, p, N_full, E_full, iters, lmax):
    """The oracle on the host cores, on bounded samples of the workload's
    element type (same order, Gauss points, coefficient and mesh generator):
      - reference-algorithm apply: the reference's PA element loop
        (partial.rs:151-208) generalised to hex, on an 8^3 sample, 1 core
        (as the reference runs it) and all cores (parity colours);
      - optimised apply: the sum-factorised oracle on stored qdata, 32^3
        sample, 1 core and all cores;
      - solve rate: oracle PCG + Chebyshev-Jacobi(4) on the 32^3 sample, all
        cores, `iters` iterations (the unit of the GPU value: N x iterations /
        time).
    Apply rates are per element, converted to the full mesh's GDOF/s
    (N_full / (time per element x E_full)); best of 5 after a warm-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import femref as fr
    model, nproc, avail = cpu_info()
    kind = fr.FR_DIFFUSION if w["form"] == "diffusion" else fr.FR_DIFF_MASS
    if isinstance(w["kappa"], str):
        kap = fr.coef_sinprod()
    else:
        kap = fr.coef_grf(w["kappa"][1])

    def sample(ns):
        d = fr.hex_desc(p, p + 1, (ns, ns, ns), kind=kind, kappa=kap, rho=1.0, dirichlet=True)
        c = fr.hex_vertices(ns, ns, ns, w["amp"], 5)
        return d, c, fr.hex_ndofs(d), ns ** 3

    full = lambda t, E: float(f"{N_full / (t / E * E_full) / 1e9:.4g}")
    d8, c8, N8, E8 = sample(8)
    x8 = np.random.default_rng(0).standard_normal(N8)
    t_ref1 = best_of(lambda: fr.hex_apply_refloop(d8, c8, x8))
    t_refa = best_of(lambda: fr.hex_apply_refloop(d8, c8, x8, nthreads=avail))
    d32, c32, N32, E32 = sample(32)
    Dq, Mq = fr.hex_qdata(d32, c32)
    x32 = np.random.default_rng(0).standard_normal(N32)
    t_sf1 = best_of(lambda: fr.hex_apply_sumfac(d32, Dq, Mq, x32, nthreads=1))
    t_sfa = best_of(lambda: fr.hex_apply_sumfac(d32, Dq, Mq, x32, nthreads=avail))
    op = fr._Op(N32, hexop=(d32, Dq, Mq, avail))
    diag = fr.hex_diagonal(d32, Dq, Mq)
    b = np.random.default_rng(11).uniform(-1.0, 1.0, N32)
    lm = 1.1 * fr.estimate_lmax_jacobi(op, diag, 20)
    x0 = np.zeros(N32)
    fr.pcg_chebyshev(op, diag, 4, lm / 30, lm, b, x0, rtol=1e-30, max_iter=1)  # warm-up
    t0 = time.perf_counter()
    _, res = fr.pcg_chebyshev(op, diag, 4, lm / 30, lm, b, x0, rtol=1e-30, max_iter=iters)
    dt = time.perf_counter() - t0
    return {"value": float(f"{N32 * res.iterations / dt / 1e9:.4g}"), "unit": "GDOF/s", "cores": avail,
            "kind": "port",
            "sample": f"oracle PCG + Chebyshev-Jacobi(4) (sum-factorised FP64 PA, OpenMP) on a 32^3 "
                      f"sample of the workload's Q{p} mesh family ({N32} DOFs), {res.iterations} "
                      f"iterations in {dt:.2f} s on {avail} threads",
            "cpu_model": model, "nproc": nproc, "cpus_available": avail,
            "apply_gdofs_full_mesh_equiv": {
                "reference_algorithm_1_core": full(t_ref1, E8),
                "reference_algorithm_all_cores": full(t_refa, E8),
                "optimised_sumfac_1_core": full(t_sf1, E32),
                "optimised_sumfac_all_cores": full(t_sfa, E32),
                "what": "element loop of partial.rs:151-208 generalised to hex (8^3 sample) and the "
                        "sum-factorised oracle (32^3 sample); best of 5; per-element time scaled to "
                        f"the full mesh ({E_full} elements, {N_full} DOFs)"}}


def coarsening(K):
    """Coarsening factor of the learned preconditioner's grid: the smallest
    divisor cf of the DOF grid's (Nd - 1) that leaves at most 64 intervals per
    direction (Q4 128^3: 512 -> cf 8, a 65^3 network grid)."""
    Nd = K.shape[0] * K.p + 1
    for cf in range(1, Nd):
        if (Nd - 1) % cf == 0 and (Nd - 1) // cf <= 64:
            return cf
    return 1


def legs(fg, ctx, K, b, x, lmax, rtol, a):
    """Whole solves on the same operator and right-hand side, each once,
    timed on the host around device-resident solves: plain FP64 to rtol and
    continued to --tight-rtol; mixed precision (continued refinement) to rtol
    and on to --tight-rtol; mixed + learned to rtol.  Errors are relative
    2-norm errors against the FP64 reference (the plain solve continued to
    --ref-rtol)."""
    from femgpu import learned
    N = K.n_dofs()
    x0 = fg.DeviceArray(ctx, N, np.float64)
    bh = b.download()
    bn = float(np.linalg.norm(bh))
    r = fg.DeviceArray(ctx, N, np.float64)

    def relres():
        K.apply_zero(x, r)
        return float(np.linalg.norm(bh - r.download()) / bn)

    def timed(fn, start=None):
        x.copy_from(start if start is not None else x0)
        ctx.synchronize()
        t0 = time.perf_counter()
        out = fn()
        ctx.synchronize()
        return out, time.perf_counter() - t0

    cfg = lambda tol: fg.SolverConfig(rtol=tol, max_iter=200000)
    say = lambda m: print(f"bench.py: {m}", file=sys.stderr, flush=True)
    if a.learned_only:
        # the learned leg alone (for profiling it): no reference solve, so no
        # errors and no ratios
        out, t8, t12 = {}, None, None
        err = lambda v: None
        K32 = K.convert(fg.F32)
    else:
        pcg = lambda tol: fg.solve_pcg_chebyshev(K, b, x, cfg(tol), lmax=lmax)[0]
        # plain FP64: to rtol, then on to the tight tolerance, then the reference
        res8, t8 = timed(lambda: pcg(rtol))
        x8 = fg.DeviceArray(ctx, N, np.float64); x8.copy_from(x)
        rr8 = relres()
        say(f"plain FP64 to {rtol:g}: {res8.iterations} it, {t8:.2f} s")
        res12, t12 = timed(lambda: pcg(a.tight_rtol), start=x8)
        x12 = fg.DeviceArray(ctx, N, np.float64); x12.copy_from(x)
        rr12 = relres()
        say(f"plain FP64 continued to {a.tight_rtol:g}: {res12.iterations} it, {t12:.2f} s")
        resr, tr = timed(lambda: pcg(a.ref_rtol), start=x12)
        rrr = relres()
        xref = x.download()
        say(f"FP64 reference to {a.ref_rtol:g}: {resr.iterations} it, {tr:.2f} s, true residual {rrr:.2e}")
        nref = float(np.linalg.norm(xref))
        err = lambda v: float(np.linalg.norm(v.download() - xref) / nref)
        e8, e12 = err(x8), err(x12)
        del x8, x12
        out = {"reference": {"method": f"plain FP64 continued to {a.ref_rtol:g} (restart with a recomputed "
                                       "residual)", "iterations": resr.iterations, "time_s": round(tr, 2),
                             "rel_residual_fp64": rrr},
               "plain_fp64": {"method": "FP64 PCG + Chebyshev-Jacobi(4)", "rtol": rtol,
                              "solve_time_ms": round(1e3 * t8, 2), "iterations": res8.iterations,
                              "rel_residual_fp64": rr8, "rel_error": e8},
               "plain_fp64_tight": {"method": f"the plain solve continued from its {rtol:g} solution to "
                                             f"{a.tight_rtol:g}",
                                    "time_to_tight_ms": round(1e3 * (t8 + t12), 2),
                                    "iterations": res8.iterations + res12.iterations,
                                    "rel_residual_fp64": rr12, "rel_error": e12,
                                    "error_le_1e-10": e12 <= 1e-10}}
        K32 = K.convert(fg.F32)
        mix = lambda tol: fg.solve_ir_mixed(K, K32, b, x, cfg(tol), lmax=lmax, mode=1, inner_rtol=1e-2,
                                            update_interval=16)
        (rm, im), tm = timed(lambda: mix(rtol))
        xm = fg.DeviceArray(ctx, N, np.float64); xm.copy_from(x)
        rrm, em = relres(), err(x)
        say(f"mixed to {rtol:g}: {im} FP32 it, {tm:.2f} s")
        (rm2, im2), tm2 = timed(lambda: mix(a.tight_rtol), start=xm)
        rrm2, em2 = relres(), err(x)
        say(f"mixed continued to {a.tight_rtol:g}: {im2} FP32 it, {tm2:.2f} s")
        del xm
        out["mixed_fp32_ir"] = {
            "method": "one FP32 PCG + Chebyshev-Jacobi(4) (FP32 apply and smoother, FP64 dots) whose residual is "
                      "replaced by the FP64 residual b - A u after each 1e-2 reduction or 16 iterations "
                      "(u updated in FP64)",
            "solve_time_ms": round(1e3 * tm, 2), "fp32_iterations": int(im), "residual_replacements": rm.iterations,
            "rel_residual_fp64": rrm, "rel_error": em, "speedup_vs_plain_fp64": round(t8 / tm, 3)}
        out["mixed_fp32_ir_tight"] = {
            "method": f"the mixed solve continued from its {rtol:g} solution to {a.tight_rtol:g}",
            "time_to_tight_ms": round(1e3 * (tm + tm2), 2), "fp32_iterations": int(im + im2),
            "rel_residual_fp64": rrm2, "rel_error": em2, "error_le_1e-10": em2 <= 1e-10,
            "speedup_vs_plain_fp64_tight": round((t8 + t12) / (tm + tm2), 3)}
        if a.no_learned:
            return out
    # Learned preconditioner: the 3-level U-Net placed on the DOF grid coarsened by cf
    # (trilinear prolongation, restriction by its transpose), trained on the operator
    # family's FP32 copy with the energy-norm loss on pairs from PCG-Chebyshev iterates
    # of two solves whose solution is known by construction (b = A x*, x* seeded).
    t0 = time.perf_counter()
    cf = a.coarsen or coarsening(K)
    pairs = []
    for sd in (101, 102):
        pairs += learned.harvest_pairs_known(K32, sd, lmax=lmax)
    ctx.synchronize()
    t_harvest = time.perf_counter() - t0
    t0 = time.perf_counter()
    net = fg.UNet.for_operator(K, seed=11, coarsen=cf)
    rng = np.random.default_rng(0)
    losses = []
    for st in range(a.train_steps):
        r_, e_ = pairs[rng.integers(len(pairs))]
        losses.append(net.train_step_energy(K32, r_, e_, 2e-3 * 0.1 ** (st / max(a.train_steps, 1))))
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    for r_, e_ in pairs:
        r_.free(); e_.free()
    say(f"harvested {len(pairs)} pairs in {t_harvest:.1f} s, trained {a.train_steps} steps in {t_train:.1f} s "
        f"(network grid {net.shape})")
    lrn = lambda tol: fg.solve_ir_learned(K, K32, net, b, x, cfg(tol), inner_rtol=1e-4, lmax=lmax,
                                          truncation=a.truncation, degree=a.learned_degree)
    (res_l, inner_l, info_l), t_l = timed(lambda: lrn(rtol))
    xl = fg.DeviceArray(ctx, N, np.float64); xl.copy_from(x)
    rrl, el = relres(), err(x)
    say(f"mixed + learned to {rtol:g}: {inner_l} inner it, {t_l:.2f} s")
    (res_l2, inner_l2, info_l2), t_l2 = timed(lambda: lrn(a.tight_rtol), start=xl)
    rrl2, el2 = relres(), err(x)
    say(f"mixed + learned continued to {a.tight_rtol:g}: {inner_l2} inner it, {t_l2:.2f} s")
    del xl
    out["mixed_learned"] = {
        "method": f"FP64 refinement (inner solves to 1e-4) around FP32 flexible CG ({a.truncation} directions) "
                  f"whose preconditioner is Chebyshev-Jacobi({a.learned_degree}) followed by the U-Net's Galerkin-scaled "
                  f"correction; the U-Net (3 levels, 16/32/64 channels, BF16 MFMA) runs on every {cf}-th DOF "
                  f"node ({net.shape[0]}^3 grid), its input restricted and its output prolongated "
                  f"trilinearly; PCG-Chebyshev fallback on stagnation",
        "solve_time_ms": round(1e3 * t_l, 2), "outer_iterations": res_l.iterations,
        "inner_iterations": int(inner_l), "fcg_iterations": info_l.fcg_iterations,
        "fallback_iterations": info_l.fallback_iterations, "fell_back": info_l.fell_back,
        "rel_residual_fp64": rrl, "rel_error": el, "speedup_vs_plain_fp64": round(t8 / t_l, 3) if t8 else None,
        "speedup_vs_plain_fp64_with_training": round(t8 / (t_l + t_harvest + t_train), 3) if t8 else None,
        "train": {"pairs": len(pairs), "steps": a.train_steps, "harvest_s": round(t_harvest, 2),
                  "train_s": round(t_train, 2), "coarsening": cf, "network_grid": list(net.shape),
                  "pairs_from": "PCG-Chebyshev iterates (1, 3, 6, 10, 20, 35, 60) of two FP32 solves with "
                                "b = A x*, x* a seeded uniform field (seeds 101, 102), smoothed as the "
                                "preconditioner sees them",
                  "energy_loss_first_last": [round(float(np.mean(losses[:20])), 4),
                                             round(float(np.mean(losses[-20:])), 4)]}}
    out["mixed_learned_tight"] = {
        "method": f"the mixed + learned solve continued from its {rtol:g} solution to {a.tight_rtol:g}",
        "time_to_tight_ms": round(1e3 * (t_l + t_l2), 2), "inner_iterations": int(inner_l + inner_l2),
        "rel_residual_fp64": rrl2, "rel_error": el2, "error_le_1e-10": el2 <= 1e-10 if el2 is not None else None,
        "speedup_vs_plain_fp64_tight": round((t8 + t12) / (t_l + t_l2), 3) if t8 else None}
    net.close()
    K32.close()
    return out


def run_unet(a):
    """The learned preconditioner's network alone: --unet-passes forward +
    backward passes of the 3-level U-Net (16/32/64 channels) on a --unet-grid^3
    grid after 3 warm-up forwards, timed on the host around synchronised
    passes.  FEMGPU_CONV_K16=1 selects one tap per MFMA in the convolutions
    (for comparison with the default two)."""
    import femgpu as fg
    ctx = fg.default_context()
    n = a.unet_grid
    net = fg.UNet((n, n, n), seed=3, ctx=ctx)
    x = np.random.default_rng(0).standard_normal(net.V).astype(np.float32)
    dout = np.random.default_rng(1).standard_normal(net.V).astype(np.float32)
    for _ in range(3):
        net.forward(x)
    ctx.synchronize()
    tf = tb = 0.0
    for _ in range(a.unet_passes):
        t0 = time.perf_counter()
        net.forward(x)
        ctx.synchronize()
        t1 = time.perf_counter()
        net.backward(dout)
        ctx.synchronize()
        tf += t1 - t0
        tb += time.perf_counter() - t1
    k = max(a.unet_passes, 1)
    print(json.dumps({"unet_grid": [n, n, n], "passes": a.unet_passes,
                      "conv_mfma": "16x16x16" if os.environ.get("FEMGPU_CONV_K16") == "1" else "16x16x32",
                      "forward_ms_incl_host_copies": round(1e3 * tf / k, 3),
                      "backward_ms_incl_host_copies": round(1e3 * tb / k, 3)}))


def run_sumfac_probe(a):
    """One batched sweep Out = [B u | G u] (lines of p + 1 nodal values, q
    Gauss points; 256 LDS-resident lines per workgroup, --probe-reps
    repetitions per launch) on each engine: the vector unit with multiply-add
    chains, the vector unit in the even-odd form (the production element
    kernel's), and the matrix cores (v_mfma_f64_16x16x4f64 /
    v_mfma_f32_16x16x4f32, 16 lines per tile).  Useful flops count 2 (p + 1)
    2q per line and repetition; the padding an MFMA tile carries is not
    counted."""
    import femgpu as fg
    ctx = fg.default_context()
    rows = []
    for p, q in ((3, 4), (4, 5)):
        for name, dt, fdt in (("f64", np.float64, fg.F64), ("f32", np.float32, fg.F32)):
            lines = a.probe_blocks * 256
            x = fg.DeviceArray(ctx, lines * (p + 1), dt).upload(
                np.random.default_rng(0).standard_normal(lines * (p + 1)).astype(dt))
            y = fg.DeviceArray(ctx, lines * 2 * q, dt)
            flops = 2.0 * (p + 1) * 2 * q * lines * a.probe_reps
            for engine in fg.SUMFAC_ENGINES:
                ms = fg.sumfac_probe(p, q, fdt, engine, a.probe_blocks, a.probe_reps, x, y, launches=5, ctx=ctx)
                rows.append({"p": p, "q": q, "dtype": name, "engine": engine, "ms_per_launch": round(ms, 4),
                             "useful_tflops": round(flops / (ms * 1e-3) / 1e12, 2)})
                print(f"bench.py: probe p={p} q={q} {name} {engine}: {ms:.4f} ms, "
                      f"{rows[-1]['useful_tflops']} TFLOP/s useful", file=sys.stderr, flush=True)
            del x, y
    print(json.dumps({"sumfac_probe": rows, "blocks": a.probe_blocks, "lines_per_block": 256,
                      "reps_per_launch": a.probe_reps}))


def main():
    a = parse()
    if a.learned_degree <= 0:
        # measured per workload: the 65^3 network of the Q4 128^3 mesh leaves less
        # to the smoother (degree 2: 8-12% less time than 4 over three network
        # seeds); the 19^3 network of the Q3 96^3 family needs degree 4 (2.90x
        # against 2.37x less time than PCG-Chebyshev)
        a.learned_degree = WORKLOADS[a.workload].get("learned_degree", 4)
    if a.truncation <= 0:
        # Q4 128^3 at degree 2: 4 kept directions take 2-12% less time than 8 over
        # network seeds 11-13; the Q3 family (FP64 flexible CG) keeps 8
        a.truncation = WORKLOADS[a.workload].get("learned_truncation", 8)
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {a.gpus}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    import femgpu as fg

    if a.unet_passes > 0:
        return run_unet(a)
    if a.sumfac_probe:
        return run_sumfac_probe(a)
    w = WORKLOADS[a.workload]
    if w["form"] == "csr":
        return run_csr(a, w)
    if w["form"] == "learned":
        return run_learned(a, w)
    n = a.n or w["n"]
    p = a.p or w["p"]
    q = p + 1
    ctx = fg.default_context()
    K = make_operator(fg, ctx, w, n, p)
    N = K.n_dofs()
    E = n ** 3
    q3 = q ** 3
    s_comp = K.qdata_per_point()  # 6 / 7 general, 1 / 2 on an axis-aligned box mesh, 1 with the geometry on the fly
    qmode = K.qdata_mode()
    dev = lambda: fg.DeviceArray(ctx, N, np.float64)
    b = dev()
    if w["rhs"] == "unit":
        K.rhs_unit(b)
    else:
        K.rhs_random(w["rhs"][1], b)
    x = dev()

    if a.copy_calib_mb:
        nc = a.copy_calib_mb * (1 << 20) // 8
        src = fg.DeviceArray(ctx, nc, np.float64, np.ones(nc))
        dst = fg.DeviceArray(ctx, nc, np.float64)
        dst.copy_from(src)  # one copy kernel of known size: reads and writes nc*8 bytes
        del src, dst

    copy_gbs = device_copy_gbs(fg, ctx)

    # PA apply rates (device vectors, HIP events)
    xv = dev().upload(np.random.default_rng(0).standard_normal(N))
    yv = dev()
    K.time_apply(xv, yv, 5)
    ms_apply, ms_kern = K.time_apply(xv, yv, a.apply_reps)
    # x read and y written once, the stored qdata once, and with the geometry on
    # the fly the mesh vertices once
    bytes_apply = 8.0 * (2 * N + s_comp * E * q3) + (24.0 * (n + 1) ** 3 if qmode == "geometry" else 0.0)
    pa = {"operator_gdofs": round(N / (ms_apply * 1e-3) / 1e9, 3), "operator_ms": ms_apply,
          "operator_kernel_ms": ms_kern, "qdata_per_point": s_comp, "qdata_mode": qmode}
    if qmode != "general":
        # the same operator with the stored general (6-value) quadrature data, for comparison
        os.environ["FEMGPU_HEX_QDATA"] = "general"
        try:
            Kg = make_operator(fg, ctx, w, n, p)
        finally:
            del os.environ["FEMGPU_HEX_QDATA"]
        Kg.time_apply(xv, yv, 5)
        msg_apply, msg_kern = Kg.time_apply(xv, yv, a.apply_reps)
        sg = Kg.qdata_per_point()
        bg = 8.0 * (2 * N + sg * E * q3)
        pa["general_qdata"] = {"qdata_per_point": sg, "operator_ms": msg_apply, "operator_kernel_ms": msg_kern,
                               "operator_gdofs": round(N / (msg_apply * 1e-3) / 1e9, 3),
                               "algorithmic_bytes": bg,
                               "frac": round(bg / (msg_apply * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        Kg.close()
    # the FP32 copy of the operator (the mixed-precision legs' inner apply)
    K32 = K.convert(fg.F32)
    xs = fg.DeviceArray(ctx, N, np.float32).upload(np.random.default_rng(0).standard_normal(N).astype(np.float32))
    ys = fg.DeviceArray(ctx, N, np.float32)
    K32.time_apply(xs, ys, 5)
    ms32_apply, ms32_kern = K32.time_apply(xs, ys, a.apply_reps)
    pa["fp32"] = {"operator_ms": ms32_apply, "operator_kernel_ms": ms32_kern,
                  "operator_gdofs": round(N / (ms32_apply * 1e-3) / 1e9, 3)}
    del xs, ys
    K32.close()
    if w["mass_apply"]:
        M = fg.HexPAOperator(p, q, n, form=fg.MASS, rho=1.0, dirichlet=False, ctx=ctx)
        msm_apply, msm_kern = M.time_apply(xv, yv, a.apply_reps)
        pa.update({"mass_gdofs": round(N / (msm_apply * 1e-3) / 1e9, 3), "mass_ms": msm_apply,
                   "mass_kernel_ms": msm_kern})
        M.close()
    del xv, yv

    # Chebyshev interval once (part of set-up, as the operator's diagonal)
    _, lmax = fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=w["rtol"], max_iter=1))
    print(f"bench.py: set-up done (N = {N}, lmax = {lmax:.4f}, apply {ms_apply:.3f} ms)", file=sys.stderr,
          flush=True)

    x0 = dev()  # zero initial guess, copied on the device inside the timed step
    chunk = w.get("chunk", 0) if a.chunk < 0 else a.chunk
    step_cfg = fg.SolverConfig(rtol=w["rtol"], max_iter=chunk if chunk > 0 else 100000)

    def step():
        x.copy_from(x0)
        return fg.solve_pcg_chebyshev(K, b, x, step_cfg, degree=4, lmax=lmax)[0]

    for _ in range(a.warmup):
        step()
    ctx.synchronize()
    times, iters = [], []
    t_all = time.perf_counter()
    for _ in range(a.steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        res = step()
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
        iters.append(res.iterations)
        print(f"bench.py: step {len(times)}: {res.iterations} iterations, {1e3 * times[-1]:.1f} ms",
              file=sys.stderr, flush=True)
        if chunk == 0 and not res.converged:
            print(f"bench.py: solve did not converge ({res})", file=sys.stderr)
            sys.exit(1)
    total = time.perf_counter() - t_all
    ms_step = 1e3 * total / a.steps
    value = N * sum(iters) / total / 1e9
    leg = None if a.no_legs else legs(fg, ctx, K, b, x, lmax, w["rtol"], a)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(w, p, N, E, a.cpu_iters, lmax)
        print(f"bench.py: CPU baseline {cpu['value']} GDOF/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)
    kform = 2 if w["form"] == "diffusion" else 3
    flops_apply = E * apply_flops_per_element(p + 1, q, kform == 3, qmode)
    kernel = ((f"PA apply = k_pa_hex_tpe<{q},{q},double,{'true' if K.cartesian else 'false'}> element kernel "
               f"(one thread per element"
               if K.element_kernel() == "thread-per-element" else
               "PA apply = k_pa_hex_quad<double> element kernel (one quad of lanes per element"
               if K.element_kernel() == "quad-per-element" else
               f"PA apply = k_pa_hex<{q},{q},double,{kform},{ {'general': 0, 'cartesian': 1, 'geometry': 2}[qmode]}> "
               f"element kernel (one workgroup per {256 // (q * q)} elements") +
              f"; {s_comp} qdata value(s) per Gauss point"
              f"{', axis-aligned mesh' if K.cartesian else ''}"
              f"{', Jacobian formed from the element vertices at every point' if qmode == 'geometry' else ''}) + "
              f"k_hex_gather<double,{p},unsigned int> DOF gather; bytes over the apply's time, flops over the "
              f"element kernel's own HIP-event time")
    # the apply's bytes over the apply (element kernel + gather); its flops,
    # all in the element kernel, over that kernel's own event time
    hbm_gbs = bytes_apply / (ms_apply * 1e-3) / 1e9
    tflops = flops_apply / ((ms_kern if ms_kern > 0 else ms_apply) * 1e-3) / 1e12
    if qmode == "geometry":
        # the geometry on the fly trades qdata bytes for FP64 vector work: the
        # apply is bound by issue, not by HBM
        roofline = {"bound": "valu", "achieved": round(tflops, 2), "peak": FP64_VALU_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(tflops / FP64_VALU_PEAK_TFLOPS, 4),
                    "traffic": committed_traffic(a.workload, n, p, s_comp), "kernel": kernel,
                    "algorithmic_bytes": bytes_apply, "flops_per_apply": flops_apply,
                    "hbm_achieved_gb_s": round(hbm_gbs, 1), "hbm_frac": round(hbm_gbs / HBM_PEAK_GBS, 4)}
    else:
        roofline = {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(hbm_gbs / HBM_PEAK_GBS, 4),
                    "traffic": committed_traffic(a.workload, n, p, s_comp), "kernel": kernel,
                    "algorithmic_bytes": bytes_apply, "flops_per_apply": flops_apply,
                    "flop_frac_fp64_valu": round(tflops / FP64_VALU_PEAK_TFLOPS, 4)}
    roofline["element_kernel_ms"] = round(ms_kern, 4)
    roofline["apply_ms"] = round(ms_apply, 4)
    roofline["device_copy_gb_s"] = round(copy_gbs, 1)
    roofline["frac_of_device_copy"] = round(hbm_gbs / copy_gbs, 4)
    out = {
        "metric": "solve time to 1e-8 rel.res + PA apply GDOF/s; plain FP64 vs mixed+learned",
        "value": round(value, 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["desc"].format(n=n, p=p, N=N, q=q), "name": a.workload,
                   "n_dofs": N, "elements": E, "order": p, "quad_points_1d": q,
                   "global_batch": 1, "parallelism": "single GPU"},
        "step": {"what": (f"{chunk} iterations of FP64 PCG + Chebyshev-Jacobi(4) from x0 = 0" if chunk > 0 else
                          f"one whole FP64 PCG + Chebyshev-Jacobi(4) solve to {w['rtol']:g} from x0 = 0"),
                 "iterations": int(iters[-1]), "ms_per_iteration": round(ms_step / max(iters[-1], 1), 4),
                 "lmax": lmax},
        "pa_apply": pa,
        "roofline": roofline,
        "legs": leg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
