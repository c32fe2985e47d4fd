#!/usr/bin/env python3
"""Benchmark of the fem-rs solve path on one MI355X.

Default workload (one step = one full solve, inputs resident in HBM):
  3-D diffusion -div(kappa grad u) = 1 on the unit cube, 64^3 hex elements,
  H1 Q2 (129^3 = 2,146,689 DOFs), kappa = 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z),
  homogeneous Dirichlet BCs, FP64 PA diffusion apply (3^3 Gauss points),
  PCG + Chebyshev-Jacobi(4) preconditioner to 1e-8 relative residual, x0 = 0.

value = solve throughput N_dofs * PCG iterations / solve time (GDOF/s).  The
line also carries the PA apply rates (the operator's form and, for the
default workload, the mass form; FP64), the solve time, the FP64 residual of
the returned solution (recomputed with a fresh apply), the roofline fraction
of the PA apply (algorithmic bytes 8 (2N + s E q^3) per apply, s = the
quadrature-data values per Gauss point: 6 diffusion / 7 diffusion + mass on a
general mesh, 1 / 2 on an axis-aligned box mesh; over the HIP-event-timed mean
duration of its two kernels; on box meshes also the same operator with
general quadrature data), the HBM traffic per apply from the committed counter summary
for that workload (profiles/*_hbm_traffic.md), the bandwidth of a 1 GiB
device-to-device copy timed in the same process (the practical streaming
ceiling, with the apply's fraction of it) and a CPU baseline (the oracle's
sum-factorised PCG-Chebyshev on the host cores, bounded sample).

"legs" compares, on the same operator and right-hand side, the plain FP64
solve with (a) FP32 PCG-Chebyshev + FP64 refinement (continued: one FP32 PCG
whose residual is replaced by the FP64 one; and restarted inner solves) and
(b) restarted refinement around the FP32 flexible CG with the learned U-Net
preconditioner, trained here first (energy-norm loss, pairs harvested from
PCG-Chebyshev iterates of two other right-hand sides; training time is
reported and is not part of the solve time).  Each leg reports its solve
time, iterations, FP64 relative residual and relative error against an FP64
PCG-Chebyshev solution to rtol / 100.

--workload selects the other configurations (one solve takes seconds to
tens of seconds at the large ones, so use --steps 1 --warmup 0 there):
  p1-256-csr        2-D Poisson on the unit square, 256^2 quads split into two
                    triangles each, P1 (66,049 DOFs), CSR and load vector
                    assembled on the GPU (assemble_bilinear / assemble_linear,
                    row zeroing by apply_dirichlet), PCG + Jacobi to 1e-10
  q3-96-grf         Q3, 96^3, kappa = exp(GRF) clipped to [1e-2, 1e2], random
                    nodal RHS, rtol 1e-10 (24,137,569 DOFs)
  q4-128-perturbed  Q4, 128^3, vertices perturbed by up to 0.2 h, diffusion +
                    mass (rho = 1), kappa = exp(GRF), random RHS, rtol 1e-8
                    (135,005,697 DOFs)
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

WORKLOADS = {
    "q3-96-learned": dict(p=3, n=96, form="learned", rtol=1e-8,
                          desc="learned preconditioner, 3-D diffusion {n}^3 hex Q{p} ({N} DOFs), kappa = exp(GRF) "
                               "per field, random nodal RHS, Dirichlet, FP64 flexible CG (Chebyshev-Jacobi(4) + "
                               "U-Net 16/32/64 BF16) vs PCG + Chebyshev-Jacobi(4) to 1e-8 on held-out fields"),
    "p1-256-csr": dict(p=1, n=256, form="csr", rtol=1e-10,
                       desc="2-D Poisson -lap u = 2 pi^2 sin(pi x) sin(pi y) on the unit square, {n}^2 quads "
                            "split into 2 triangles, P1 ({N} DOFs, nnz {nnz}), assembled CSR, homogeneous "
                            "Dirichlet by row zeroing, FP64 PCG + Jacobi to 1e-10"),
    "q2-64-sinprod": dict(p=2, n=64, form="diffusion", kappa="sinprod", amp=0.0, rhs="unit",
                          rtol=1e-8, mass_apply=True,
                          desc="3-D diffusion, {n}^3 hex Q{p} ({N} DOFs), kappa = 1 + 0.5 sin sin sin, "
                               "Dirichlet, FP64 PA ({q}^3 Gauss points), PCG + Chebyshev-Jacobi(4) to 1e-8"),
    "q3-96-grf": dict(p=3, n=96, form="diffusion", kappa=("grf", 7), amp=0.0, rhs=("random", 11),
                      rtol=1e-10, mass_apply=False,
                      desc="3-D diffusion, {n}^3 hex Q{p} ({N} DOFs), kappa = exp(GRF, l = 0.1, sigma = 1.5) "
                           "in [1e-2, 1e2] (seed 7), random nodal RHS (seed 11), Dirichlet, FP64 PA "
                           "({q}^3 Gauss points), PCG + Chebyshev-Jacobi(4) to 1e-10"),
    "q4-128-perturbed": dict(p=4, n=128, form="diff_mass", kappa=("grf", 7), amp=0.2, rhs=("random", 11),
                             rtol=1e-8, mass_apply=False,
                             desc="3-D diffusion + mass (rho = 1), {n}^3 hex Q{p} ({N} DOFs), vertices "
                                  "perturbed <= 0.2 h (seed 5), kappa = exp(GRF) in [1e-2, 1e2] (seed 7), "
                                  "random nodal RHS (seed 11), Dirichlet, FP64 PA ({q}^3 Gauss points), "
                                  "PCG + Chebyshev-Jacobi(4) to 1e-8"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="q2-64-sinprod", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override elements per direction")
    ap.add_argument("--p", type=int, default=0, help="override the order")
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (counter calibration)")
    ap.add_argument("--no-legs", action="store_true", help="skip the mixed / learned comparison")
    ap.add_argument("--train-steps", type=int, default=300)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--train-fields", type=int, default=4, help="q3-96-learned: training fields")
    ap.add_argument("--test-fields", type=int, default=3, help="q3-96-learned: held-out fields")
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
    Ab.apply_dirichlet_row_zeroing(bnd2, np.zeros(len(bnd2)))  # as config 1: A 1 != 0, rho > 0
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
    """Learned preconditioner on held-out fields of the Q3 GRF family (a
    reduced form of the 256-field training set: --train-fields fields)."""
    import femgpu as fg
    from femgpu import learned
    n = a.n or w["n"]
    p = a.p or w["p"]
    ctx = fg.default_context()
    steps = a.train_steps

    def field(seed):
        K = fg.HexPAOperator(p, p + 1, n, form=fg.DIFFUSION, kappa=fg.grf(seed), dirichlet=True, ctx=ctx)
        b = fg.DeviceArray(ctx, K.n_dofs(), np.float64)
        K.rhs_random(1000 + seed, b)
        x = fg.DeviceArray(ctx, K.n_dofs(), np.float64)
        _, lmax = fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=1e-8, max_iter=1))
        return K, b, lmax

    t0 = time.perf_counter()
    train_ops, pairs = [], []
    for sd in range(a.train_fields):
        K, b, lmax = field(sd)
        ps = learned.harvest_pairs(K, b.download(), iters=(2, 6, 15, 40, 100), lmax=lmax, tol=1e-10)
        pairs += [(len(train_ops), r_, e_) for r_, e_ in ps]
        train_ops.append(K)
        print(f"bench.py: field {sd}: {len(ps)} pairs", file=sys.stderr, flush=True)
    t_harvest = time.perf_counter() - t0
    net = fg.UNet.for_operator(train_ops[0], seed=11)
    rng = np.random.default_rng(0)
    losses = []
    t0 = time.perf_counter()
    cur = -1
    for st in range(steps):
        k, r_, e_ = pairs[rng.integers(len(pairs))]
        if k != cur:
            net.aux_from(train_ops[k])
            cur = k
        losses.append(net.train_step_energy(train_ops[k], r_, e_, 2e-3 * 0.1 ** (st / max(steps, 1))))
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    print(f"bench.py: trained {steps} steps in {t_train:.1f} s", file=sys.stderr, flush=True)
    for K in train_ops:
        K.close()
    del train_ops
    cfg = fg.SolverConfig(rtol=w["rtol"], max_iter=100000)
    rows, it_l, t_l_sum = [], 0, 0.0
    N = 0
    for j in range(a.test_fields):
        sd = 256 + j
        K, b, lmax = field(sd)
        N = K.n_dofs()
        x = fg.DeviceArray(ctx, N, np.float64)
        fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=1e-8, max_iter=3), lmax=lmax)  # warm
        ctx.synchronize()
        x.upload(np.zeros(N))
        ctx.synchronize(); t0 = time.perf_counter()
        rp, _ = fg.solve_pcg_chebyshev(K, b, x, cfg, lmax=lmax)
        ctx.synchronize(); tp = time.perf_counter() - t0
        net.aux_from(K)
        x.upload(np.zeros(N))
        ctx.synchronize(); t0 = time.perf_counter()
        rl, info = fg.solve_fcg_learned(K, net, b, x, cfg, lmax=lmax)
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
        print(f"bench.py: held-out field {sd}: PCG-Chebyshev {rp.iterations} it {1e3 * tp:.0f} ms, "
              f"learned FCG {rl.iterations} it {1e3 * tl:.0f} ms", file=sys.stderr, flush=True)
        it_l += rl.iterations
        t_l_sum += tl
        K.close()
        del r, x, b
    out = {
        "metric": "solve time to 1e-8 rel.res + PA apply GDOF/s; plain FP64 vs mixed+learned",
        "value": round(N * it_l / t_l_sum / 1e9, 4), "unit": "GDOF/s", "n_gpus": 1,
        "steps": a.test_fields, "warmup": 0, "ms_per_step": round(1e3 * t_l_sum / a.test_fields, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["desc"].format(n=n, p=p, N=N), "name": a.workload, "n_dofs": N,
                   "train_fields": a.train_fields, "test_fields": a.test_fields, "global_batch": 1,
                   "parallelism": "single GPU"},
        "train": {"pairs": len(pairs), "steps": steps, "harvest_s": round(t_harvest, 1),
                  "train_s": round(t_train, 1),
                  "energy_loss_first_last": [round(float(np.mean(losses[:20])), 4),
                                             round(float(np.mean(losses[-20:])), 4)]},
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


def cpu_baseline(n, p, iters, lmax):
    """Oracle PCG-Chebyshev(4) (sum-factorised PA, OpenMP over element colours)
    on the host cores, capped at `iters` iterations (default workload only)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import femref as fr
    threads = min(os.cpu_count() or 1, 16)
    d = fr.hex_desc(p, p + 1, n, kind=fr.FR_DIFFUSION, kappa=fr.coef_sinprod(), dirichlet=True)
    c = fr.hex_vertices(n, n, n, 0.0, 0)
    Dq, Mq = fr.hex_qdata(d, c)
    N = fr.hex_ndofs(d)
    op = fr._Op(N, hexop=(d, Dq, Mq, threads))
    diag = fr.hex_diagonal(d, Dq, Mq)
    b = fr.hex_rhs_unit(d, c)
    x0 = np.zeros(N)
    fr.pcg_chebyshev(op, diag, 4, lmax / 30, lmax, b, x0, rtol=1e-8, max_iter=1)  # warm-up
    t0 = time.perf_counter()
    _, res = fr.pcg_chebyshev(op, diag, 4, lmax / 30, lmax, b, x0, rtol=1e-8, max_iter=iters)
    dt = time.perf_counter() - t0
    return {"value": N * res.iterations / dt / 1e9, "unit": "GDOF/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle PCG + Chebyshev-Jacobi(4), sum-factorised FP64 PA, same Q{p} {n}^3 "
                      f"problem, {res.iterations} iterations ({dt:.2f} s)"}


def legs(fg, ctx, K, b, x, x_plain, lmax, t_plain, it_plain, rtol, a, small_op=None):
    from femgpu import learned
    N = K.n_dofs()
    x0 = fg.DeviceArray(ctx, N, np.float64)
    # FP64 reference solution (not timed)
    ref_rtol = rtol / 100
    x.copy_from(x0)
    fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=ref_rtol, max_iter=100000), lmax=lmax)
    xref = x.download()
    err = lambda v: float(np.linalg.norm(v - xref) / np.linalg.norm(xref))
    bh = b.download()
    relres = lambda v: float(np.linalg.norm(bh - K.matvec(v)) / np.linalg.norm(bh))
    K32 = K.convert(fg.F32)
    cfg = fg.SolverConfig(rtol=rtol, max_iter=100000)
    big = N > 10_000_000

    def timed(fn):
        reps = 1 if big else 3
        if not big:
            fn()
        ts = []
        for _ in range(reps):
            x.copy_from(x0)
            ctx.synchronize()
            t0 = time.perf_counter()
            out = fn()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        return out, float(np.median(ts))

    (res_m, inner_m), t_m = timed(lambda: fg.solve_ir_mixed(K, K32, b, x, cfg, lmax=lmax, mode=1,
                                                            inner_rtol=1e-2, update_interval=16))
    xm = x.download()
    print(f"bench.py: mixed leg {1e3 * t_m:.1f} ms", file=sys.stderr, flush=True)
    mixed = {"method": "one FP32 PCG + Chebyshev-Jacobi(4) (FP64 dots) whose residual is replaced by the "
                       "FP64 residual b - A u after each 1e-2 reduction or 16 iterations (u updated in FP64)",
             "solve_time_ms": round(1e3 * t_m, 3), "residual_replacements": res_m.iterations,
             "fp32_iterations": int(inner_m), "rel_residual_fp64": relres(xm), "rel_error": err(xm),
             "speedup_vs_plain_fp64": round(t_plain / t_m, 3)}
    (res_r, inner_r), t_r = timed(lambda: fg.solve_ir_mixed(K, K32, b, x, cfg, lmax=lmax, mode=0, inner_rtol=1e-4))
    xr = x.download()
    mixed_r = {"method": "restarted refinement: FP32 PCG-Chebyshev inner solves to 1e-4, FP64 outer",
               "solve_time_ms": round(1e3 * t_r, 3), "outer_iterations": res_r.iterations,
               "fp32_iterations": int(inner_r), "rel_error": err(xr), "speedup_vs_plain_fp64": round(t_plain / t_r, 3)}
    out = {"plain_fp64": {"method": "FP64 PCG + Chebyshev-Jacobi(4)", "solve_time_ms": round(1e3 * t_plain, 3),
                          "iterations": int(it_plain), "rel_error": err(x_plain)},
           "mixed_fp32_ir": mixed, "mixed_fp32_ir_restarted": mixed_r, "reference_rtol": ref_rtol}
    if a.no_learned:
        return out
    # learned preconditioner: pairs from two other right-hand sides, energy-norm training.
    # Above 10M DOFs the network is trained on a 4x coarser mesh of the same family
    # (the convolutions are resolution-independent) and its weights are moved over.
    t0 = time.perf_counter()
    Kt = small_op() if big and small_op else K
    lmax_t = lmax if Kt is K else fg.solve_pcg_chebyshev(Kt, Kt.rhs_random(101), np.zeros(Kt.n_dofs()),
                                                         fg.SolverConfig(rtol=1e-8, max_iter=1))[1]
    pairs = []
    for sd in (101, 102):
        pairs += learned.harvest_pairs(Kt, Kt.rhs_random(sd), iters=(1, 3, 6, 10, 20, 35, 60), lmax=lmax_t)
    net_t = fg.UNet.for_operator(Kt, seed=11)
    rng = np.random.default_rng(0)
    losses = []
    for st in range(a.train_steps):
        r_, e_ = pairs[rng.integers(len(pairs))]
        losses.append(net_t.train_step_energy(Kt, r_, e_, 2e-3 * 0.1 ** (st / max(a.train_steps, 1))))
    if Kt is K:
        net = net_t
    else:
        net = fg.UNet.for_operator(K, seed=11)
        net.set_params(net_t.get_params())
        net_t.close()
        Kt.close()
    t_train = time.perf_counter() - t0
    print(f"bench.py: trained {a.train_steps} steps in {t_train:.1f} s", file=sys.stderr, flush=True)
    (res_l, inner_l, info_l), t_l = timed(
        lambda: fg.solve_ir_learned(K, K32, net, b, x, cfg, inner_rtol=1e-4, lmax=lmax))
    xl = x.download()
    ml = {"method": "FP32 flexible CG (Chebyshev-Jacobi(4) + U-Net correction, 16 directions) inner, "
                    "FP64 refinement",
          "solve_time_ms": round(1e3 * t_l, 3), "outer_iterations": res_l.iterations,
          "inner_iterations": int(inner_l), "fell_back": info_l.fell_back,
          "rel_residual_fp64": relres(xl), "rel_error": err(xl),
          "speedup_vs_plain_fp64": round(t_plain / t_l, 3),
          "train": {"pairs": len(pairs), "steps": a.train_steps, "time_s": round(t_train, 2),
                    "trained_on": "the operator solved" if Kt is K else "a 4x coarser mesh of the same family",
                    "energy_loss_first_last": [round(float(np.mean(losses[:20])), 4),
                                               round(float(np.mean(losses[-20:])), 4)]}}
    out["mixed_learned"] = ml
    if not big:
        (res_f, info_f), t_f = timed(lambda: fg.solve_fcg_learned(K, net, b, x, cfg, lmax=lmax))
        xf = x.download()
        out["learned_fcg_fp64"] = {"method": "FP64 flexible CG with the U-Net (no refinement)",
                                   "solve_time_ms": round(1e3 * t_f, 3), "iterations": res_f.iterations,
                                   "rel_residual_fp64": relres(xf), "rel_error": err(xf),
                                   "speedup_vs_plain_fp64": round(t_plain / t_f, 3)}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {a.gpus}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    import femgpu as fg

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
    s_comp = K.qdata_per_point()  # 6 / 7 general, 1 / 2 on an axis-aligned box mesh
    dev = lambda: fg.DeviceArray(ctx, N, np.float64)
    b = dev()
    if w["rhs"] == "unit":
        K.rhs_unit(b)
    else:
        K.rhs_random(w["rhs"][1], b)
    x = dev()
    cfg = fg.SolverConfig(rtol=w["rtol"], max_iter=100000)

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
    bytes_apply = 8.0 * (2 * N + s_comp * E * q3)
    pa = {"operator_gdofs": round(N / (ms_apply * 1e-3) / 1e9, 3), "operator_ms": ms_apply,
          "operator_kernel_ms": ms_kern, "qdata_per_point": s_comp}
    if K.cartesian:
        # the same operator with the general (6-value) quadrature data, for comparison
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
                               "frac": round(bg / (msg_kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        Kg.close()
    if w["mass_apply"]:
        M = fg.HexPAOperator(p, q, n, form=fg.MASS, rho=1.0, dirichlet=False, ctx=ctx)
        msm_apply, msm_kern = M.time_apply(xv, yv, a.apply_reps)
        pa.update({"mass_gdofs": round(N / (msm_apply * 1e-3) / 1e9, 3), "mass_ms": msm_apply,
                   "mass_kernel_ms": msm_kern})
        M.close()
    del xv, yv

    # Chebyshev interval once (part of set-up, as the operator's diagonal)
    _, lmax = fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=w["rtol"], max_iter=1))

    x0 = dev()  # zero initial guess, copied on the device inside the timed step

    def step():
        x.copy_from(x0)
        return fg.solve_pcg_chebyshev(K, b, x, cfg, degree=4, lmax=lmax)[0]

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
        if not res.converged:
            print(f"bench.py: solve did not converge ({res})", file=sys.stderr)
            sys.exit(1)
    total = time.perf_counter() - t_all
    # independent FP64 residual of the last solution
    r = dev()
    K.apply_zero(x, r)
    xs, rs, bs = x.download(), r.download(), b.download()
    rel_res = float(np.linalg.norm(bs - rs) / np.linalg.norm(bs))
    del r, rs

    ms_step = 1e3 * total / a.steps
    value = N * sum(iters) / total / 1e9
    t_plain = float(np.median(times))
    leg = None if a.no_legs else legs(fg, ctx, K, b, x, xs, lmax, t_plain, iters[-1], w["rtol"], a,
                                      small_op=lambda: make_operator(fg, ctx, w, max(n // 4, 8), p))
    cpu = None
    if not a.no_cpu and a.workload == "q2-64-sinprod":
        cpu = cpu_baseline(n, p, a.cpu_iters, lmax)
    kform = 2 if w["form"] == "diffusion" else 3
    flops_apply = 2.0 * E * sumfac_fma_per_element(p + 1, q, True, kform == 3, K.cartesian)
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
        "solve": {"iterations": int(iters[-1]), "solve_time_ms": round(1e3 * t_plain, 4),
                  "rel_residual_fp64": rel_res, "lmax": lmax},
        "pa_apply": pa,
        "roofline": {
            "bound": "hbm",
            "achieved": round(bytes_apply / (ms_kern * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(bytes_apply / (ms_kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": committed_traffic(a.workload, n, p, s_comp),
            "kernel": (f"PA apply = k_pa_hex_tpe<{q},{q},double,{'true' if K.cartesian else 'false'}> element kernel "
                       f"(one thread per element"
                       if K.element_kernel() == "thread-per-element" else
                       "PA apply = k_pa_hex_quad<double> element kernel (one quad of lanes per element"
                       if K.element_kernel() == "quad-per-element" else
                       f"PA apply = k_pa_hex<{q},{q},double,{kform},{'true' if K.cartesian else 'false'}> element "
                       f"kernel (one workgroup per {256 // (q * q)} elements") +
                      f"; {s_comp} qdata value(s) per Gauss point"
                      f"{', axis-aligned mesh' if K.cartesian else ''}) + "
                      f"k_hex_gather<double,{p},unsigned int> DOF gather (HIP events around both)",
            "algorithmic_bytes": bytes_apply,
            "flops_per_apply": flops_apply,
            "flop_frac_fp64_valu": round(flops_apply / (ms_kern * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS, 4),
            "device_copy_gb_s": round(copy_gbs, 1),
            "frac_of_device_copy": round(bytes_apply / (ms_kern * 1e-3) / 1e9 / copy_gbs, 4),
        },
        "legs": leg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
