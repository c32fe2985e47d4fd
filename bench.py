#!/usr/bin/env python3
"""Benchmark of the fem-rs solve path on one MI355X.

Workload (one step = one full solve, inputs resident in HBM):
  3-D diffusion -div(kappa grad u) = 1 on the unit cube, 64^3 hex elements,
  H1 Q2 (129^3 = 2,146,689 DOFs), kappa = 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z),
  homogeneous Dirichlet BCs, FP64 PA diffusion apply (3^3 Gauss points),
  PCG + Chebyshev-Jacobi(4) preconditioner to 1e-8 relative residual, x0 = 0.

value = solve throughput N_dofs * PCG iterations / solve time (GDOF/s).  The
line also carries the PA apply rates (diffusion and mass, FP64), the solve
time, the FP64 residual of the returned solution (recomputed with a fresh
apply), the roofline fraction of the PA diffusion kernel (algorithmic bytes
8 * (2N + 6 E 27) per launch over its HIP-event-timed mean duration) and a CPU
baseline (the oracle's sum-factorised PCG-Chebyshev on the host cores, bounded
sample).

"legs" compares, on the same operator and right-hand side, the plain FP64
solve with (a) FP32-inner PCG-Chebyshev + FP64 iterative refinement and (b)
the same refinement around the FP32 flexible CG with the learned U-Net
preconditioner, trained here first (energy-norm loss, pairs harvested from
PCG-Chebyshev iterates of two other right-hand sides; training time is
reported and is not part of the solve time).  Each leg reports its solve
time, iterations, FP64 relative residual and relative error against an FP64
PCG-Chebyshev solution to 1e-12.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fem-rs_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="elements per direction")
    ap.add_argument("--p", type=int, default=2)
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (counter calibration)")
    ap.add_argument("--no-legs", action="store_true", help="skip the mixed / learned comparison")
    ap.add_argument("--train-steps", type=int, default=300)
    return ap.parse_args()


def cpu_baseline(n, p, iters, lmax):
    """Oracle PCG-Chebyshev(4) (sum-factorised PA, OpenMP over element colours)
    on the host cores, capped at `iters` iterations."""
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


def legs(fg, ctx, K, b, x, x_plain, lmax, t_plain, it_plain, a):
    from femgpu import learned
    N = K.n_dofs()
    dev = lambda dt=np.float64: fg.DeviceArray(ctx, N, dt)
    x0 = dev()
    # FP64 reference solution (not timed)
    x.copy_from(x0)
    fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=1e-12, max_iter=20000), lmax=lmax)
    xref = x.download()
    err = lambda v: float(np.linalg.norm(v - xref) / np.linalg.norm(xref))
    bh = b.download()
    relres = lambda v: float(np.linalg.norm(bh - K.matvec(v)) / np.linalg.norm(bh))
    K32 = K.convert(fg.F32)
    cfg = fg.SolverConfig(rtol=1e-8, max_iter=20000)

    def timed(fn, reps=3):
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

    (res_m, inner_m), t_m = timed(lambda: fg.solve_ir_mixed(K, K32, b, x, cfg, inner_rtol=1e-4, lmax=lmax))
    xm = x.download()
    mixed = {"method": "FP32 PCG + Chebyshev-Jacobi(4) inner (FP64 dots), FP64 refinement",
             "solve_time_ms": round(1e3 * t_m, 3), "outer_iterations": res_m.iterations,
             "inner_iterations": int(inner_m), "rel_residual_fp64": relres(xm), "rel_error": err(xm),
             "speedup_vs_plain_fp64": round(t_plain / t_m, 3)}
    # learned preconditioner: pairs from two other right-hand sides, energy-norm training
    t0 = time.perf_counter()
    pairs = []
    for sd in (101, 102):
        pairs += learned.harvest_pairs(K, K.rhs_random(sd), iters=(1, 3, 6, 10, 20, 35, 60), lmax=lmax)
    net = fg.UNet.for_operator(K, seed=11)
    rng = np.random.default_rng(0)
    losses = []
    for st in range(a.train_steps):
        r_, e_ = pairs[rng.integers(len(pairs))]
        losses.append(net.train_step_energy(K, r_, e_, 2e-3 * 0.1 ** (st / max(a.train_steps, 1))))
    t_train = time.perf_counter() - t0
    (res_l, inner_l, info_l), t_l = timed(lambda: fg.solve_ir_learned(K, K32, net, b, x, cfg, inner_rtol=1e-4, lmax=lmax))
    xl = x.download()
    (res_f, info_f), t_f = timed(lambda: fg.solve_fcg_learned(K, net, b, x, cfg, lmax=lmax))
    xf = x.download()
    ml = {"method": "FP32 flexible CG (Chebyshev-Jacobi(4) + U-Net correction, 16 directions) inner, "
                    "FP64 refinement",
          "solve_time_ms": round(1e3 * t_l, 3), "outer_iterations": res_l.iterations,
          "inner_iterations": int(inner_l), "fell_back": info_l.fell_back,
          "rel_residual_fp64": relres(xl), "rel_error": err(xl),
          "speedup_vs_plain_fp64": round(t_plain / t_l, 3),
          "train": {"pairs": len(pairs), "steps": a.train_steps, "time_s": round(t_train, 2),
                    "energy_loss_first_last": [round(float(np.mean(losses[:20])), 4),
                                               round(float(np.mean(losses[-20:])), 4)]}}
    fcg64 = {"method": "FP64 flexible CG with the U-Net (no refinement)",
             "solve_time_ms": round(1e3 * t_f, 3), "iterations": res_f.iterations,
             "rel_residual_fp64": relres(xf), "rel_error": err(xf),
             "speedup_vs_plain_fp64": round(t_plain / t_f, 3)}
    return {"plain_fp64": {"method": "FP64 PCG + Chebyshev-Jacobi(4)", "solve_time_ms": round(1e3 * t_plain, 3),
                           "iterations": int(it_plain), "rel_error": err(x_plain)},
            "mixed_fp32_ir": mixed, "mixed_learned": ml, "learned_fcg_fp64": fcg64}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {a.gpus}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    import femgpu as fg

    ctx = fg.default_context()
    n, p = a.n, a.p
    K = fg.HexPAOperator(p, p + 1, n, form=fg.DIFFUSION, kappa="sinprod", dirichlet=True, ctx=ctx)
    M = fg.HexPAOperator(p, p + 1, n, form=fg.MASS, rho=1.0, dirichlet=False, ctx=ctx)
    N = K.n_dofs()
    E = n ** 3
    q3 = (p + 1) ** 3
    dev = lambda: fg.DeviceArray(ctx, N, np.float64)
    b = dev()
    K.rhs_unit(b)
    x = dev()
    zero = np.zeros(N)
    cfg = fg.SolverConfig(rtol=1e-8, max_iter=2000)

    if a.copy_calib_mb:
        nc = a.copy_calib_mb * (1 << 20) // 8
        src = fg.DeviceArray(ctx, nc, np.float64, np.ones(nc))
        dst = fg.DeviceArray(ctx, nc, np.float64)
        dst.copy_from(src)  # one copy kernel of known size: reads and writes nc*8 bytes
        del src, dst

    # PA apply rates (device vectors, HIP events)
    xv = dev().upload(np.random.default_rng(0).standard_normal(N))
    yv = dev()
    K.time_apply(xv, yv, 5)
    ms_apply, ms_kern = K.time_apply(xv, yv, a.apply_reps)
    msm_apply, msm_kern = M.time_apply(xv, yv, a.apply_reps)
    bytes_diff = 8.0 * (2 * N + 6 * E * q3)
    bytes_mass = 8.0 * (2 * N + 1 * E * q3)

    # Chebyshev interval once (part of set-up, as the operator's diagonal)
    x.upload(zero)
    _, lmax = fg.solve_pcg_chebyshev(K, b, x, fg.SolverConfig(rtol=1e-8, max_iter=1))

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
        if not res.converged:
            print(f"bench.py: solve did not converge ({res})", file=sys.stderr)
            sys.exit(1)
    total = time.perf_counter() - t_all
    # independent FP64 residual of the last solution
    r = dev()
    K.apply_zero(x, r)
    xs, rs, bs = x.download(), r.download(), b.download()
    rel_res = float(np.linalg.norm(bs - rs) / np.linalg.norm(bs))

    ms_step = 1e3 * total / a.steps
    value = N * sum(iters) / total / 1e9
    t_plain = float(np.median(times))
    leg = None if a.no_legs else legs(fg, ctx, K, b, x, xs, lmax, t_plain, iters[-1], a)
    cpu = None if a.no_cpu else cpu_baseline(n, p, a.cpu_iters, lmax)
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
        "config": {"workload": f"3-D diffusion, {n}^3 hex Q{p} ({N} DOFs), kappa = 1 + 0.5 sin sin sin, "
                               "Dirichlet, FP64 PA (3^3 Gauss points), PCG + Chebyshev-Jacobi(4) to 1e-8",
                   "n_dofs": N, "elements": E, "order": p, "quad_points_1d": p + 1,
                   "global_batch": 1, "parallelism": "single GPU"},
        "solve": {"iterations": int(iters[-1]), "solve_time_ms": round(1e3 * float(np.median(times)), 4),
                  "rel_residual_fp64": rel_res, "lmax": lmax},
        "pa_apply": {
            "diffusion_gdofs": round(N / (ms_apply * 1e-3) / 1e9, 3),
            "diffusion_ms": ms_apply, "diffusion_kernel_ms": ms_kern,
            "mass_gdofs": round(N / (msm_apply * 1e-3) / 1e9, 3),
            "mass_ms": msm_apply, "mass_kernel_ms": msm_kern,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(bytes_diff / (ms_kern * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(bytes_diff / (ms_kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": f"PA diffusion apply = k_pa_hex<{p + 1},{p + 1},double,2> element kernel + "
                      f"k_hex_gather<double,{p}> (HIP events around both)",
            "algorithmic_bytes": bytes_diff,
        },
        "legs": leg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
