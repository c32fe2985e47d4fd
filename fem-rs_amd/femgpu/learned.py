"""Training-pair harvesting and training for the learned preconditioner.

Pairs come from plain PCG + Chebyshev-Jacobi solves (no reference
counterpart): with x* the converged solution and x_k the k-th PCG iterate,
the error e_k = x* - x_k and the residual r_k = b - A x_k = A e_k form an
exact (input, target) pair.

femgpu_solve_fcg_learned smooths before it calls the network: z1 = C r,
r1 = r - A z1, and the network corrects r1.  harvest_pairs(smoothed=True)
therefore returns the pairs the network sees there, (r1, e1) with
e1 = e_k - C r_k = A^-1 r1: residuals whose rough part the Chebyshev
polynomial has already removed.  Every operation runs through the libfemgpu
C ABI.
"""
import numpy as np

from . import SolverConfig, solve_pcg_chebyshev, chebyshev_jacobi_apply, F64, F32, DeviceArray


def harvest_pairs(op, b, iters=(1, 2, 3, 4, 6, 8, 11, 15, 20, 27, 35), degree=4, ratio=30.0,
                  lmax=0.0, x_star=None, tol=1e-11, smoothed=True):
    """-> list of (r, e) float32 pairs (A e = r) for one operator / right-hand side."""
    n = op.n_dofs()
    dt = np.float64 if op.dtype == F64 else np.float32
    b = np.ascontiguousarray(b, dt)
    if x_star is None:
        x_star = np.zeros(n, dt)
        solve_pcg_chebyshev(op, b, x_star, SolverConfig(rtol=tol, max_iter=20000), degree=degree,
                            ratio=ratio, lmax=lmax)
    pairs = []
    for k in iters:
        xk = np.zeros(n, dt)
        solve_pcg_chebyshev(op, b, xk, SolverConfig(rtol=1e-30, max_iter=k), degree=degree,
                            ratio=ratio, lmax=lmax)
        ek = x_star - xk
        rk = b - op.matvec(xk)
        if smoothed:
            z1 = np.zeros(n, dt)
            chebyshev_jacobi_apply(op, rk, z1, degree=degree, ratio=ratio, lmax=lmax)
            rk = rk - op.matvec(z1)
            ek = ek - z1
        if np.linalg.norm(rk) > 0 and np.linalg.norm(ek) > 0:
            pairs.append((rk.astype(np.float32), ek.astype(np.float32)))
    return pairs


def train(net, pairs, epochs=10, lr=2e-3, seed=0, lr_decay=0.97, op=None):
    """Adam over the pairs in shuffled order; returns the per-epoch mean loss.

    With op given the loss is the energy-norm error of the Galerkin-scaled
    correction (UNet.train_step_energy), otherwise the L2 error of the
    unit-RMS output (UNet.train_step)."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        tot = 0.0
        for i in order:
            r, e = pairs[i]
            tot += net.train_step_energy(op, r, e, lr) if op is not None else net.train_step(r, e, lr)
        hist.append(tot / max(len(pairs), 1))
        lr *= lr_decay
    return hist


def harvest_pairs_known(op, seed, iters=(1, 3, 6, 10, 20, 35, 60), degree=4, ratio=30.0, lmax=0.0,
                        smoothed=True):
    """Pairs from PCG + Chebyshev-Jacobi iterates of a solve whose solution is
    known by construction, all on the device: x* = a seeded uniform [-1, 1]
    nodal field (zero on essential DOFs; the hex operator's rhs_random),
    b = A x*, x_k the k-th iterate from 0, e_k = x* - x_k, r_k = b - A x_k (then
    smoothed as femgpu_solve_fcg_learned sees them, as harvest_pairs does).
    No converged solve is needed, so this scales to the largest operators.
    -> list of (r, e) float32 DeviceArrays."""
    ctx = op.ctx
    n = op.n_dofs()
    dt = np.float64 if op.dtype == F64 else np.float32
    xs = DeviceArray(ctx, n, dt)
    op.rhs_random(seed, xs)
    b = DeviceArray(ctx, n, dt)
    op.apply_zero(xs, b)
    zero = DeviceArray(ctx, n, dt)
    xk = DeviceArray(ctx, n, dt)
    rk = DeviceArray(ctx, n, dt)
    z1 = DeviceArray(ctx, n, dt)
    t = DeviceArray(ctx, n, dt)
    pairs = []
    for k in iters:
        xk.copy_from(zero)
        solve_pcg_chebyshev(op, b, xk, SolverConfig(rtol=1e-30, max_iter=k), degree=degree, ratio=ratio,
                            lmax=lmax)
        ek = DeviceArray(ctx, n, dt)
        ek.copy_from(xs)
        ek.axpby(-1.0, xk, 1.0)                 # e_k = x* - x_k
        op.apply_zero(xk, rk)
        rk.axpby(1.0, b, -1.0)                  # r_k = b - A x_k
        if smoothed:
            chebyshev_jacobi_apply(op, rk, z1, degree=degree, ratio=ratio, lmax=lmax)
            op.apply_zero(z1, t)
            rk.axpby(-1.0, t, 1.0)              # r1 = r_k - A z1
            ek.axpby(-1.0, z1, 1.0)             # e1 = e_k - z1
        pairs.append((rk.astype(np.float32), ek.astype(np.float32)))
        ek.free()
    for v in (xs, b, zero, xk, rk, z1, t):
        v.free()
    return pairs
