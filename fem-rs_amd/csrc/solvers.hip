// Device-resident Krylov solvers and smoothers.
//
// The host loop only enqueues kernels; every scalar (rho, alpha, beta, norms,
// the convergence test) lives in device memory and is computed by one-block
// finalise kernels from fixed-grid partial sums, so results are bitwise
// reproducible.  Once a finalise kernel decides to stop it sets a device flag
// that makes every later kernel of the iteration a no-op; the host polls that
// flag every few iterations, so no host round trip sits on the critical path.
//
// Semantics follow the reference:
//  - CG            solve_cg_operator          (solver/src/lib.rs:198-278)
//  - PCG-Jacobi    par_solve_pcg_jacobi       (parallel/src/par_solver.rs:262-346),
//                  the in-tree definition behind solve_pcg_jacobi (lib.rs:281-297)
//  - Chebyshev     chebyshev_smooth           (amg/src/lib.rs:53-103)
//  - power method  estimate_spectral_radius   (amg/src/lib.rs:108-123)
// plus the textbook Chebyshev-Jacobi preconditioner and the FP32-inner /
// FP64-outer iterative refinement, which have no reference counterpart.
#include "internal.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <type_traits>

namespace femgpu {

enum Scal {
    S_RZ = 0, S_RZNEW, S_PAP, S_RR, S_ALPHA, S_BETA, S_BNORM, S_TOL, S_RES, S_ITERS, S_STATUS,
    S_BB, S_RR0, S_XPEND, S_COUNT
};
enum Status { ST_RUNNING = 0, ST_CONVERGED = 1, ST_BREAKDOWN = 2 };
enum Mode { MODE_CG = 0, MODE_PCG = 1 };
enum Prec { PREC_NONE = 0, PREC_JACOBI = 1, PREC_CHEB = 2, PREC_CB = 3 };  // CB: the pc callback

static inline unsigned nblk(size_t n) { return (unsigned)red_blocks(n); }

// ------------------------------------------------------------ kernels --
// a*b + c with one rounding (written out so that every kernel forming the
// same update rounds it the same way)
template <class T> __device__ __forceinline__ T fmad(T a, T b, T c);
template <> __device__ __forceinline__ double fmad<double>(double a, double b, double c) { return fma(a, b, c); }
template <> __device__ __forceinline__ float fmad<float>(float a, float b, float c) { return fmaf(a, b, c); }

// r = b - ax;  z = inv*r (or r);  p = z;  partials: r.z, b.b, r.r
template <class T, bool JAC>
__global__ __launch_bounds__(256) void k_init(size_t n, const T *__restrict__ b,
                                              const T *__restrict__ ax, const T *__restrict__ inv,
                                              T *__restrict__ r, T *__restrict__ z,
                                              T *__restrict__ p, double *__restrict__ part,
                                              bool write_zp) {
    double v[3] = {0, 0, 0};
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const T ri = b[i] - ax[i];
        r[i] = ri;
        const T zi = JAC ? inv[i] * ri : ri;
        if (write_zp) { z[i] = zi; p[i] = zi; }
        v[0] += (double)ri * (double)zi;
        v[1] += (double)b[i] * (double)b[i];
        v[2] += (double)ri * (double)ri;
    }
    block_sum<3, 256>(v);
    if (threadIdx.x == 0) {
        part[0 * kRedBlocks + blockIdx.x] = v[0];
        part[1 * kRedBlocks + blockIdx.x] = v[1];
        part[2 * kRedBlocks + blockIdx.x] = v[2];
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_dot(size_t n, const T *__restrict__ a,
                                             const T *__restrict__ b, double *__restrict__ part,
                                             int slot, const int *skip) {
    if (skip && *skip) return;
    double v[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        v[0] += (double)a[i] * (double)b[i];
    block_sum<1, 256>(v);
    if (threadIdx.x == 0) part[slot * kRedBlocks + blockIdx.x] = v[0];
}

// x += alpha p; r -= alpha ap; [z = inv r]; partials r.r, r.z
template <class T, bool JAC>
__global__ __launch_bounds__(256) void k_update(size_t n, const double *__restrict__ S,
                                                T *__restrict__ x, const T *__restrict__ p,
                                                T *__restrict__ r, const T *__restrict__ ap,
                                                const T *__restrict__ inv, T *__restrict__ z,
                                                double *__restrict__ part, const int *skip) {
    if (*skip) return;
    const T alpha = (T)S[S_ALPHA];
    double v[2] = {0, 0};
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        x[i] = fmad<T>(alpha, p[i], x[i]);
        const T ri = fmad<T>(-alpha, ap[i], r[i]);
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
        if (JAC) {
            const T zi = inv[i] * ri;
            z[i] = zi;
            v[1] += (double)ri * (double)zi;
        }
    }
    block_sum<2, 256>(v);
    if (threadIdx.x == 0) {
        part[0 * kRedBlocks + blockIdx.x] = v[0];
        part[1 * kRedBlocks + blockIdx.x] = v[1];
    }
}

template <class T>
__global__ void k_pupdate(size_t n, const double *__restrict__ S, const T *__restrict__ z,
                          T *__restrict__ p, const int *skip) {
    if (*skip) return;
    const T beta = (T)S[S_BETA];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = fmad<T>(beta, p[i], z[i]);
}

__device__ __forceinline__ void fin_sum(const double *part, int nb, double *out, int nv) {
    for (int k = 0; k < nv; ++k) {
        double v[1] = {threadIdx.x < (unsigned)nb ? part[k * kRedBlocks + threadIdx.x] : 0.0};
        block_sum<1, 1024>(v);
        if (threadIdx.x == 0) out[k] = v[0];
        __syncthreads();
    }
}

// After k_init: rz, bnorm, initial residual; decides the early exit.
__global__ __launch_bounds__(1024) void k_fin_init(const double *part, int nb, double *S,
                                                   int *flags, int mode, double rtol, double atol) {
    __shared__ double o[3];
    fin_sum(part, nb, o, 3);
    if (threadIdx.x) return;
    S[S_RZ] = o[0];
    S[S_BB] = o[1];
    const double bnorm = sqrt(o[1]);
    S[S_BNORM] = bnorm;
    S[S_RR] = o[2];
    S[S_ITERS] = 0;
    S[S_STATUS] = ST_RUNNING;
    S[S_XPEND] = 0;
    flags[0] = 0;
    if (mode == MODE_CG) {
        // tol = max(atol, rtol * max(|b|, 1e-32)); return if |r0| <= tol  (lib.rs:229-240)
        const double tol = fmax(atol, rtol * fmax(bnorm, 1e-32));
        S[S_TOL] = tol;
        const double res = sqrt(o[2]);
        S[S_RES] = res;
        if (res <= tol) { S[S_STATUS] = ST_CONVERGED; flags[0] = 1; }
    } else {
        // |b| < 1e-30: converged, 0 iterations, residual 0 (par_solver.rs:291-294)
        S[S_RES] = bnorm > 0 ? sqrt(o[2]) / bnorm : 0.0;
        if (bnorm < 1e-30) { S[S_STATUS] = ST_CONVERGED; S[S_RES] = 0.0; flags[0] = 1; }
    }
}

// pAp -> alpha, breakdown test.
__global__ __launch_bounds__(1024) void k_fin_alpha(const double *part, int nb, double *S,
                                                    int *flags, double brk) {
    if (flags[0]) return;
    __shared__ double o[1];
    fin_sum(part, nb, o, 1);
    if (threadIdx.x) return;
    const double pap = o[0];
    S[S_PAP] = pap;
    if (fabs(pap) < brk) { S[S_STATUS] = ST_BREAKDOWN; flags[0] = 1; return; }
    S[S_ALPHA] = S[S_RZ] / pap;
}

// r.r (and r.z for Jacobi) -> convergence test, beta.
// have_rz: 1 if slot 1 holds r.z; CG uses r.r as r.z.
__global__ __launch_bounds__(1024) void k_fin_conv(const double *part, int nb, double *S,
                                                   int *flags, int mode, int have_rz,
                                                   int compute_beta, double rtol, double atol,
                                                   double iter1) {
    if (flags[0]) return;
    __shared__ double o[2];
    fin_sum(part, nb, o, have_rz ? 2 : 1);
    if (threadIdx.x) return;
    const double rr = o[0];
    S[S_RR] = rr;
    S[S_ITERS] = iter1;
    if (mode == MODE_CG) {
        const double res = sqrt(rr);
        S[S_RES] = res;
        if (res <= S[S_TOL]) { S[S_STATUS] = ST_CONVERGED; flags[0] = 1; return; }
    } else {
        const double res = sqrt(rr) / S[S_BNORM];
        S[S_RES] = res;
        if (res < rtol || sqrt(rr) < atol) { S[S_STATUS] = ST_CONVERGED; flags[0] = 1; return; }
    }
    if (compute_beta) {
        const double rz_new = have_rz ? o[1] : rr;
        S[S_BETA] = rz_new / S[S_RZ];
        S[S_RZ] = rz_new;
    }
}

// r.z from the preconditioner -> beta.
__global__ __launch_bounds__(1024) void k_fin_beta(const double *part, int nb, double *S,
                                                   int *flags) {
    if (flags[0]) return;
    __shared__ double o[1];
    fin_sum(part, nb, o, 1);
    if (threadIdx.x) return;
    S[S_BETA] = o[0] / S[S_RZ];
    S[S_RZ] = o[0];
}

// ---- Chebyshev-Jacobi (textbook; Saad, Iterative Methods, Alg. 12.1, x0 = 0)
// z = inv r / theta; d = z  [if last: partial r.z]
template <class T>
__global__ __launch_bounds__(256) void k_cheb0(size_t n, const T *__restrict__ r,
                                               const T *__restrict__ inv, T *__restrict__ z,
                                               T *__restrict__ d, double inv_theta, int last,
                                               double *__restrict__ part, const int *skip, int write_z) {
    if (skip && *skip) return;
    double v[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const T zi = (T)inv_theta * (inv[i] * r[i]);
        if (write_z) z[i] = zi;
        d[i] = zi;
        if (last) v[0] += (double)r[i] * (double)zi;
    }
    if (last) {
        block_sum<1, 256>(v);
        if (threadIdx.x == 0) part[blockIdx.x] = v[0];
    }
}

// d = c1 d + c2 inv (r - t); z += d  [if last: partial r.z, and d is not
// stored]; with first, z and d equal (1/theta) inv r on entry: neither is read,
// the value is recomputed from r and inv (read here anyway)
template <class T, bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_chebk(size_t n, const T *__restrict__ r,
                                               const T *__restrict__ t, const T *__restrict__ inv,
                                               T *__restrict__ z, T *__restrict__ d, double c1,
                                               double c2, double *__restrict__ part, const int *skip,
                                               double inv_theta) {
    if (skip && *skip) return;
    double v[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const T dk = FIRST ? cheb_first<T>(inv_theta, inv[i], r[i]) : d[i];
        const T di = (T)c1 * dk + (T)c2 * (inv[i] * (r[i] - t[i]));
        if constexpr (!LAST) d[i] = di;  // the last step's d is not read again
        const T zi = (FIRST ? dk : z[i]) + di;
        z[i] = zi;
        if constexpr (LAST) v[0] += (double)r[i] * (double)zi;
    }
    if constexpr (LAST) {
        block_sum<1, 256>(v);
        if (threadIdx.x == 0) part[blockIdx.x] = v[0];
    }
}

// Reference chebyshev_smooth step (amg/src/lib.rs:80-100), r holds A x on entry:
// r = b - r;  k == 0: d = (2/alpha) dinv r  else d = rho ((2/alpha) dinv r + (rho-1) d);  x += d
template <class T>
__global__ void k_cheb_ref(size_t n, const T *__restrict__ b, T *__restrict__ r,
                           const T *__restrict__ dinv, T *__restrict__ d, T *__restrict__ x,
                           double two_over_alpha, double rho, int first) {
#pragma clang fp contract(off)
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const T ri = b[i] - r[i];
        r[i] = ri;
        T di;
        if (first) di = (T)two_over_alpha * dinv[i] * ri;
        else di = (T)rho * ((T)two_over_alpha * dinv[i] * ri + ((T)rho - (T)1) * d[i]);
        d[i] = di;
        x[i] += di;
    }
}

template <class T>
__global__ void k_inv_tiny(size_t n, const T *__restrict__ d, T *__restrict__ inv, double tiny) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        inv[i] = fabs((double)d[i]) > tiny ? (T)1 / d[i] : (T)1;
}

template <class T>
__global__ void k_scale_to(size_t n, const T *__restrict__ w, T *__restrict__ v,
                           const double *__restrict__ S, int slot) {
    // v = w / sqrt(S[slot]) if > 1e-14
    const double rho = sqrt(S[slot]);
    if (!(rho > 1e-14)) return;
    const double inv = 1.0 / rho;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        v[i] = (T)((double)w[i] * inv);
}

template <class T>
__global__ void k_div(size_t n, T *__restrict__ w, const T *__restrict__ d) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        w[i] = w[i] / d[i];
}

template <class T>
__global__ void k_sub_from(size_t n, const T *__restrict__ b, T *__restrict__ r) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        r[i] = b[i] - r[i];
}

__global__ __launch_bounds__(1024) void k_fin_store(const double *part, int nb, double *S, int slot) {
    __shared__ double o[1];
    fin_sum(part, nb, o, 1);
    if (threadIdx.x == 0) S[slot] = o[0];
}

// Residual history (print level Iterations and above): after iteration k
// (k = 0: the initial residual) hist[k] = the residual the stopping test
// used.  Re-recording an index after the solver has stopped writes the same
// value again.
__global__ void k_hist(const double *S, double *hist, uint64_t cap) {
    const uint64_t k = (uint64_t)S[S_ITERS];
    if (k < cap) hist[k] = S[S_RES];
}

// Start of a continued PCG segment: running, no iterations yet, |b| given.
__global__ void k_set_cont(double *S, int *flags, double bnorm) {
    S[S_BNORM] = bnorm;
    S[S_STATUS] = ST_RUNNING;
    S[S_ITERS] = 0;
    S[S_XPEND] = 0;
    flags[0] = 0;
}

// ------------------------------------------- fused PCG-Jacobi on CSR --
// Two kernels per iteration instead of six (SpMV, p.Ap, alpha, update,
// convergence/beta, p update).  Each 256-thread block finalises the scalars
// it needs itself from the previous kernel's block partials (rsum256), in
// exactly the order of the one-block finalisers, so every block sees the same
// alpha / beta / stopping decision and the iterates are bitwise those of the
// separate-kernel loop.  The new search direction p_k = z + beta p_{k-1} is
// formed inside the SpMV (row i reads z_j and p_{k-1,j} of its columns and
// writes p_k,i into the other buffer of a pair), which removes the p update
// pass and its dependency.
//
// Partial slots: kFzRz[0] / kFzRz[1] hold r.z of even / odd iterations (slot
// 0 is where k_init leaves r0.z0), kFzRr holds r.r, kFzPap holds p.Ap.
constexpr int kFzRz[2] = {0, 3};
constexpr int kFzRr = 4, kFzPap = 5;

// Sums of the nb partials of NS slots, each in fin_sum's order (16
// wavefront butterflies over 64-entry chunks, then the 16 chunk sums in
// sequence), computed by a 256-thread block and returned in every thread.
// All loads are issued before the first reduction.
//
// It also returns flags[0] as read by thread 0, the same in every thread of
// the block: the flag can be set by block 0 of the kernel that is running, so
// reading it per wavefront could split a block at a barrier.
template <int NS>
__device__ __forceinline__ int rsum256(const double *part, int nb, const int (&slot)[NS],
                                       double (&out)[NS], double (*lds)[16], const int *flags,
                                       int *stop) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int f = threadIdx.x == 0 ? flags[0] : 0;
    double v[NS][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = (w + 4 * q) * 64 + lane;
#pragma unroll
        for (int t = 0; t < NS; ++t) v[t][q] = i < nb ? part[slot[t] * kRedBlocks + i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const double sum = wave_sum(v[t][q]);
            if (lane == 0) lds[t][w + 4 * q] = sum;
        }
    if (threadIdx.x == 0) *stop = f;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        double acc = 0.0;
        for (int c = 0; c < 16; ++c) acc += lds[t][c];
        out[t] = acc;
    }
    return *stop;
}

// flags[0] as read by thread 0, in every thread of the block.
__device__ __forceinline__ int block_flag(const int *flags, int *stop) {
    if (threadIdx.x == 0) *stop = flags[0];
    __syncthreads();
    return *stop;
}

// Iteration k, first half.  For k > 1: records iteration k-1 (r.r -> relative
// residual, the stopping test of par_solver.rs:318-324) and stops there if it
// has converged; beta = r_{k-1}.z_{k-1} / r_{k-2}.z_{k-2}.  Then
// p_k = z (k = 1) or z + beta p_{k-1}, Ap_k = A p_k, partial p_k.Ap_k.
// The loads of a thread's first row (up to kPre entries), which do not depend
// on beta, are issued before the partial sums are reduced, so the two
// latency chains overlap.
constexpr int kPre = 8;
template <class T>
__global__ __launch_bounds__(256) void k_pcgj_spmv(size_t n, const uint64_t *__restrict__ rp,
                                                   const uint32_t *__restrict__ ci,
                                                   const T *__restrict__ v, const T *__restrict__ z,
                                                   const T *__restrict__ pold, T *__restrict__ pnew,
                                                   T *__restrict__ ap, double *__restrict__ part,
                                                   int nb, double *S, int *flags, int first,
                                                   int rz_new, int rz_old, double rtol, double atol) {
    __shared__ double lds[3][16];
    __shared__ int stop;
    const size_t i0 = blockIdx.x * (size_t)256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    T beta = 0;
    // preloaded first row: cnt = its length, or -1 when not preloaded
    int cnt = -1;
    T vv[kPre], zz[kPre], pp[kPre], z0 = 0, p0 = 0;
    if (!first && i0 < n) {
        const uint64_t s0 = rp[i0], len = rp[i0 + 1] - s0;
        z0 = z[i0];
        p0 = pold[i0];
        if (len <= (uint64_t)kPre) {
            cnt = (int)len;
#pragma unroll
            for (int k = 0; k < kPre; ++k) {
                if (k < cnt) {
                    const uint32_t c = ci[s0 + k];
                    vv[k] = v[s0 + k];
                    zz[k] = z[c];
                    pp[k] = pold[c];
                } else {
                    vv[k] = zz[k] = pp[k] = 0;
                }
            }
        }
    }
    if (first) {
        if (block_flag(flags, &stop)) return;
    } else {  // first is uniform: every thread of the block reaches the barriers
        double o[3];
        if (rsum256<3>(part, nb, {kFzRr, rz_new, rz_old}, o, lds, flags, &stop)) return;
        const double rr = o[0], rzn = o[1], rzo = o[2];
        const double res = sqrt(rr) / S[S_BNORM];
        const bool conv = res < rtol || sqrt(rr) < atol;
        const double b = rzn / rzo;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S[S_RR] = rr;
            S[S_RES] = res;
            S[S_ITERS] = S[S_ITERS] + 1.0;
            if (conv) {
                S[S_STATUS] = ST_CONVERGED;
                flags[0] = 1;
            } else {
                S[S_BETA] = b;
                S[S_RZ] = rzn;
            }
        }
        if (conv) return;
        beta = (T)b;
    }
    double d[1] = {0.0};
    for (size_t i = i0; i < n; i += stride) {
        T pi, s;
        if (first) {
            pi = z[i];
            s = row_dot_ref<T>(rp, ci, v, [z](uint32_t j) { return z[j]; }, i);
        } else if (i == i0 && cnt >= 0) {
            pi = fmad<T>(beta, p0, z0);
            T xk[kPre];
#pragma unroll
            for (int k = 0; k < kPre; ++k) xk[k] = fmad<T>(beta, pp[k], zz[k]);
            {  // row_dot_ref's order on the preloaded entries, without contraction
#pragma clang fp contract(off)
                s = 0;
                int done = 0;
                if constexpr (sizeof(T) == 8) {
                    done = cnt / 4 * 4;
#pragma unroll
                    for (int k = 0; k + 4 <= kPre; k += 4)
                        if (k + 4 <= done)
                            s += vv[k] * xk[k] + vv[k + 1] * xk[k + 1] + vv[k + 2] * xk[k + 2] +
                                 vv[k + 3] * xk[k + 3];
                }
#pragma unroll
                for (int k = 0; k < kPre; ++k)
                    if (k >= done && k < cnt) s += vv[k] * xk[k];
            }
        } else {
            pi = fmad<T>(beta, pold[i], z[i]);
            s = row_dot_ref<T>(rp, ci, v, [z, pold, beta](uint32_t j) { return fmad<T>(beta, pold[j], z[j]); }, i);
        }
        pnew[i] = pi;
        ap[i] = s;
        d[0] += (double)pi * (double)s;
    }
    block_sum<1, 256>(d);
    if (threadIdx.x == 0) part[kFzPap * kRedBlocks + blockIdx.x] = d[0];
}

// Iteration k, second half: alpha = r_{k-1}.z_{k-1} / p.Ap (breakdown below
// brk), x += alpha p, r -= alpha Ap, z = inv r, partials r.r and r.z.  The
// first entry's operands are loaded before the partial sums are reduced.
template <class T>
__global__ __launch_bounds__(256) void k_pcgj_update(size_t n, T *__restrict__ x,
                                                     const T *__restrict__ p, T *__restrict__ r,
                                                     const T *__restrict__ ap,
                                                     const T *__restrict__ inv, T *__restrict__ z,
                                                     double *__restrict__ part, int nb, double *S,
                                                     int *flags, int rz_in, int rz_out, double brk) {
    __shared__ double lds[2][16];
    __shared__ int stop;
    const size_t i0 = blockIdx.x * (size_t)256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    T x0 = 0, p0 = 0, r0 = 0, a0 = 0, q0 = 0;
    if (i0 < n) {
        x0 = x[i0];
        p0 = p[i0];
        r0 = r[i0];
        a0 = ap[i0];
        q0 = inv[i0];
    }
    double o[2];
    if (rsum256<2>(part, nb, {kFzPap, rz_in}, o, lds, flags, &stop)) return;
    const double pap = o[0], rz = o[1];
    const bool brkdn = fabs(pap) < brk;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S[S_PAP] = pap;
        if (brkdn) {
            S[S_STATUS] = ST_BREAKDOWN;
            flags[0] = 1;
        } else {
            S[S_ALPHA] = rz / pap;
        }
    }
    if (brkdn) return;
    const T alpha = (T)(rz / pap);
    double v[2] = {0, 0};
    for (size_t i = i0; i < n; i += stride) {
        T xi, pi, ri, ai, qi;
        if (i == i0) { xi = x0; pi = p0; ri = r0; ai = a0; qi = q0; }
        else { xi = x[i]; pi = p[i]; ri = r[i]; ai = ap[i]; qi = inv[i]; }
        x[i] = fmad<T>(alpha, pi, xi);
        ri = fmad<T>(-alpha, ai, ri);
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
        const T zi = qi * ri;
        z[i] = zi;
        v[1] += (double)ri * (double)zi;
    }
    block_sum<2, 256>(v);
    if (threadIdx.x == 0) {
        part[kFzRr * kRedBlocks + blockIdx.x] = v[0];
        part[rz_out * kRedBlocks + blockIdx.x] = v[1];
    }
}

// ------------------------------ fused PCG + Chebyshev-Jacobi iteration --
// The same finalise-in-every-workgroup scheme for the Chebyshev-preconditioned
// PCG: k_update_cheb0 takes alpha from the p.Ap partials (slot 0) and the
// previous r.z partials, updates x and r (partial r.r into slot 1) and starts
// the Chebyshev polynomial (z = d = inv r / theta); k_pupdate_beta takes beta
// from the r.z partials of this and the previous iteration and forms p.  With
// DEFER the update of x moves from k_update_cheb0 to k_pupdate_beta, which
// reads the old p anyway (one vector read fewer per iteration; the same
// x += alpha p_old, bit for bit): S_XPEND marks an x update that is pending,
// and k_x_fixup applies it after the loop when the iteration that set it was
// the last (a stop skips k_pupdate_beta).  The
// r.z partials alternate between slots kFcRz[0] / kFcRz[1] by iteration.  The
// one-block k_fin_conv between them keeps the stopping test ahead of the
// Chebyshev applies.  Iterates are bitwise those of the separate kernels.
constexpr int kFcRz[2] = {2, 3};

template <class T, bool DEFER>
__global__ __launch_bounds__(256) void k_update_cheb0(size_t n, double *S, T *__restrict__ x,
                                                      const T *__restrict__ p, T *__restrict__ r,
                                                      const T *__restrict__ ap, const T *__restrict__ inv,
                                                      T *__restrict__ z, T *__restrict__ d,
                                                      double *__restrict__ part, int nb, int *flags,
                                                      int rz_in, double inv_theta, int rz_out, double brk,
                                                      int write_z) {
    __shared__ double lds[2][16];
    __shared__ int stop;
    const size_t i0 = blockIdx.x * (size_t)256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    T x0 = 0, p0 = 0, r0 = 0, a0 = 0, q0 = 0;
    if (i0 < n) {
        if constexpr (!DEFER) {
            x0 = x[i0];
            p0 = p[i0];
        }
        r0 = r[i0];
        a0 = ap[i0];
        q0 = inv[i0];
    }
    double o[2];
    if (rsum256<2>(part, nb, {0, rz_in}, o, lds, flags, &stop)) return;
    const double pap = o[0], rz = o[1];
    const bool brkdn = fabs(pap) < brk;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S[S_PAP] = pap;
        if (brkdn) {
            S[S_STATUS] = ST_BREAKDOWN;
            flags[0] = 1;
        } else {
            S[S_ALPHA] = rz / pap;
            if (DEFER) S[S_XPEND] = 1;
        }
    }
    if (brkdn) return;
    const T alpha = (T)(rz / pap);
    double v[2] = {0, 0};
    for (size_t i = i0; i < n; i += stride) {
        T ri, ai, qi;
        if (i == i0) { ri = r0; ai = a0; qi = q0; }
        else { ri = r[i]; ai = ap[i]; qi = inv[i]; }
        if constexpr (!DEFER) {
            const T xi = i == i0 ? x0 : x[i], pi = i == i0 ? p0 : p[i];
            x[i] = fmad<T>(alpha, pi, xi);
        }
        ri = fmad<T>(-alpha, ai, ri);
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
        const T zi = (T)inv_theta * (qi * ri);
        if (write_z) z[i] = zi;
        d[i] = zi;
        if (rz_out >= 0) v[1] += (double)ri * (double)zi;
    }
    block_sum<2, 256>(v);
    if (threadIdx.x == 0) {
        part[1 * kRedBlocks + blockIdx.x] = v[0];
        if (rz_out >= 0) part[rz_out * kRedBlocks + blockIdx.x] = v[1];
    }
}

template <class T, bool DEFER>
__global__ __launch_bounds__(256) void k_pupdate_beta(size_t n, double *S, const T *__restrict__ z,
                                                      T *__restrict__ p, const double *__restrict__ part,
                                                      int nb, int *flags, int rz_new, int rz_old,
                                                      T *__restrict__ x) {
    __shared__ double lds[2][16];
    __shared__ int stop;
    const size_t i0 = blockIdx.x * (size_t)256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    T z0 = 0, p0 = 0, x0 = 0;
    if (i0 < n) {
        z0 = z[i0];
        p0 = p[i0];
        if constexpr (DEFER) x0 = x[i0];
    }
    double o[2];
    if (rsum256<2>(part, nb, {rz_new, rz_old}, o, lds, flags, &stop)) return;
    const double b = o[0] / o[1];
    // DEFER: alpha as k_update_cheb0 rounded it, (T)(rz / pAp)
    const T alpha = DEFER ? (T)S[S_ALPHA] : T(0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S[S_BETA] = b;
        S[S_RZ] = o[0];
        if (DEFER) S[S_XPEND] = 0;  // read only by k_x_fixup, after the loop
    }
    const T beta = (T)b;
    for (size_t i = i0; i < n; i += stride) {
        const T zi = i == i0 ? z0 : z[i];
        const T pi = i == i0 ? p0 : p[i];
        if constexpr (DEFER) x[i] = fmad<T>(alpha, pi, i == i0 ? x0 : x[i]);
        p[i] = fmad<T>(beta, pi, zi);
    }
}

// The x update of the last iteration, when k_pupdate_beta did not run (see
// k_update_cheb0); not skipped by the stop flag.
template <class T>
__global__ __launch_bounds__(256) void k_x_fixup(size_t n, const double *S, T *__restrict__ x,
                                                 const T *__restrict__ p) {
    if (S[S_XPEND] == 0.0) return;
    const T alpha = (T)S[S_ALPHA];
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        x[i] = fmad<T>(alpha, p[i], x[i]);
}

// After the last enqueued iteration: records it and applies the stopping test.
__global__ __launch_bounds__(256) void k_pcgj_end(const double *part, int nb, double *S, int *flags,
                                                  double rtol, double atol) {
    __shared__ double lds[1][16];
    __shared__ int stop;
    double o[1];
    if (rsum256<1>(part, nb, {kFzRr}, o, lds, flags, &stop)) return;
    const double rr = o[0];
    if (threadIdx.x) return;
    const double res = sqrt(rr) / S[S_BNORM];
    S[S_RR] = rr;
    S[S_RES] = res;
    S[S_ITERS] = S[S_ITERS] + 1.0;
    if (res < rtol || sqrt(rr) < atol) {
        S[S_STATUS] = ST_CONVERGED;
        flags[0] = 1;
    }
}

// FEMGPU_PCG_FUSED=0 selects the separate-kernel PCG-Jacobi loop (for
// comparisons); the fused loop is the default.
static bool fused_pcg_enabled() {
    const char *e = std::getenv("FEMGPU_PCG_FUSED");
    return !(e && e[0] == '0');
}
// The fused PCG-Jacobi loop replays captured chunks of iterations from a HIP
// graph only when FEMGPU_GRAPHS=1.  The replay saved 0.4 us of a 15 us
// iteration (66K rows); a profiled run (rocprofv3 --kernel-trace) of a solve
// with the replay on died with a host SIGSEGV inside the runtime's graph
// launch, which the code here could not be shown to cause, so direct
// launches are the default.
static bool graphs_enabled() {
    const char *e = std::getenv("FEMGPU_GRAPHS");
    return e && e[0] == '1';
}

static unsigned vgrid(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

// ------------------------------------------------ allocation-free workspace --
// The work vectors of femgpu_solve_pcg_jacobi_noalloc: one file-scope device
// array, which each GPU holds (its own copy, at its own address) from the
// loading of this code object until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host it is only a stand-in for the
// symbol): noalloc_scratch() looks up the device address once per device and
// every use goes through that pointer.
static __device__ __align__(256) unsigned char scratch[FEMGPU_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[64];

// Address of this device's scratch array; `device` must be current.  Called
// at context creation, so that the lookup is done before any solve.
void *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    FG_REQUIRE(device >= 0 && device < 64, FEMGPU_ERR_INVALID_ARG, "noalloc: device index out of range");
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        FG_REQUIRE(e == hipSuccess && ptr, FEMGPU_ERR_HIP,
                   std::string("noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[device] = ptr;
    }
    return g_scratch_dev[device];
}

// ------------------------------------------------------------ engine --
struct ChebParams {
    int degree = 4;
    double lmin = 0, lmax = 0;
};

template <class T> struct Krylov {
    femgpu_op *op;
    femgpu_ctx *c;
    size_t n;
    hipStream_t s;
    double *S;
    int *flags;
    double *part;
    // workspace slots
    enum { W_R = 0, W_Z, W_P, W_AP, W_INV, W_D, W_T, W_DIAG, W_X, W_B, W_COUNT };
    // preconditioner for PREC_CB: z = M^-1 r (kernels may honour the stop flag)
    std::function<void(const T *, T *)> pc;
    // Caller-provided workspace of W_COUNT vectors, ext_stride entries apart
    // (the allocation-free entry point), or null: the operator's workspace.
    T *ext = nullptr;
    size_t ext_stride = 0;
    // Reporting (SolverConfig::effective_print_level, solver/src/lib.rs:133-142):
    // 0 silent, 1 summary line, 2+ per-iteration residuals, recorded on the
    // device into the context's history buffer and printed after the solve.
    int level = 0;
    const char *label = "pcg";
    uint64_t hist_cap = 0;
    void record() {
        if (hist_cap)
            hipLaunchKernelGGL(k_hist, dim3(1), dim3(1), 0, s, (const double *)S, c->hist.as<double>(), hist_cap);
    }
    void start_history(uint64_t max_iter) {
        hist_cap = 0;
        if (level < 2) return;
        hist_cap = std::min<uint64_t>(max_iter + 1, kHistMax);
        c->hist.ensure(hist_cap * sizeof(double));
        c->hist_len = 0;
    }
    T *vec(int slot) {
        return ext ? ext + (size_t)slot * ext_stride : (T *)op->wsvec(slot, n * sizeof(T));
    }

    explicit Krylov(femgpu_op *o) : op(o), c(o->ctx), n(o->n), s(o->ctx->stream) {
        S = c->scalars.as<double>();
        flags = c->flags.as<int>();
        part = c->partials.as<double>();
        FG_REQUIRE(o->esize() == sizeof(T), FEMGPU_ERR_INVALID_ARG, "solver: operator dtype mismatch");
    }

    void apply(const T *x, T *y, bool use_skip) { op->apply_dev(x, y, false, use_skip ? flags : nullptr); }

    // diag(A) in the W_DIAG slot, computed once per operator
    T *diag() {
        T *d = vec(W_DIAG);
        if (ext) {
            op->diagonal_dev(d);  // the shared workspace keeps no state between calls
        } else if (!op->diag_ready) {
            op->diagonal_dev(d);
            op->diag_ready = true;
        }
        return d;
    }

    // inv = 1/diag (1 where |diag| <= tiny); returns pointer to the W_INV slot
    T *build_inv(double tiny) {
        T *d = diag(), *inv = vec(W_INV);
        hipLaunchKernelGGL(k_inv_tiny<T>, dim3(vgrid(n)), dim3(256), 0, s, n, (const T *)d, inv, tiny);
        return inv;
    }

    void read_scalars() {
        FG_HIP(hipMemcpyAsync(c->pinned, S, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
    }
    int read_flag() {
        int f = 0;
        FG_HIP(hipMemcpyAsync(c->pinned + 40, flags, sizeof(int), hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
        std::memcpy(&f, c->pinned + 40, sizeof(int));
        return f;
    }

    // z = cheb(r) with partial r.z into part slot rz_slot (skippable).  With
    // from1 the first step (z = d = inv r / theta) has been done by the caller,
    // which for degree >= 2 wrote d only.  For degree >= 2 the first step
    // leaves z unwritten and the next one applies A to d and forms z = d + d'
    // (one vector write and one read fewer, the same values).
    void cheb_apply(const ChebParams &cp, const T *r, const T *inv, T *z, int rz_slot = 0,
                    bool from1 = false) {
        T *d = vec(W_D), *t = vec(W_T);
        double *pz = part + (size_t)rz_slot * kRedBlocks;
        const double theta = 0.5 * (cp.lmax + cp.lmin), delta = 0.5 * (cp.lmax - cp.lmin);
        const double sigma = theta / delta;
        double rho_prev = 1.0 / sigma;
        const unsigned nb = nblk(n);
        if (!from1)
            hipLaunchKernelGGL(k_cheb0<T>, dim3(nb), dim3(256), 0, s, n, r, inv, z, d, 1.0 / theta,
                               cp.degree == 1 ? 1 : 0, pz, flags, cp.degree == 1 ? 1 : 0);
        for (int k = 1; k < cp.degree; ++k) {
            const double rho = 1.0 / (2.0 * sigma - rho_prev);
            const int last = k == cp.degree - 1 ? 1 : 0;
            const int first = k == 1 ? 1 : 0;
            const T *xin = first ? d : z;
            if (!op->apply_cheb_dev(xin, r, inv, z, d, rho * rho_prev, 2.0 * rho / delta, last, pz, flags,
                                    first, 1.0 / theta)) {
                apply(xin, t, true);
                auto kk = first ? (last ? k_chebk<T, true, true> : k_chebk<T, true, false>)
                                : (last ? k_chebk<T, false, true> : k_chebk<T, false, false>);
                hipLaunchKernelGGL(kk, dim3(nb), dim3(256), 0, s, n, r, (const T *)t, inv, z, d,
                                   rho * rho_prev, 2.0 * rho / delta, pz, flags, 1.0 / theta);
            }
            rho_prev = rho;
        }
    }

    // Preconditioned CG on device vectors b, x (x: initial guess in, solution out).
    // Returns status code (0 ok), fills result.
    int solve(int mode, int prec, const ChebParams &cp, const T *b, T *x, double rtol, double atol,
              uint64_t max_iter, femgpu_solve_result *res) {
        const unsigned nb = nblk(n);
        T *r = vec(W_R), *z = vec(W_Z), *p = vec(W_P), *ap = vec(W_AP);
        const T *inv = nullptr;
        if (prec == PREC_JACOBI) inv = build_inv(1e-30);
        else if (prec == PREC_CHEB) inv = vec(W_INV);  // caller prepared it
        // r0 = b - A x0
        apply(x, ap, false);
        if (prec == PREC_JACOBI)
            hipLaunchKernelGGL((k_init<T, true>), dim3(nb), dim3(256), 0, s, n, b, (const T *)ap, inv, r,
                               z, p, part, true);
        else
            hipLaunchKernelGGL((k_init<T, false>), dim3(nb), dim3(256), 0, s, n, b, (const T *)ap, inv,
                               r, z, p, part, prec == PREC_NONE);
        hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(1024), 0, s, (const double *)part, (int)nb, S,
                           flags, mode, rtol, atol);
        start_history(max_iter);
        record();
        CsrView cv;
        if (mode == MODE_PCG && prec == PREC_JACOBI && op->csr_view(&cv) && fused_pcg_enabled())
            return run_pcgj_fused(cv, x, inv, rtol, atol, max_iter, res);
        if (prec == PREC_CHEB) {
            // z = M^-1 r0; p = z; rz = r.z (partials in the slot the fused loop reads first)
            cheb_apply(cp, r, inv, z, kFcRz[0]);
            hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(1024), 0, s,
                               (const double *)(part + (size_t)kFcRz[0] * kRedBlocks), (int)nb, S, (int)S_RZ);
            copy<T>(s, p, z, n);
        } else if (prec == PREC_CB) {
            pc(r, z);
            hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, s, n, (const T *)r, (const T *)z, part, 0, flags);
            hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(1024), 0, s, (const double *)part, (int)nb, S, (int)S_RZ);
            copy<T>(s, p, z, n);
        }
        return run(mode, prec, cp, x, inv, rtol, atol, max_iter, res, false);
    }

    // PCG-Jacobi on an operator that exposes its CSR arrays: the two-kernel
    // iteration (k_pcgj_spmv, k_pcgj_update), same iterates as run().  After
    // the first iteration, whole chunks of kPcgChunk iterations are replayed
    // from a HIP graph cached on the operator (launch-bound small systems);
    // the host reads the stop flag after every chunk.
    static constexpr uint64_t kPcgChunk = 32;
    int run_pcgj_fused(const CsrView &cv, T *x, const T *inv, double rtol, double atol,
                       uint64_t max_iter, femgpu_solve_result *res) {
        const unsigned nb = nblk(n);
        T *r = vec(W_R), *z = vec(W_Z), *ap = vec(W_AP);
        T *pb[2] = {vec(W_P), vec(W_D)};
        const T *vals = (const T *)cv.vals;
        const double brk = 1e-30;
        auto iter = [&](uint64_t k) {
            T *pn = pb[k & 1];
            const T *po = pb[(k + 1) & 1];
            hipLaunchKernelGGL(k_pcgj_spmv<T>, dim3(nb), dim3(256), 0, s, n, cv.rp, cv.ci, vals,
                               (const T *)z, po, pn, ap, part, (int)nb, S, flags, k == 1 ? 1 : 0,
                               kFzRz[(k - 1) & 1], kFzRz[k & 1], rtol, atol);
            hipLaunchKernelGGL(k_pcgj_update<T>, dim3(nb), dim3(256), 0, s, n, x, (const T *)pn, r,
                               (const T *)ap, inv, z, part, (int)nb, S, flags, kFzRz[(k - 1) & 1],
                               kFzRz[k & 1], brk);
            if (k > 1) record();  // the spmv finalised iteration k - 1
        };
        // instantiating a graph allocates; the history records between the kernels
        const bool graphs = graphs_enabled() && !ext && !hist_cap;
        uint64_t it = 0, since = 0;
        const uint64_t check = n < (1u << 20) ? kPcgChunk : 8;
        bool stopped = read_flag() != 0;
        while (!stopped && it < max_iter) {
            // a chunk starts at an even k, so its slot parities are fixed
            if (graphs && (it & 1) && max_iter - it >= kPcgChunk) {
                FG_HIP(hipGraphLaunch(chunk_graph(cv, x, inv, rtol, atol, iter), s));
                it += kPcgChunk;
                since = check;
            } else {
                iter(it + 1);
                FG_HIP(hipGetLastError());
                ++it;
                ++since;
            }
            if (since >= check || it == max_iter) {
                stopped = read_flag() != 0;
                since = 0;
            }
        }
        if (it > 0) {
            hipLaunchKernelGGL(k_pcgj_end, dim3(1), dim3(256), 0, s, (const double *)part, (int)nb, S,
                               flags, rtol, atol);
            record();
        }
        return finish(MODE_PCG, max_iter, res);
    }

    // The graph of iterations k = 2 .. kPcgChunk + 1 (any chunk starting at an
    // even k launches the same kernels), captured once per set of buffers.
    template <class Iter>
    hipGraphExec_t chunk_graph(const CsrView &cv, T *x, const T *inv, double rtol, double atol,
                               Iter &iter) {
        IterGraph &g = op->pcg_graph;
        std::vector<const void *> key = {x, inv, vec(W_R), vec(W_Z), vec(W_AP), vec(W_P), vec(W_D),
                                         part, S, flags, cv.vals, cv.rp, cv.ci};
        std::vector<double> params = {rtol, atol, (double)n};
        if (g.exec && g.key == key && g.params == params) return g.exec;
        g.reset();
        hipGraph_t graph = nullptr;
        FG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (uint64_t j = 0; j < kPcgChunk; ++j) iter(2 + j);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        FG_HIP(le);
        FG_HIP(ce);
        const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) g.exec = nullptr;
        FG_HIP(ie);
        g.key = key;
        g.params = params;
        return g.exec;
    }

    // Continues PCG + Chebyshev-Jacobi on the residual the caller put in
    // vec(W_R) (e.g. a residual recomputed in higher precision), with x
    // accumulating the correction.  first: p = z; otherwise the previous search
    // direction is kept and p = z + beta p, beta = r.z / (previous r.z), so the
    // Krylov recurrence goes on across the replacement.  Relative residuals are
    // against bnorm.
    int solve_cont(bool first, const ChebParams &cp, T *x, double bnorm, double rtol, uint64_t max_iter,
                   femgpu_solve_result *res) {
        const unsigned nb = nblk(n);
        T *r = vec(W_R), *z = vec(W_Z), *p = vec(W_P);
        const T *inv = vec(W_INV);
        hipLaunchKernelGGL(k_set_cont, dim3(1), dim3(1), 0, s, S, flags, bnorm);
        cheb_apply(cp, r, inv, z, kFcRz[0]);
        const double *prz = part + (size_t)kFcRz[0] * kRedBlocks;
        if (first) {
            hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(1024), 0, s, prz, (int)nb, S, (int)S_RZ);
            copy<T>(s, p, z, n);
        } else {
            hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(1024), 0, s, prz, (int)nb, S, flags);
            hipLaunchKernelGGL(k_pupdate<T>, dim3(vgrid(n)), dim3(256), 0, s, n, (const double *)S, (const T *)z,
                               p, flags);
        }
        return run(MODE_PCG, PREC_CHEB, cp, x, inv, rtol, 0.0, max_iter, res, true);
    }

    // The PCG iteration loop and its result, shared by solve / solve_cont.
    // segment: the last iteration stops after the update of x and r (no
    // preconditioning and no new search direction), so that solve_cont can
    // go on from a replaced residual with the last search direction.
    int last_status = ST_RUNNING;
    int run(int mode, int prec, const ChebParams &cp, T *x, const T *inv, double rtol, double atol,
            uint64_t max_iter, femgpu_solve_result *res, bool segment) {
        const unsigned nb = nblk(n);
        T *r = vec(W_R), *z = vec(W_Z), *p = vec(W_P), *ap = vec(W_AP);
        const double brk = mode == MODE_CG ? 1e-32 : 1e-30;
        if (prec == PREC_CHEB && mode == MODE_PCG && fused_pcg_enabled())
            return run_cheb_fused(cp, x, inv, rtol, atol, max_iter, res, segment);
        uint64_t it = 0;
        bool stopped = read_flag() != 0;
        const uint64_t check = n < (1u << 20) ? 16 : 4;
        while (!stopped && it < max_iter) {
            if (!op->apply_dot_dev(p, ap, part, flags)) {
                apply(p, ap, true);
                hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, s, n, (const T *)p, (const T *)ap,
                                   part, 0, flags);
            }
            hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(1024), 0, s, (const double *)part, (int)nb, S,
                               flags, brk);
            if (prec == PREC_JACOBI)
                hipLaunchKernelGGL((k_update<T, true>), dim3(nb), dim3(256), 0, s, n, (const double *)S,
                                   x, (const T *)p, r, (const T *)ap, inv, z, part, flags);
            else
                hipLaunchKernelGGL((k_update<T, false>), dim3(nb), dim3(256), 0, s, n, (const double *)S,
                                   x, (const T *)p, r, (const T *)ap, inv, z, part, flags);
            hipLaunchKernelGGL(k_fin_conv, dim3(1), dim3(1024), 0, s, (const double *)part, (int)nb, S,
                               flags, mode, prec == PREC_JACOBI ? 1 : 0, prec == PREC_NONE || prec == PREC_JACOBI,
                               rtol, atol, (double)(it + 1));
            record();
            if (segment && it + 1 == max_iter) {
                ++it;
                break;
            }
            if (prec == PREC_CHEB) {
                cheb_apply(cp, r, inv, z);
                hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(1024), 0, s, (const double *)part, (int)nb, S,
                                   flags);
            } else if (prec == PREC_CB) {
                pc(r, z);
                hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, s, n, (const T *)r, (const T *)z, part, 0,
                                   flags);
                hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(1024), 0, s, (const double *)part, (int)nb, S,
                                   flags);
            }
            hipLaunchKernelGGL(k_pupdate<T>, dim3(vgrid(n)), dim3(256), 0, s, n, (const double *)S,
                               (const T *)(prec == PREC_NONE ? r : z), p, flags);
            FG_HIP(hipGetLastError());
            ++it;
            if (it % check == 0 || it == max_iter) stopped = read_flag() != 0;
        }
        return finish(mode, max_iter, res);
    }

    // run() for PCG + Chebyshev-Jacobi with the fused vector kernels
    // (k_update_cheb0, k_pupdate_beta); same iterates, same segment semantics.
    int run_cheb_fused(const ChebParams &cp, T *x, const T *inv, double rtol, double atol,
                       uint64_t max_iter, femgpu_solve_result *res, bool segment) {
        const unsigned nb = nblk(n);
        T *r = vec(W_R), *z = vec(W_Z), *p = vec(W_P), *ap = vec(W_AP), *d = vec(W_D);
        const double brk = 1e-30;
        const double inv_theta = 1.0 / (0.5 * (cp.lmax + cp.lmin));
        uint64_t it = 0;
        bool stopped = read_flag() != 0;
        const uint64_t check = n < (1u << 20) ? 16 : 4;
        while (!stopped && it < max_iter) {
            const uint64_t k = it + 1;
            const int rz_prev = kFcRz[(k - 1) & 1], rz_new = kFcRz[k & 1];
            if (!op->apply_dot_dev(p, ap, part, flags)) {
                apply(p, ap, true);
                hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, s, n, (const T *)p, (const T *)ap,
                                   part, 0, flags);
            }
            hipLaunchKernelGGL((k_update_cheb0<T, true>), dim3(nb), dim3(256), 0, s, n, S, x, (const T *)p, r,
                               (const T *)ap, inv, z, d, part, (int)nb, flags, rz_prev, inv_theta,
                               cp.degree == 1 ? rz_new : -1, brk,
                               (cp.degree == 1 || (segment && it + 1 == max_iter)) ? 1 : 0);
            hipLaunchKernelGGL(k_fin_conv, dim3(1), dim3(1024), 0, s, (const double *)(part + kRedBlocks),
                               (int)nb, S, flags, (int)MODE_PCG, 0, 0, rtol, atol, (double)(it + 1));
            record();
            if (segment && it + 1 == max_iter) {
                ++it;
                break;
            }
            cheb_apply(cp, r, inv, z, rz_new, true);
            hipLaunchKernelGGL((k_pupdate_beta<T, true>), dim3(nb), dim3(256), 0, s, n, S, (const T *)z, p,
                               (const double *)part, (int)nb, flags, rz_new, rz_prev, x);
            FG_HIP(hipGetLastError());
            ++it;
            if (it % check == 0 || it == max_iter) stopped = read_flag() != 0;
        }
        hipLaunchKernelGGL(k_x_fixup<T>, dim3(nb), dim3(256), 0, s, n, (const double *)S, x, (const T *)p);
        return finish(MODE_PCG, max_iter, res);
    }

    // Iterations the last solve completed (x and r updated), whatever its
    // status: at a breakdown the result reports max_iter, as the reference
    // does, but fewer ran.  Valid after finish().
    uint64_t iterations_done() const { return (uint64_t)c->pinned[S_ITERS]; }

    // Copies the recorded residuals to the context and prints what the level asks for.
    void report(int status, uint64_t iters, double resid, bool relative) {
        if (level >= 2 && hist_cap) {
            const uint64_t len = std::min<uint64_t>(iters + 1, hist_cap);
            c->hist_host.resize(len);
            FG_HIP(hipMemcpyAsync(c->hist_host.data(), c->hist.ptr, len * sizeof(double), hipMemcpyDeviceToHost, s));
            FG_HIP(hipStreamSynchronize(s));
            c->hist_len = len;
            for (uint64_t k = 1; k < len; ++k)
                std::fprintf(stderr, "femgpu %s iter %llu: residual = %.3e\n", label, (unsigned long long)k,
                             c->hist_host[k]);
        }
        if (level >= 1)
            std::fprintf(stderr, "femgpu %s: %s after %llu iterations, %s residual = %.3e\n", label,
                         status == ST_CONVERGED ? "converged" : status == ST_BREAKDOWN ? "breakdown" : "not converged",
                         (unsigned long long)iters, relative ? "relative" : "absolute", resid);
    }

    // Reads the scalars after a solve and fills the result.
    int finish(int mode, uint64_t max_iter, femgpu_solve_result *res) {
        read_scalars();
        const double *h = c->pinned;
        const int status = (int)h[S_STATUS];
        last_status = status;
        if (level >= 1)
            report(status, (uint64_t)h[S_ITERS], h[S_RES], mode == MODE_PCG);
        res->converged = status == ST_CONVERGED;
        res->final_residual_is_relative = mode == MODE_PCG;
        res->error_estimate = 0.0;
        if (status == ST_CONVERGED) {
            res->iterations = (uint64_t)h[S_ITERS];
            res->final_residual = h[S_RES];
            return FEMGPU_OK;
        }
        if (status == ST_BREAKDOWN) {
            if (mode == MODE_CG) {
                res->iterations = (uint64_t)h[S_ITERS];
                res->final_residual = h[S_RES];
                return FEMGPU_ERR_BREAKDOWN;
            }
            // par_solve_pcg_jacobi: break, then converged = false, iterations = max_iter
            // and the relative residual of the current r (par_solver.rs:339-345)
        }
        res->iterations = max_iter;
        res->final_residual = mode == MODE_CG ? h[S_RES] : std::sqrt(h[S_RR]) / h[S_BNORM];
        return FEMGPU_ERR_NOT_CONVERGED;
    }

    // Largest eigenvalue of D^-1 A by power iteration from 1/sqrt(n).
    double power_dinv_a(const T *diag, int iters) {
        T *v = vec(W_R), *w = vec(W_Z);
        fill<T>(s, v, (T)(1.0 / std::sqrt((double)n)), n);
        const unsigned nb = nblk(n);
        for (int k = 0; k < iters; ++k) {
            apply(v, w, false);
            hipLaunchKernelGGL(k_div<T>, dim3(vgrid(n)), dim3(256), 0, s, n, w, diag);
            hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, s, n, (const T *)w, (const T *)w, part, 0,
                               nullptr);
            hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(1024), 0, s, (const double *)part, (int)nb, S, (int)S_RR0);
            hipLaunchKernelGGL(k_scale_to<T>, dim3(vgrid(n)), dim3(256), 0, s, n, (const T *)w, v,
                               (const double *)S, (int)S_RR0);
        }
        read_scalars();
        return std::sqrt(c->pinned[S_RR0]);
    }

    // Prepares W_INV = 1/diag and the Chebyshev interval.
    ChebParams prepare_cheb(const femgpu_cheb_config *cc) {
        ChebParams cp;
        cp.degree = cc && cc->degree > 0 ? cc->degree : 4;
        const double ratio = cc && cc->ratio > 1.0 ? cc->ratio : 30.0;
        const double safety = cc && cc->safety > 0 ? cc->safety : 1.1;
        const int iters = cc && cc->power_iters > 0 ? cc->power_iters : 20;
        T *d = diag(), *inv = vec(W_INV);
        hipLaunchKernelGGL(k_inv_tiny<T>, dim3(vgrid(n)), dim3(256), 0, s, n, (const T *)d, inv, 0.0);
        const double est = cc && cc->lmax > 0 ? cc->lmax / safety : power_dinv_a(d, iters);
        cp.lmax = safety * est;
        cp.lmin = cp.lmax / ratio;
        return cp;
    }
};

// Host/device vector staging for the C entry points.
struct Staged {
    femgpu_op *op;
    const void *b = nullptr;
    void *x = nullptr, *x_host = nullptr;
    size_t bytes;
    Staged(femgpu_op *o, const void *b_, void *x_, femgpu_mem mem, int slot_b, int slot_x)
        : op(o), bytes(o->n * o->esize()) {
        if (mem == FEMGPU_DEVICE) { b = b_; x = x_; return; }
        void *bd = op->wsvec(slot_b, bytes), *xd = op->wsvec(slot_x, bytes);
        FG_HIP(hipMemcpyAsync(bd, b_, bytes, hipMemcpyHostToDevice, op->ctx->stream));
        FG_HIP(hipMemcpyAsync(xd, x_, bytes, hipMemcpyHostToDevice, op->ctx->stream));
        b = bd; x = xd; x_host = x_;
    }
    void finish() {
        if (x_host) FG_HIP(hipMemcpyAsync(x_host, x, bytes, hipMemcpyDeviceToHost, op->ctx->stream));
        FG_HIP(hipStreamSynchronize(op->ctx->stream));
    }
};

// Work vectors of a Krylov solve (its workspace slots) and the largest n the
// allocation-free entry point takes in each precision.
constexpr int KrylovSlots = Krylov<double>::W_COUNT;
static size_t noalloc_max_n(femgpu_dtype dt) {
    const size_t per = (size_t)FEMGPU_NOALLOC_SCRATCH_BYTES / KrylovSlots / (dt == FEMGPU_F64 ? 8 : 4);
    return per / 64 * 64;
}

template <class F> static int dispatch(femgpu_op *op, F &&f) {
    if (op->dtype == FEMGPU_F64) { Krylov<double> k(op); return f(k); }
    Krylov<float> k(op);
    return f(k);
}

static void check_cfg(const femgpu_solver_config *cfg) {
    FG_REQUIRE(cfg, FEMGPU_ERR_INVALID_ARG, "solver: cfg is NULL");
    FG_REQUIRE(cfg->rtol >= 0 && cfg->atol >= 0, FEMGPU_ERR_INVALID_ARG, "solver: negative tolerance");
}

// SolverConfig::effective_print_level (solver/src/lib.rs:133-142): print_level
// when it is not Silent, else Iterations when verbose, else Silent.
static int print_level(const femgpu_solver_config *cfg) {
    if (cfg->print_level > 0) return cfg->print_level;
    return cfg->verbose ? 2 : 0;
}

}  // namespace femgpu

using namespace femgpu;

// FP64 iterative refinement around an FP32 inner solver:
// inner(K32, cp, r32, e32, rtol, max_iter, &result) solves A32 e = r32.
template <class Inner>
static int ir_refine(femgpu_op *op64, femgpu_op *op32, const femgpu_ir_config *ir, const void *b,
                     void *x, femgpu_mem mem, const femgpu_solver_config *cfg,
                     femgpu_solve_result *res, uint64_t *inner_iterations, Inner &&inner) {
    FG_REQUIRE(op64 && op32 && ir && b && x && res, FEMGPU_ERR_INVALID_ARG,
               "solve_ir_mixed: NULL argument");
    FG_REQUIRE(op64->dtype == FEMGPU_F64 && op32->dtype == FEMGPU_F32, FEMGPU_ERR_INVALID_ARG,
               "solve_ir_mixed: need an f64 and an f32 operator");
    FG_REQUIRE(op64->n == op32->n && op64->ctx == op32->ctx, FEMGPU_ERR_DIM_MISMATCH,
               "solve_ir_mixed: operators differ in size or context");
    check_cfg(cfg);
    femgpu_ctx *c = op64->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    FG_HIP(hipSetDevice(c->device));
    Staged st(op64, b, x, mem, 8, 9);
    Krylov<double> K64(op64);
    Krylov<float> K32(op32);
    const size_t n = op64->n;
    hipStream_t s = c->stream;
    double *r64 = K64.vec(K64.W_R);
    float *r32 = K32.vec(K32.W_B), *e32 = K32.vec(K32.W_X);
    const double *bd = (const double *)st.b;
    double *xd = (double *)st.x;
    ChebParams cp = K32.prepare_cheb(&ir->cheb);
    const double bnorm = std::sqrt(dot_host(c, bd, bd, n, FEMGPU_F64));
    uint64_t inner_total = 0, outer = 0;
    double rel = 0.0;
    auto residual = [&]() {
        op64->apply_dev(xd, r64, false, nullptr);
        hipLaunchKernelGGL(k_sub_from<double>, dim3(vgrid(n)), dim3(256), 0, s, n, bd, r64);
        return std::sqrt(dot_host(c, r64, r64, n, FEMGPU_F64));
    };
    double rnorm = residual();
    rel = bnorm > 0 ? rnorm / bnorm : 0.0;
    const uint64_t max_outer = ir->max_outer ? ir->max_outer : 50;
    const double inner_rtol = ir->inner_rtol > 0 ? ir->inner_rtol : 1e-4;
    const uint64_t inner_max = ir->inner_max_iter ? ir->inner_max_iter : cfg->max_iter;
    while (bnorm > 0 && !(rel < cfg->rtol) && outer < max_outer && inner_total < cfg->max_iter) {
        convert_f64_to_f32(s, r32, r64, n, 1.0 / rnorm);
        fill<float>(s, e32, 0.f, n);
        femgpu_solve_result ir_res{};
        // The inner problem is scaled to |r| = 1, so an inner relative
        // residual t leaves about t |r| outside: solve no further than the
        // outer target needs (half of it, for the FP32 / FP64 residual gap).
        const double t = std::max(inner_rtol, 0.5 * cfg->rtol * bnorm / rnorm);
        const uint64_t cap = std::min<uint64_t>(inner_max, cfg->max_iter - inner_total);
        const int rc = inner(K32, cp, r32, e32, t, cap, &ir_res);
        inner_total += ir_res.iterations;
        (void)rc;  // an inner solve that stops early still yields a correction
        convert_f32_to_f64_axpy(s, xd, e32, n, rnorm);
        const double prev = rel;
        rnorm = residual();
        rel = rnorm / bnorm;
        ++outer;
        if (!(rel < prev)) break;  // stagnation: stop refining
    }
    st.finish();
    if (inner_iterations) *inner_iterations = inner_total;
    res->iterations = outer;
    res->final_residual = rel;
    res->final_residual_is_relative = 1;
    res->converged = bnorm == 0 || rel < cfg->rtol;
    res->error_estimate = 0.0;
    if (!res->converged)
        return fail(FEMGPU_ERR_NOT_CONVERGED,
                    "mixed-precision refinement stopped at relative residual " + std::to_string(rel));
    return 0;
}

// Continued refinement: one FP32 PCG + Chebyshev-Jacobi run whose recursive
// residual is replaced by the FP64 residual r = b - A64 u (after the FP32
// correction has been added to u in FP64) each time it has dropped by the
// factor delta since the last replacement (or after max_seg iterations).  The
// search direction survives each replacement (the Krylov recurrence goes on),
// so the FP32 iterations are not lost to restarts as in ir_refine, and the
// FP32 recursive residual never has to be accurate far below the residual it
// started from.  Stops on the FP64 relative residual; a breakdown in a segment
// restarts with p = z.
static int ir_continued(femgpu_op *op64, femgpu_op *op32, const femgpu_ir_config *ir, const void *b,
                        void *x, femgpu_mem mem, const femgpu_solver_config *cfg,
                        femgpu_solve_result *res, uint64_t *inner_iterations) {
    femgpu_ctx *c = op64->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    FG_HIP(hipSetDevice(c->device));
    Staged st(op64, b, x, mem, 8, 9);
    Krylov<double> K64(op64);
    Krylov<float> K32(op32);
    const size_t n = op64->n;
    hipStream_t s = c->stream;
    double *r64 = K64.vec(K64.W_R);
    float *r32 = K32.vec(K32.W_R), *x32 = K32.vec(K32.W_X);
    const double *bd = (const double *)st.b;
    double *xd = (double *)st.x;
    ChebParams cp = K32.prepare_cheb(&ir->cheb);
    const double bnorm = std::sqrt(dot_host(c, bd, bd, n, FEMGPU_F64));
    auto residual = [&]() {
        op64->apply_dev(xd, r64, false, nullptr);
        hipLaunchKernelGGL(k_sub_from<double>, dim3(vgrid(n)), dim3(256), 0, s, n, bd, r64);
        return std::sqrt(dot_host(c, r64, r64, n, FEMGPU_F64));
    };
    double rel = bnorm > 0 ? residual() / bnorm : 0.0;
    const double delta = ir->inner_rtol > 0 && ir->inner_rtol < 1 ? ir->inner_rtol : 1e-2;
    const uint64_t max_seg = ir->update_interval > 0 ? (uint64_t)ir->update_interval : cfg->max_iter;
    uint64_t total = 0, updates = 0;
    bool first = true;
    double best = rel;
    int since_best = 0;
    while (bnorm > 0 && !(rel < cfg->rtol) && total < cfg->max_iter) {
        convert_f64_to_f32(s, r32, r64, n, 1.0);
        fill<float>(s, x32, 0.f, n);
        femgpu_solve_result seg{};
        const uint64_t len = std::min<uint64_t>(max_seg, cfg->max_iter - total);
        K32.solve_cont(first, cp, x32, bnorm, std::max(0.5 * cfg->rtol, delta * rel), len, &seg);
        total += K32.iterations_done();
        first = K32.last_status == ST_BREAKDOWN;
        convert_f32_to_f64_axpy(s, xd, x32, n, 1.0);
        rel = residual() / bnorm;
        ++updates;
        if (!std::isfinite(rel)) break;
        if (rel < best) { best = rel; since_best = 0; }
        else if (++since_best >= 8) break;  // no progress over 8 replacements
    }
    st.finish();
    if (inner_iterations) *inner_iterations = total;
    res->iterations = updates;
    res->final_residual = rel;
    res->final_residual_is_relative = 1;
    res->converged = bnorm == 0 || rel < cfg->rtol;
    res->error_estimate = 0.0;
    if (!res->converged)
        return fail(FEMGPU_ERR_NOT_CONVERGED,
                    "mixed-precision refinement stopped at relative residual " + std::to_string(rel));
    return 0;
}

extern "C" {

int femgpu_solve_cg_operator(femgpu_op *op, const void *b, void *x, femgpu_mem mem,
                             const femgpu_solver_config *cfg, femgpu_solve_result *res) {
    return guard([&] {
        FG_REQUIRE(op && b && x && res, FEMGPU_ERR_INVALID_ARG, "solve_cg_operator: NULL argument");
        check_cfg(cfg);
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        Staged st(op, b, x, mem, 8, 9);
        const int rc = dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            k.level = print_level(cfg);
            k.label = "cg_operator";
            return k.solve(MODE_CG, PREC_NONE, ChebParams{}, (const T *)st.b, (T *)st.x, cfg->rtol,
                           cfg->atol, cfg->max_iter, res);
        });
        st.finish();
        if (rc == FEMGPU_ERR_BREAKDOWN) return fail(rc, "CG operator breakdown: p^T A p is near zero");
        if (rc == FEMGPU_ERR_NOT_CONVERGED)
            return fail(rc, "solver did not converge in " + std::to_string(cfg->max_iter) +
                                " iterations (residual = " + std::to_string(res->final_residual) + ")");
        return rc;
    });
}

int femgpu_solve_pcg_jacobi(femgpu_op *op, const void *b, void *x, femgpu_mem mem,
                            const femgpu_solver_config *cfg, femgpu_solve_result *res) {
    return guard([&] {
        FG_REQUIRE(op && b && x && res, FEMGPU_ERR_INVALID_ARG, "solve_pcg_jacobi: NULL argument");
        check_cfg(cfg);
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        Staged st(op, b, x, mem, 8, 9);
        const int rc = dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            k.level = print_level(cfg);
            k.label = "pcg_jacobi";
            return k.solve(MODE_PCG, PREC_JACOBI, ChebParams{}, (const T *)st.b, (T *)st.x, cfg->rtol,
                           cfg->atol, cfg->max_iter, res);
        });
        st.finish();
        if (rc == FEMGPU_ERR_NOT_CONVERGED)
            return fail(rc, "solver did not converge in " + std::to_string(cfg->max_iter) +
                                " iterations (residual = " + std::to_string(res->final_residual) + ")");
        return rc;
    });
}

int femgpu_solve_pcg_jacobi_noalloc(femgpu_op *op, const void *b, void *x, const femgpu_solver_config *cfg,
                                    femgpu_solve_result *res) {
    return guard([&] {
        FG_REQUIRE(op && b && x && res, FEMGPU_ERR_INVALID_ARG, "solve_pcg_jacobi_noalloc: NULL argument");
        check_cfg(cfg);
        CsrView cv;
        FG_REQUIRE(op->csr_arrays(&cv), FEMGPU_ERR_INVALID_ARG,
                   "solve_pcg_jacobi_noalloc: needs an assembled CSR operator (its apply uses no workspace)");
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        void *ws = op->ctx->noalloc_scratch;
        FG_REQUIRE(ws, FEMGPU_ERR_HIP, "solve_pcg_jacobi_noalloc: the context has no scratch address");
        // W_COUNT work vectors, each starting on a 256-byte boundary
        const size_t stride = (op->n + 63) / 64 * 64;
        const size_t need = (size_t)KrylovSlots * stride * op->esize();
        FG_REQUIRE(need <= (size_t)FEMGPU_NOALLOC_SCRATCH_BYTES, FEMGPU_ERR_INVALID_ARG,
                   "solve_pcg_jacobi_noalloc: n = " + std::to_string(op->n) + " needs " + std::to_string(need) +
                       " bytes of workspace, more than the " + std::to_string((size_t)FEMGPU_NOALLOC_SCRATCH_BYTES) +
                       "-byte scratch array (at most " + std::to_string(noalloc_max_n(op->dtype)) +
                       " unknowns in this precision)");
        const int rc = dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            k.ext = (T *)ws;
            k.ext_stride = stride;
            return k.solve(MODE_PCG, PREC_JACOBI, ChebParams{}, (const T *)b, (T *)x, cfg->rtol, cfg->atol,
                           cfg->max_iter, res);
        });
        if (rc == FEMGPU_ERR_NOT_CONVERGED)
            return fail(rc, "solver did not converge in " + std::to_string(cfg->max_iter) +
                                " iterations (residual = " + std::to_string(res->final_residual) + ")");
        return rc;
    });
}

int femgpu_solve_amg_cg(femgpu_amg *amg, const void *b, void *x, femgpu_mem mem,
                        const femgpu_solver_config *cfg, femgpu_solve_result *res) {
    return guard([&] {
        FG_REQUIRE(amg && b && x && res, FEMGPU_ERR_INVALID_ARG, "solve_amg_cg: NULL argument");
        check_cfg(cfg);
        femgpu_op *op = amg_operator(amg);
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        Staged st(op, b, x, mem, 8, 9);
        Krylov<double> k(op);
        k.pc = [&](const double *r, double *z) { amg_apply(amg, r, z, k.flags); };
        k.level = print_level(cfg);
        k.label = "amg_pcg";
        const int rc = k.solve(MODE_PCG, PREC_CB, ChebParams{}, (const double *)st.b, (double *)st.x, cfg->rtol,
                               cfg->atol, cfg->max_iter, res);
        st.finish();
        if (rc == FEMGPU_ERR_NOT_CONVERGED)
            return fail(rc, "solver did not converge in " + std::to_string(cfg->max_iter) +
                                " iterations (residual = " + std::to_string(res->final_residual) + ")");
        return rc;
    });
}

int femgpu_solve_pcg_chebyshev(femgpu_op *op, const femgpu_cheb_config *cheb, const void *b,
                               void *x, femgpu_mem mem, const femgpu_solver_config *cfg,
                               femgpu_solve_result *res, double *lmax_used) {
    return guard([&] {
        FG_REQUIRE(op && b && x && res, FEMGPU_ERR_INVALID_ARG, "solve_pcg_chebyshev: NULL argument");
        check_cfg(cfg);
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        Staged st(op, b, x, mem, 10, 11);
        const int rc = dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            ChebParams cp = k.prepare_cheb(cheb);
            if (lmax_used) *lmax_used = cp.lmax;
            k.level = print_level(cfg);
            k.label = "pcg_chebyshev";
            return k.solve(MODE_PCG, PREC_CHEB, cp, (const T *)st.b, (T *)st.x, cfg->rtol, cfg->atol,
                           cfg->max_iter, res);
        });
        st.finish();
        if (rc == FEMGPU_ERR_NOT_CONVERGED)
            return fail(rc, "solver did not converge in " + std::to_string(cfg->max_iter) +
                                " iterations (residual = " + std::to_string(res->final_residual) + ")");
        return rc;
    });
}

int femgpu_chebyshev_smooth(femgpu_op *op, void *x, const void *b, double lmin, double lmax,
                            int n_iter, femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(op && b && x, FEMGPU_ERR_INVALID_ARG, "chebyshev_smooth: NULL argument");
        FG_REQUIRE(n_iter >= 0, FEMGPU_ERR_INVALID_ARG, "chebyshev_smooth: n_iter < 0");
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        Staged st(op, b, x, mem, 8, 9);
        dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            const size_t n = k.n;
            T *dg = k.diag(), *dinv = k.vec(k.W_INV), *r = k.vec(k.W_R), *d = k.vec(k.W_D);
            hipLaunchKernelGGL(k_inv_tiny<T>, dim3(vgrid(n)), dim3(256), 0, k.s, n, (const T *)dg, dinv, 1e-14);
            fill<T>(k.s, d, T(0), n);
            const double alpha = (lmax + lmin) / 2.0, delta = (lmax - lmin) / 2.0;
            const double two_over_alpha = 2.0 / alpha;
            const double q = std::pow(delta / (2.0 * alpha), 2);
            double rho_prev = 0.0;
            for (int it = 0; it < n_iter; ++it) {
                k.apply((const T *)st.x, r, false);
                double rho = 1.0;
                if (it == 0) rho_prev = 1.0;
                else { rho = 1.0 / (1.0 - q * rho_prev); rho_prev = rho; }
                hipLaunchKernelGGL(k_cheb_ref<T>, dim3(vgrid(n)), dim3(256), 0, k.s, n, (const T *)st.b, r,
                                   (const T *)dinv, d, (T *)st.x, two_over_alpha, rho, it == 0 ? 1 : 0);
            }
            FG_HIP(hipGetLastError());
            return 0;
        });
        st.finish();
        return 0;
    });
}

int femgpu_estimate_spectral_radius(femgpu_op *op, int n_iter, double *rho_out) {
    return guard([&] {
        FG_REQUIRE(op && rho_out, FEMGPU_ERR_INVALID_ARG, "estimate_spectral_radius: NULL argument");
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            const size_t n = k.n;
            T *v = k.vec(k.W_R), *w = k.vec(k.W_Z);
            fill<T>(k.s, v, (T)(1.0 / std::sqrt((double)n)), n);
            const unsigned nb = nblk(n);
            double rho = 1.0;
            for (int it = 0; it < n_iter; ++it) {
                k.apply(v, w, false);
                hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, k.s, n, (const T *)w, (const T *)w,
                                   k.part, 0, nullptr);
                hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(1024), 0, k.s, (const double *)k.part, (int)nb,
                                   k.S, (int)S_RR0);
                hipLaunchKernelGGL(k_scale_to<T>, dim3(vgrid(n)), dim3(256), 0, k.s, n, (const T *)w, v,
                                   (const double *)k.S, (int)S_RR0);
            }
            if (n_iter > 0) {
                k.read_scalars();
                rho = std::sqrt(op->ctx->pinned[S_RR0]);
            }
            *rho_out = rho;
            return 0;
        });
        return 0;
    });
}

int femgpu_solve_ir_mixed(femgpu_op *op64, femgpu_op *op32, const femgpu_ir_config *ir,
                          const void *b, void *x, femgpu_mem mem, const femgpu_solver_config *cfg,
                          femgpu_solve_result *res, uint64_t *inner_iterations) {
    return guard([&] {
        if (ir && ir->mode == 1) {
            FG_REQUIRE(op64 && op32 && b && x && res, FEMGPU_ERR_INVALID_ARG, "solve_ir_mixed: NULL argument");
            FG_REQUIRE(op64->dtype == FEMGPU_F64 && op32->dtype == FEMGPU_F32, FEMGPU_ERR_INVALID_ARG,
                       "solve_ir_mixed: need an f64 and an f32 operator");
            FG_REQUIRE(op64->n == op32->n && op64->ctx == op32->ctx, FEMGPU_ERR_DIM_MISMATCH,
                       "solve_ir_mixed: operators differ in size or context");
            check_cfg(cfg);
            return ir_continued(op64, op32, ir, b, x, mem, cfg, res, inner_iterations);
        }
        return ir_refine(op64, op32, ir, b, x, mem, cfg, res, inner_iterations,
                         [&](Krylov<float> &K32, const ChebParams &cp, const float *r32, float *e32,
                             double t, uint64_t cap, femgpu_solve_result *ir_res) {
                             const int rc = K32.solve(MODE_PCG, PREC_CHEB, cp, r32, e32, t, 0.0, cap, ir_res);
                             // a breakdown reports max_iter; count what ran
                             ir_res->iterations = K32.iterations_done();
                             return rc;
                         });
    });
}

}  // extern "C"

// ------------------------------------------------- learned preconditioner --
namespace femgpu {

// Flexible CG (Notay, SIAM J. Sci. Comput. 22 (2000)) truncated to the last
// `trunc` directions (A-orthogonalisation window), with a two-level
// preconditioner: Chebyshev-Jacobi smoothing z1 = C r, then the network's
// correction of the smooth residual r1 = r - A z1 that the smoother leaves,
// scaled by its Galerkin step (the network's output scale does not matter).
constexpr int kFcgTruncMax = 64;

// ---- fused flexible-CG vector kernels (partials on the fixed red_blocks grid) ----
// The solver's scalars live in a small device state (FcgState slots); the
// one-block kernels k_fcgd_* finalise the partials into it, and the vector
// kernels read gamma, beta and alpha from it, so an iteration reads back one
// status record instead of every dot product.
enum FcgState {
    F_GAMMA = 0, F_ALPHA, F_RR, F_BEST, F_SINCE, F_STATUS, F_NET, F_RC2,
    F_BETA = 8,                  // kFcgTruncMax coefficients
    F_PQ = F_BETA + kFcgTruncMax,  // p_j.q_j by history slot
    F_COUNT = F_PQ + kFcgTruncMax
};
enum FcgStatus { FS_RUN = 0, FS_CONV = 1, FS_NONFINITE = 2, FS_PQ = 3, FS_STAG = 4 };

constexpr int kFcgPack = 16;  // directions per k_fcg_pq pass
template <class T> struct FcgDirs {
    const T *p[kFcgPack];
    const T *q[kFcgPack];
    int m, j0;  // directions in this pass, index of the first one's coefficient
};
struct FcgSlots {
    int s[kPack];  // history slots of the pack's directions
    int m, j0;
};

// The flexible-CG vector kernels stream up to 2 + 2 kFcgPack vectors.  With
// every pointer 16-byte aligned (W4) each thread takes 4 consecutive entries
// per grid-stride step, loaded as 16-byte vectors (twice as many bytes in
// flight per wave as 4-byte loads); the n mod 4 tail goes to thread 0 of
// block 0.  Either way every
// thread sums its entries in a fixed order, so the partials are deterministic.
template <class T> struct Q4 { T v[4]; };
template <class T> __device__ __forceinline__ Q4<T> ld4(const T *p, size_t c) {
    Q4<T> o;
    if constexpr (sizeof(T) == 4) {
        const float4 f = reinterpret_cast<const float4 *>(p)[c];
        o.v[0] = f.x; o.v[1] = f.y; o.v[2] = f.z; o.v[3] = f.w;
    } else {
        const double2 a = reinterpret_cast<const double2 *>(p)[2 * c];
        const double2 b = reinterpret_cast<const double2 *>(p)[2 * c + 1];
        o.v[0] = a.x; o.v[1] = a.y; o.v[2] = b.x; o.v[3] = b.y;
    }
    return o;
}
template <class T> __device__ __forceinline__ void st4(T *p, size_t c, const Q4<T> &o) {
    if constexpr (sizeof(T) == 4) {
        reinterpret_cast<float4 *>(p)[c] = make_float4(o.v[0], o.v[1], o.v[2], o.v[3]);
    } else {
        reinterpret_cast<double2 *>(p)[2 * c] = make_double2(o.v[0], o.v[1]);
        reinterpret_cast<double2 *>(p)[2 * c + 1] = make_double2(o.v[2], o.v[3]);
    }
}
// Runs body(i) over [0, n): W4 in chunks of 4 (vector body4(c) for chunk c),
// else one entry per step.
template <bool W4, class F1, class F4>
__device__ __forceinline__ void fcg_stride(size_t n, F1 &&body, F4 &&body4) {
    const size_t t0 = blockIdx.x * (size_t)256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    if constexpr (W4) {
        const size_t n4 = n / 4;
        for (size_t c = t0; c < n4; c += stride) body4(c);
        if (t0 == 0)
            for (size_t i = 4 * n4; i < n; ++i) body(i);
    } else {
        for (size_t i = t0; i < n; i += stride) body(i);
    }
}

// r1 = r - az1; partials of r1.r1
template <class T, bool W4>
__global__ __launch_bounds__(256) void k_fcg_resid(size_t n, const T *__restrict__ r, const T *__restrict__ az1,
                                                   T *__restrict__ r1, double *__restrict__ part) {
    double v[1] = {0.0};
    auto one = [&](size_t i) {
        const T d = r[i] - az1[i];
        r1[i] = d;
        v[0] += (double)d * (double)d;
    };
    fcg_stride<W4>(n, one, [&](size_t c) {
        const Q4<T> a = ld4(r, c), b = ld4(az1, c);
        Q4<T> d;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            d.v[e] = a.v[e] - b.v[e];
            v[0] += (double)d.v[e] * (double)d.v[e];
        }
        st4(r1, c, d);
    });
    block_sum<1, 256>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// partials of nn.An (row 0) and nn.r1 (row 1)
template <class T, bool W4>
__global__ __launch_bounds__(256) void k_fcg_dots2(size_t n, const T *__restrict__ nn, const T *__restrict__ An,
                                                   const T *__restrict__ r1, double *__restrict__ part) {
    double v[2] = {0.0, 0.0};
    auto one = [&](size_t i) {
        const double a = (double)nn[i];
        v[0] += a * (double)An[i];
        v[1] += a * (double)r1[i];
    };
    fcg_stride<W4>(n, one, [&](size_t c) {
        const Q4<T> a = ld4(nn, c), b = ld4(An, c), d = ld4(r1, c);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v[0] += (double)a.v[e] * (double)b.v[e];
            v[1] += (double)a.v[e] * (double)d.v[e];
        }
    });
    block_sum<2, 256>(v);
    if (threadIdx.x == 0) { part[blockIdx.x] = v[0]; part[kRedBlocks + blockIdx.x] = v[1]; }
}

// partials of (z1 + g nn).q_j, j < m <= kPack (row j), g = st[F_GAMMA]; nn
// unused when g == 0
template <class T, bool W4>
__global__ __launch_bounds__(256) void k_fcg_zq(size_t n, const T *__restrict__ z1, const T *__restrict__ nn,
                                                const double *__restrict__ st, VecPack<T> qs,
                                                double *__restrict__ part) {
    const double g = st[F_GAMMA];
    double v[kPack];
#pragma unroll
    for (int j = 0; j < kPack; ++j) v[j] = 0.0;
    auto one = [&](size_t i) {
        const double z = g != 0.0 ? (double)(T)((double)z1[i] + g * (double)nn[i]) : (double)z1[i];
#pragma unroll
        for (int j = 0; j < kPack; ++j)
            if (j < qs.m) v[j] += z * (double)qs.v[j][i];
    };
    fcg_stride<W4>(n, one, [&](size_t c) {
        const Q4<T> a = ld4(z1, c);
        double z[4];
        if (g != 0.0) {
            const Q4<T> b = ld4(nn, c);
#pragma unroll
            for (int e = 0; e < 4; ++e) z[e] = (double)(T)((double)a.v[e] + g * (double)b.v[e]);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) z[e] = (double)a.v[e];
        }
#pragma unroll
        for (int j = 0; j < kPack; ++j)
            if (j < qs.m) {
                const Q4<T> q = ld4(qs.v[j], c);
#pragma unroll
                for (int e = 0; e < 4; ++e) v[j] += z[e] * (double)q.v[e];
            }
    });
    block_sum<kPack, 256>(v);
    if (threadIdx.x == 0)
        for (int j = 0; j < qs.m; ++j) part[(size_t)j * kRedBlocks + blockIdx.x] = v[j];
}

// one entry of k_fcg_pq
template <class T>
__device__ __forceinline__ void fcg_pq_one(size_t i, const T *__restrict__ z1, const T *__restrict__ nn,
                                           const T *__restrict__ az1, const T *__restrict__ An, double g,
                                           const T *__restrict__ r, const FcgDirs<T> &d, const double (&b)[kFcgPack],
                                           T *p, T *q, int first, int last, double (&v)[2]) {
    {
        double sp, sq;
        if (first) {
            // the preconditioned residual z and A z, rounded to T as a stored
            // z would be
            sp = g != 0.0 ? (double)(T)((double)z1[i] + g * (double)nn[i]) : (double)z1[i];
            sq = g != 0.0 ? (double)(T)((double)az1[i] + g * (double)An[i]) : (double)az1[i];
        } else {
            sp = (double)p[i];
            sq = (double)q[i];
        }
#pragma unroll
        for (int j = 0; j < kFcgPack; ++j)
            if (j < d.m) {
                sp += b[j] * (double)d.p[j][i];
                sq += b[j] * (double)d.q[j][i];
            }
        const T pt = (T)sp, qt = (T)sq;
        p[i] = pt;
        q[i] = qt;
        if (last) {
            v[0] += (double)pt * (double)qt;
            v[1] += (double)pt * (double)r[i];
        }
    }
}

// p = base_p + sum_j beta_j p_j, q = base_q + sum_j beta_j q_j with base
// (z1 + g nn, az1 + g An) on the first pass and (p, q) after; on the last pass
// partials of p.q (row 0) and p.r (row 1)
template <class T, bool W4>
__global__ __launch_bounds__(256) void k_fcg_pq(size_t n, const T *__restrict__ z1, const T *__restrict__ nn,
                                                const T *__restrict__ az1, const T *__restrict__ An,
                                                const double *__restrict__ st, const T *__restrict__ r,
                                                FcgDirs<T> d, T *p, T *q, int first, int last,
                                                double *__restrict__ part) {
    const double g = st[F_GAMMA];
    double b[kFcgPack];
#pragma unroll
    for (int j = 0; j < kFcgPack; ++j) b[j] = j < d.m ? st[F_BETA + d.j0 + j] : 0.0;
    double v[2] = {0.0, 0.0};
    fcg_stride<W4>(n, [&](size_t i) { fcg_pq_one(i, z1, nn, az1, An, g, r, d, b, p, q, first, last, v); },
                   [&](size_t c) {
        double sp[4], sq[4];
        if (first) {
            const Q4<T> a = ld4(z1, c), aa = ld4(az1, c);
            if (g != 0.0) {
                const Q4<T> bn = ld4(nn, c), bA = ld4(An, c);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    sp[e] = (double)(T)((double)a.v[e] + g * (double)bn.v[e]);
                    sq[e] = (double)(T)((double)aa.v[e] + g * (double)bA.v[e]);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) { sp[e] = (double)a.v[e]; sq[e] = (double)aa.v[e]; }
            }
        } else {
            const Q4<T> a = ld4((const T *)p, c), aa = ld4((const T *)q, c);
#pragma unroll
            for (int e = 0; e < 4; ++e) { sp[e] = (double)a.v[e]; sq[e] = (double)aa.v[e]; }
        }
#pragma unroll
        for (int j = 0; j < kFcgPack; ++j)
            if (j < d.m) {
                const Q4<T> pj = ld4(d.p[j], c), qj = ld4(d.q[j], c);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    sp[e] += b[j] * (double)pj.v[e];
                    sq[e] += b[j] * (double)qj.v[e];
                }
            }
        Q4<T> pt, qt;
#pragma unroll
        for (int e = 0; e < 4; ++e) { pt.v[e] = (T)sp[e]; qt.v[e] = (T)sq[e]; }
        st4(p, c, pt);
        st4(q, c, qt);
        if (last) {
            const Q4<T> rv = ld4(r, c);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[0] += (double)pt.v[e] * (double)qt.v[e];
                v[1] += (double)pt.v[e] * (double)rv.v[e];
            }
        }
    });
    if (last) {
        block_sum<2, 256>(v);
        if (threadIdx.x == 0) { part[blockIdx.x] = v[0]; part[kRedBlocks + blockIdx.x] = v[1]; }
    }
}

// x += alpha p, r -= alpha q (alpha = st[F_ALPHA]); partials of r.r.  Nothing
// when the iteration has already stopped (p.q <= 0 or not finite).
template <class T, bool W4>
__global__ __launch_bounds__(256) void k_fcg_xr(size_t n, const double *__restrict__ st, const T *__restrict__ p,
                                                const T *__restrict__ q, T *__restrict__ x, T *__restrict__ r,
                                                double *__restrict__ part) {
    if (st[F_STATUS] != FS_RUN) return;
    const double a = st[F_ALPHA];
    double v[1] = {0.0};
    auto one = [&](size_t i) {
        x[i] = (T)((double)x[i] + a * (double)p[i]);
        const T ri = (T)((double)r[i] - a * (double)q[i]);
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
    };
    fcg_stride<W4>(n, one, [&](size_t c) {
        Q4<T> xv = ld4((const T *)x, c), rv = ld4((const T *)r, c);
        const Q4<T> pv = ld4(p, c), qv = ld4(q, c);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            xv.v[e] = (T)((double)xv.v[e] + a * (double)pv.v[e]);
            rv.v[e] = (T)((double)rv.v[e] - a * (double)qv.v[e]);
            v[0] += (double)rv.v[e] * (double)rv.v[e];
        }
        st4(x, c, xv);
        st4(r, c, rv);
    });
    block_sum<1, 256>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// out[j] = sum of partial row j (fixed order), j < m
__global__ __launch_bounds__(1024) void k_fcg_fin(const double *__restrict__ partials, int nb, int m,
                                                  double *__restrict__ out) {
    for (int j = 0; j < m; ++j) {
        double v[1] = {threadIdx.x < (unsigned)nb ? partials[(size_t)j * kRedBlocks + threadIdx.x] : 0.0};
        block_sum<1, 1024>(v);
        if (threadIdx.x == 0) out[j] = v[0];
        __syncthreads();
    }
}

// the m sums of the partial rows, copied to the host (one synchronisation)
static void fcg_fin_host(femgpu_ctx *c, int nb, int m, double *out) {
    double *o = c->scalars.as<double>() + kMultiDotBase;
    hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(1024), 0, c->stream, (const double *)c->partials.as<double>(), nb,
                       m, o);
    FG_HIP(hipMemcpyAsync(c->pinned + kMultiDotBase, o, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->pinned + kMultiDotBase, sizeof(double) * m);
}

// ---- one-block finalisers into the state (fixed summation order) ----
__global__ __launch_bounds__(1024) void k_fcgd_init(double *st, double best) {
    for (int i = threadIdx.x; i < F_COUNT; i += 1024) st[i] = i >= F_PQ ? 1.0 : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) st[F_BEST] = best;
}
// after the true-residual check of a recurrence that claimed convergence
__global__ void k_fcgd_resume(double *st, double rr, double best, double since) {
    st[F_RR] = rr; st[F_BEST] = best; st[F_SINCE] = since; st[F_STATUS] = FS_RUN;
}
__global__ void k_fcgd_no_net(double *st) { st[F_GAMMA] = 0.0; }
// gamma = n.r1 / n.An, or 0 when that is unusable (or the network's input
// scale |R r1|^2 in st[F_RC2] is, with use_rc2)
__global__ __launch_bounds__(1024) void k_fcgd_gamma(const double *part, int nb, double *st, int use_rc2) {
    __shared__ double o[2];
    fin_sum(part, nb, o, 2);
    if (threadIdx.x) return;
    const double nAn = o[0], nr = o[1], rc2 = st[F_RC2];
    const bool ok = nAn > 0.0 && isfinite(nAn) && isfinite(nr) && (!use_rc2 || (rc2 > 0.0 && isfinite(rc2)));
    st[F_GAMMA] = ok ? nr / nAn : 0.0;
    if (ok) st[F_NET] += 1.0;
}
// beta[j0 + j] = -(z.q_j) / (p_j.q_j), j < m
__global__ __launch_bounds__(1024) void k_fcgd_beta(const double *part, int nb, FcgSlots sl, double *st) {
    __shared__ double o[kPack];
    fin_sum(part, nb, o, sl.m);
    if (threadIdx.x) return;
    for (int j = 0; j < sl.m; ++j) st[F_BETA + sl.j0 + j] = -o[j] / st[F_PQ + sl.s[j]];
}
// alpha = p.r / p.q and p.q into the history slot; the stop codes of a
// non-finite or non-positive p.q
__global__ __launch_bounds__(1024) void k_fcgd_alpha(const double *part, int nb, int slot, double *st) {
    __shared__ double o[2];
    fin_sum(part, nb, o, 2);
    if (threadIdx.x) return;
    const double pq = o[0], pr = o[1];
    if (!isfinite(pq)) { st[F_STATUS] = FS_NONFINITE; return; }
    if (!(pq > 0.0)) { st[F_STATUS] = FS_PQ; return; }
    st[F_PQ + slot] = pq;
    st[F_ALPHA] = pr / pq;
}
// |r|, the stopping test on the recurrence residual and the stagnation count
__global__ __launch_bounds__(1024) void k_fcgd_conv(const double *part, int nb, double *st, double bnorm,
                                                    double rtol, double atol, int patience) {
    __shared__ double o[1];
    fin_sum(part, nb, o, 1);
    if (threadIdx.x) return;
    if (st[F_STATUS] != FS_RUN) return;
    const double rr = sqrt(o[0]);
    st[F_RR] = rr;
    if (!isfinite(rr)) { st[F_STATUS] = FS_NONFINITE; return; }
    if (rr / bnorm < rtol || rr < atol) { st[F_STATUS] = FS_CONV; return; }
    if (rr < st[F_BEST]) {
        st[F_BEST] = rr;
        st[F_SINCE] = 0.0;
    } else {
        st[F_SINCE] += 1.0;
        if (st[F_SINCE] >= patience) st[F_STATUS] = FS_STAG;
    }
}

// Flexible CG (Notay, SIAM J. Sci. Comput. 22 (2000)) truncated to the last
// `trunc` directions (A-orthogonalisation window), with a two-level
// preconditioner: Chebyshev-Jacobi smoothing z1 = C r, then the network's
// correction of the smooth residual r1 = r - A z1 that the smoother leaves,
// scaled by its Galerkin step (the network's output scale does not matter).
// z = z1 + gamma n and A z = A z1 + gamma A n are never stored: the fused
// kernels above read their parts.  One status read-back per iteration (two
// with a network on the full DOF grid, which needs |r1| on the host).
template <class T>
static int fcg_learned(Krylov<T> &K, femgpu_unet *net, const ChebParams &cp, const T *b, T *x,
                       double rtol, double atol, uint64_t max_iter, int patience, int trunc,
                       femgpu_solve_result *res, femgpu_fcg_info *info) {
    femgpu_ctx *c = K.c;
    hipStream_t s = K.s;
    const size_t n = K.n;
    const femgpu_dtype dt = sizeof(T) == 8 ? FEMGPU_F64 : FEMGPU_F32;
    T *r = K.vec(K.W_R), *tmp = K.vec(K.W_AP);
    T *nn = K.vec(12), *An = K.vec(13), *r1 = K.vec(14), *z1 = K.vec(15), *az1 = K.vec(16);
    std::vector<T *> P(trunc), Q(trunc);
    for (int j = 0; j < trunc; ++j) { P[j] = K.vec(20 + 2 * j); Q[j] = K.vec(21 + 2 * j); }
    const T *inv = K.vec(K.W_INV);
    const int nb = red_blocks(n);
    const dim3 grid((unsigned)nb), block(256);
    double *part = c->partials.as<double>();
    const bool coarse = unet_coarse(net);
    // the vector kernels' 16-byte path needs every vector they touch aligned
    // (the workspace always is; x is the caller's)
    auto al16 = [](const void *q) { return ((uintptr_t)q & 15u) == 0; };
    bool w4 = al16(x) && al16(r) && al16(nn) && al16(An) && al16(r1) && al16(z1) && al16(az1);
    for (int j = 0; j < trunc; ++j) w4 = w4 && al16(P[j]) && al16(Q[j]);
    auto kf_resid = w4 ? k_fcg_resid<T, true> : k_fcg_resid<T, false>;
    auto kf_dots2 = w4 ? k_fcg_dots2<T, true> : k_fcg_dots2<T, false>;
    auto kf_zq = w4 ? k_fcg_zq<T, true> : k_fcg_zq<T, false>;
    auto kf_pq = w4 ? k_fcg_pq<T, true> : k_fcg_pq<T, false>;
    auto kf_xr = w4 ? k_fcg_xr<T, true> : k_fcg_xr<T, false>;
    DevBuf stb(sizeof(double) * F_COUNT);
    double *st = stb.as<double>();
    double *hst = c->pinned + kMultiDotBase;  // read-back of the first 8 state slots
    auto read_state = [&]() {
        FG_HIP(hipMemcpyAsync(hst, st, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
    };
    auto dot = [&](const T *a, const T *bb) { return dot_host(c, a, bb, n, dt); };
    FG_HIP(hipMemsetAsync(K.flags, 0, sizeof(int), s));
    // Two-level preconditioner, z = B r, together with A z:
    //   z1 = C r              (Chebyshev-Jacobi smoothing)
    //   r1 = r - A z1         (the smooth residual the smoother leaves)
    //   n  = N(r1), gamma = n.r1 / n.An   (network correction, Galerkin-scaled)
    //   z  = z1 + gamma n,    A z = A z1 + gamma A n
    // Without a usable network output gamma = 0, z = C r.
    auto precond = [&](const T *rin) {
        K.cheb_apply(cp, rin, inv, z1);
        K.apply(z1, az1, false);
        hipLaunchKernelGGL(kf_resid, grid, block, 0, s, n, rin, (const T *)az1, r1, part);
        if (coarse) {
            unet_apply_coarse_dev(net, r1, nn, dt, st + F_RC2);  // the scale stays on the device
        } else {
            double rn;
            fcg_fin_host(c, nb, 1, &rn);
            if (!unet_apply_dev(net, r1, nn, dt, std::sqrt(rn))) {
                hipLaunchKernelGGL(k_fcgd_no_net, dim3(1), dim3(1), 0, s, st);
                return;
            }
        }
        K.apply(nn, An, false);
        hipLaunchKernelGGL(kf_dots2, grid, block, 0, s, n, (const T *)nn, (const T *)An, (const T *)r1, part);
        hipLaunchKernelGGL(k_fcgd_gamma, dim3(1), dim3(1024), 0, s, (const double *)part, nb, st, coarse ? 1 : 0);
    };
    info->fcg_iterations = 0; info->fallback_iterations = 0; info->fell_back = 0; info->fallback_reason = 0;
    info->net_corrections = 0;
    auto true_residual = [&]() {
        K.apply(x, tmp, false);
        copy<T>(s, r, b, n);
        axpby<T>(s, r, -1.0, tmp, 1.0, n);              // r = b - A x
        return std::sqrt(dot(r, r));
    };
    double rr = true_residual();
    const double bnorm = std::sqrt(dot(b, b));
    res->final_residual_is_relative = 1;
    res->error_estimate = 0.0;
    if (bnorm < 1e-30) {
        res->converged = 1; res->iterations = 0; res->final_residual = 0.0;
        return FEMGPU_OK;
    }
    hipLaunchKernelGGL(k_fcgd_init, dim3(1), dim3(1024), 0, s, st, rr);
    int reason = 0, replacements = 0;
    auto finish_info = [&]() {
        read_state();
        info->net_corrections = (uint64_t)hst[F_NET];
    };
    precond(r);
    uint64_t it = 0, stored = 0;  // stored: directions kept so far (<= trunc)
    for (; it < max_iter; ++it) {
        // p = z - sum_j (z.q_j / p_j.q_j) p_j and q = A p = A z - sum_j (...) q_j
        // over the kept directions (A-orthogonalisation, Notay's FCG).
        const int slot = (int)(it % trunc);
        T *p = P[slot], *qk = Q[slot];
        const int m = (int)std::min<uint64_t>(stored, (uint64_t)trunc - 1);
        std::vector<int> sj(m);
        for (int j = 0; j < m; ++j) sj[j] = (int)((it - 1 - j) % trunc);
        for (int j0 = 0; j0 < m; j0 += kPack) {  // z.q_j, kPack directions per pass
            VecPack<T> pk{};
            FcgSlots sl{};
            pk.m = sl.m = std::min(kPack, m - j0);
            sl.j0 = j0;
            for (int j = 0; j < pk.m; ++j) { pk.v[j] = Q[sj[j0 + j]]; sl.s[j] = sj[j0 + j]; }
            hipLaunchKernelGGL(kf_zq, grid, block, 0, s, n, (const T *)z1, (const T *)nn, (const double *)st,
                               pk, part);
            hipLaunchKernelGGL(k_fcgd_beta, dim3(1), dim3(1024), 0, s, (const double *)part, nb, sl, st);
        }
        const int passes = std::max(1, (m + kFcgPack - 1) / kFcgPack);
        for (int ps = 0; ps < passes; ++ps) {  // p, q and p.q, p.r, kFcgPack directions per pass
            FcgDirs<T> dd{};
            dd.j0 = ps * kFcgPack;
            dd.m = std::max(0, std::min(kFcgPack, m - dd.j0));
            for (int j = 0; j < dd.m; ++j) { dd.p[j] = P[sj[dd.j0 + j]]; dd.q[j] = Q[sj[dd.j0 + j]]; }
            hipLaunchKernelGGL(kf_pq, grid, block, 0, s, n, (const T *)z1, (const T *)nn, (const T *)az1,
                               (const T *)An, (const double *)st, (const T *)r, dd, p, qk, ps == 0 ? 1 : 0,
                               ps == passes - 1 ? 1 : 0, part);
        }
        hipLaunchKernelGGL(k_fcgd_alpha, dim3(1), dim3(1024), 0, s, (const double *)part, nb, slot, st);
        hipLaunchKernelGGL(kf_xr, grid, block, 0, s, n, (const double *)st, (const T *)p, (const T *)qk, x, r,
                           part);
        hipLaunchKernelGGL(k_fcgd_conv, dim3(1), dim3(1024), 0, s, (const double *)part, nb, st, bnorm, rtol, atol,
                           patience);
        read_state();
        const int status = (int)hst[F_STATUS];
        rr = hst[F_RR];
        if (status == FS_NONFINITE) { reason = 2; break; }
        if (status == FS_PQ) { reason = 3; break; }
        if (stored < (uint64_t)trunc) ++stored;
        if (status == FS_STAG) { reason = 1; ++it; break; }
        if (status == FS_CONV) {
            // The recurrence residual says converged: confirm with b - A x and,
            // if that is not yet below the tolerance, continue from it.
            rr = true_residual();
            if (rr / bnorm < rtol || rr < atol) {
                info->fcg_iterations = it + 1;
                finish_info();
                res->converged = 1; res->iterations = it + 1; res->final_residual = rr / bnorm;
                return FEMGPU_OK;
            }
            if (++replacements > 8) { reason = 1; ++it; break; }
            double best = hst[F_BEST], since = hst[F_SINCE];
            if (rr < best) { best = rr; since = 0.0; }
            else if ((since += 1.0) >= patience) { reason = 1; ++it; break; }
            hipLaunchKernelGGL(k_fcgd_resume, dim3(1), dim3(1), 0, s, st, rr, best, since);
        }
        precond(r);
    }
    info->fcg_iterations = it;
    finish_info();
    if (!reason) {
        res->converged = 0; res->iterations = it; res->final_residual = rr / bnorm;
        return FEMGPU_ERR_NOT_CONVERGED;
    }
    // Fallback: plain PCG + Chebyshev-Jacobi from the current iterate.  A
    // non-finite iterate is reset to zero first.
    info->fell_back = 1; info->fallback_reason = reason;
    if (reason == 2) fill<T>(s, x, T(0), n);
    const uint64_t left = max_iter > it ? max_iter - it : 1;
    femgpu_solve_result r2{};
    const int rc = K.solve(MODE_PCG, PREC_CHEB, cp, b, x, rtol, atol, left, &r2);
    info->fallback_iterations = r2.iterations;
    *res = r2;
    res->iterations = it + r2.iterations;
    return rc;
}

}  // namespace femgpu

extern "C" {

int femgpu_chebyshev_jacobi_apply(femgpu_op *op, const femgpu_cheb_config *cheb, const void *r,
                                  void *z, femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(op && r && z, FEMGPU_ERR_INVALID_ARG, "chebyshev_jacobi_apply: NULL argument");
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        const size_t bytes = op->n * op->esize();
        const void *rd = r;
        void *zd = z;
        if (mem == FEMGPU_HOST) {
            rd = op->wsvec(15, bytes); zd = op->wsvec(16, bytes);
            FG_HIP(hipMemcpyAsync((void *)rd, r, bytes, hipMemcpyHostToDevice, op->ctx->stream));
        }
        dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            ChebParams cp = k.prepare_cheb(cheb);
            FG_HIP(hipMemsetAsync(k.flags, 0, sizeof(int), k.s));
            k.cheb_apply(cp, (const T *)rd, (const T *)k.vec(k.W_INV), (T *)zd);
            return 0;
        });
        FG_HIP(hipGetLastError());
        if (mem == FEMGPU_HOST) FG_HIP(hipMemcpyAsync(z, zd, bytes, hipMemcpyDeviceToHost, op->ctx->stream));
        FG_HIP(hipStreamSynchronize(op->ctx->stream));
        return 0;
    });
}

int femgpu_solve_fcg_learned(femgpu_op *op, femgpu_unet *net, const femgpu_cheb_config *cheb,
                             const void *b, void *x, femgpu_mem mem,
                             const femgpu_solver_config *cfg, femgpu_solve_result *res,
                             femgpu_fcg_info *info) {
    return guard([&] {
        FG_REQUIRE(op && net && b && x && res && info, FEMGPU_ERR_INVALID_ARG, "solve_fcg_learned: NULL argument");
        check_cfg(cfg);
        FG_REQUIRE(unet_voxels(net) == (long long)op->n, FEMGPU_ERR_DIM_MISMATCH,
                   "solve_fcg_learned: network grid and operator size differ");
        FG_REQUIRE(unet_ctx(net) == op->ctx, FEMGPU_ERR_INVALID_ARG, "solve_fcg_learned: different contexts");
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        FG_HIP(hipSetDevice(op->ctx->device));
        Staged st(op, b, x, mem, 17, 18);
        const int patience = info->patience > 0 ? info->patience : 20;
        const int trunc = info->truncation > 0 ? (info->truncation < kFcgTruncMax ? info->truncation : kFcgTruncMax) : 8;
        const int rc = dispatch(op, [&](auto &k) {
            using T = std::remove_pointer_t<decltype(k.vec(0))>;
            ChebParams cp = k.prepare_cheb(cheb);
            return fcg_learned<T>(k, net, cp, (const T *)st.b, (T *)st.x, cfg->rtol, cfg->atol,
                                  cfg->max_iter, patience, trunc, res, info);
        });
        st.finish();
        if (rc == FEMGPU_ERR_NOT_CONVERGED)
            return fail(rc, "solver did not converge in " + std::to_string(cfg->max_iter) +
                                " iterations (residual = " + std::to_string(res->final_residual) + ")");
        return rc;
    });
}

int femgpu_solve_ir_learned(femgpu_op *op64, femgpu_op *op32, femgpu_unet *net,
                            const femgpu_ir_config *ir, const void *b, void *x, femgpu_mem mem,
                            const femgpu_solver_config *cfg, femgpu_solve_result *res,
                            uint64_t *inner_iterations, femgpu_fcg_info *info) {
    return guard([&] {
        FG_REQUIRE(net && info, FEMGPU_ERR_INVALID_ARG, "solve_ir_learned: NULL argument");
        FG_REQUIRE(op32 && unet_voxels(net) == (long long)op32->n, FEMGPU_ERR_DIM_MISMATCH,
                   "solve_ir_learned: network grid and operator size differ");
        FG_REQUIRE(unet_ctx(net) == op32->ctx, FEMGPU_ERR_INVALID_ARG, "solve_ir_learned: different contexts");
        const int patience = info->patience > 0 ? info->patience : 20;
        const int trunc = info->truncation > 0 ? std::min(info->truncation, kFcgTruncMax) : 8;
        femgpu_fcg_info tot = *info;
        tot.fcg_iterations = tot.fallback_iterations = tot.net_corrections = 0;
        tot.fell_back = 0; tot.fallback_reason = 0;
        const int rc = ir_refine(op64, op32, ir, b, x, mem, cfg, res, inner_iterations,
                                 [&](Krylov<float> &K32, const ChebParams &cp, const float *r32, float *e32,
                                     double t, uint64_t cap, femgpu_solve_result *ir_res) {
                                     femgpu_fcg_info one = *info;
                                     const int r = fcg_learned<float>(K32, net, cp, r32, e32, t, 0.0, cap,
                                                                      patience, trunc, ir_res, &one);
                                     tot.fcg_iterations += one.fcg_iterations;
                                     tot.fallback_iterations += one.fallback_iterations;
                                     tot.net_corrections += one.net_corrections;
                                     if (one.fell_back) { tot.fell_back = 1; tot.fallback_reason = one.fallback_reason; }
                                     return r;
                                 });
        *info = tot;
        return rc;
    });
}

}  // extern "C"
