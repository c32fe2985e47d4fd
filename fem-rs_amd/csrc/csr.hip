// CsrMatrix<T> on the device and its SpMV (linalg/src/csr.rs).
//
// Layout is the reference's own: row_ptr u64[n+1], col_idx u32[nnz],
// values T[nnz].  Two kernels:
//  - row-per-thread, used for short rows (the assembled P1/P2 matrices): the
//    f64 row sum runs in blocks of four exactly as csr_row_dot_f64
//    (csr.rs:38-64) and f32 sums sequentially as spmv_serial (csr.rs:195-205),
//    with FMA contraction off, so results are bitwise those of the reference;
//  - row-per-subgroup (8..64 lanes, shuffle-reduced), used when rows are long
//    (assembled high-order hex matrices); equal within rounding.
#include "internal.hpp"
#include "csr_row.hpp"

#include <algorithm>
#include <vector>
#include "reduce.hpp"

#include <cstdlib>

namespace femgpu {

// mode 0: y = Ax; 1: y += Ax; 2: y = alpha*Ax + beta*y
// STAGE: the workgroup's entries go through LDS first (large matrices).
template <class T, bool STAGE>
__global__ __launch_bounds__(256) void k_spmv_row(size_t n, const uint64_t *__restrict__ rp,
                                                  const uint32_t *__restrict__ ci,
                                                  const T *__restrict__ v,
                                                  const T *__restrict__ x, T *__restrict__ y,
                                                  int mode, T alpha, T beta, const int *skip) {
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const size_t r0 = blockIdx.x * (size_t)256;
    const size_t r = r0 + threadIdx.x;
    auto xat = [x](uint32_t j) { return x[j]; };
    T s;
    if constexpr (STAGE) {
        __shared__ RowStage<T> st;
        stage_rows<T>(st, rp, ci, v, n, r0);
        if (r >= n) return;
        s = row_staged<T>(st, rp, ci, v, xat, r);
    } else {
        if (r >= n) return;
        s = row_dot_ref<T>(rp, ci, v, xat, r);
    }
    if (mode == 0) y[r] = s;
    else if (mode == 1) y[r] += s;
    else y[r] = alpha * s + beta * y[r];
}

template <class T, int W>
__global__ __launch_bounds__(256) void k_spmv_sub(size_t n, const uint64_t *__restrict__ rp,
                                                  const uint32_t *__restrict__ ci,
                                                  const T *__restrict__ v,
                                                  const T *__restrict__ x, T *__restrict__ y,
                                                  int mode, T alpha, T beta, const int *skip) {
    if (skip && *skip) return;
    const size_t r = (blockIdx.x * (size_t)256 + threadIdx.x) / W;
    const int lane = threadIdx.x % W;
    T s = 0;
    if (r < n) {
        for (uint64_t k = rp[r] + lane; k < rp[r + 1]; k += W) s += v[k] * x[ci[k]];
    }
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, W);
    if (r < n && lane == 0) {
        if (mode == 0) y[r] = s;
        else if (mode == 1) y[r] += s;
        else y[r] = alpha * s + beta * y[r];
    }
}

template <class T>
__global__ void k_csr_diag(size_t n, const uint64_t *__restrict__ rp,
                           const uint32_t *__restrict__ ci, const T *__restrict__ v,
                           T *__restrict__ d) {
    const size_t r = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    T out = 0;
    for (uint64_t k = rp[r]; k < rp[r + 1]; ++k)
        if (ci[k] == r) { out = v[k]; break; }
    d[r] = out;
}

// apply_dirichlet_row_zeroing (linalg/src/csr.rs:297-312) for the m
// constrained rows: zero the row, set its first diagonal entry to 1 and
// rhs[row] = value.  Rows are distinct (the host keeps each DOF's last value,
// as the reference's sequential loop does), so the rows are independent.
template <class T>
__global__ void k_csr_dirichlet(size_t m, const uint32_t *__restrict__ dofs,
                                const double *__restrict__ vals, const uint64_t *__restrict__ rp,
                                const uint32_t *__restrict__ ci, T *__restrict__ v,
                                T *__restrict__ rhs) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t d = dofs[i];
    bool diag = false;
    for (uint64_t k = rp[d]; k < rp[d + 1]; ++k) {
        const bool first = !diag && ci[k] == d;
        v[k] = first ? T(1) : T(0);
        diag = diag || first;
    }
    if (rhs) rhs[d] = (T)vals[i];
}

struct CsrOp final : femgpu_op {
    uint64_t ncols = 0, nnz = 0;
    DevBuf rp, ci, vals;
    int sub = 1;  // lanes per row (1 = reference-order row kernel)
    bool stage = false;  // row kernel stages entries in LDS

    const char *name() const override { return "csr"; }

    template <class T>
    void launch(const T *x, T *y, int mode, T alpha, T beta, const int *skip) {
        hipStream_t s = ctx->stream;
        const uint64_t *r = rp.as<uint64_t>();
        const uint32_t *c = ci.as<uint32_t>();
        const T *v = vals.as<T>();
        if (n == 0) return;
        if (sub == 1) {
            if (stage)
                hipLaunchKernelGGL((k_spmv_row<T, true>), dim3((n + 255) / 256), dim3(256), 0, s, n, r, c, v, x,
                                   y, mode, alpha, beta, skip);
            else
                hipLaunchKernelGGL((k_spmv_row<T, false>), dim3((n + 255) / 256), dim3(256), 0, s, n, r, c, v,
                                   x, y, mode, alpha, beta, skip);
            return;
        }
        const size_t threads = n * (size_t)sub;
        const dim3 grid((threads + 255) / 256);
        switch (sub) {
        case 8: hipLaunchKernelGGL((k_spmv_sub<T, 8>), grid, dim3(256), 0, s, n, r, c, v, x, y, mode, alpha, beta, skip); break;
        case 16: hipLaunchKernelGGL((k_spmv_sub<T, 16>), grid, dim3(256), 0, s, n, r, c, v, x, y, mode, alpha, beta, skip); break;
        case 32: hipLaunchKernelGGL((k_spmv_sub<T, 32>), grid, dim3(256), 0, s, n, r, c, v, x, y, mode, alpha, beta, skip); break;
        default: hipLaunchKernelGGL((k_spmv_sub<T, 64>), grid, dim3(256), 0, s, n, r, c, v, x, y, mode, alpha, beta, skip); break;
        }
    }

    void apply_dev(const void *x, void *y, bool acc, const int *skip) override {
        if (dtype == FEMGPU_F64)
            launch<double>((const double *)x, (double *)y, acc ? 1 : 0, 1.0, 0.0, skip);
        else
            launch<float>((const float *)x, (float *)y, acc ? 1 : 0, 1.f, 0.f, skip);
    }

    bool csr_arrays(CsrView *out) const override {
        out->n = n;
        out->rp = rp.as<uint64_t>();
        out->ci = ci.as<uint32_t>();
        out->vals = vals.ptr;
        return true;
    }
    bool csr_view(CsrView *out) const override { return sub == 1 && csr_arrays(out); }

    void diagonal_dev(void *d) override {
        if (!n) return;
        if (dtype == FEMGPU_F64)
            hipLaunchKernelGGL(k_csr_diag<double>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                               n, rp.as<uint64_t>(), ci.as<uint32_t>(), vals.as<double>(),
                               (double *)d);
        else
            hipLaunchKernelGGL(k_csr_diag<float>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                               n, rp.as<uint64_t>(), ci.as<uint32_t>(), vals.as<float>(),
                               (float *)d);
    }
};

// lanes per row of the SpMV from the mean row length
static int csr_lanes(uint64_t nrows, uint64_t nnz) {
    const double avg = nrows ? (double)nnz / nrows : 0.0;
    int sub = avg <= 16 ? 1 : avg <= 48 ? 8 : avg <= 128 ? 16 : avg <= 400 ? 32 : 64;
    if (const char *f = std::getenv("FEMGPU_CSR_LANES")) sub = std::atoi(f);
    if (sub != 1 && sub != 8 && sub != 16 && sub != 32) sub = 64;
    return sub;
}

// LDS staging in the row kernel pays where a launch has many more workgroups
// than the GPU has CUs (2^19 rows: 2048 workgroups); on smaller matrices its
// extra pass and barrier only add latency.  FEMGPU_CSR_STAGE=1 / 0 forces it
// on / off.
static bool csr_stage(uint64_t nrows) {
    bool on = nrows >= (1u << 19);
    if (const char *f = std::getenv("FEMGPU_CSR_STAGE")) on = std::atoi(f) != 0;
    return on;
}

femgpu_op *csr_adopt(femgpu_ctx *c, uint64_t n, uint64_t nnz, DevBuf &&rp, DevBuf &&ci, DevBuf &&vals) {
    auto op = new CsrOp();
    op->ctx = c; op->dtype = FEMGPU_F64; op->n = n; op->ncols = n; op->nnz = nnz;
    op->rp = std::move(rp);
    op->ci = std::move(ci);
    op->vals = std::move(vals);
    op->sub = csr_lanes(n, nnz);
    op->stage = csr_stage(n);
    return op;
}

}  // namespace femgpu

using namespace femgpu;

extern "C" {

int femgpu_csr_create(femgpu_ctx *c, uint64_t nrows, uint64_t ncols, const uint64_t *row_ptr,
                      const uint32_t *col_idx, const void *values, femgpu_dtype dtype,
                      femgpu_op **out) {
    return guard([&] {
        FG_REQUIRE(c && row_ptr && out, FEMGPU_ERR_INVALID_ARG, "csr_create: NULL argument");
        FG_REQUIRE(dtype == FEMGPU_F64 || dtype == FEMGPU_F32, FEMGPU_ERR_INVALID_ARG,
                   "csr_create: bad dtype");
        FG_REQUIRE(row_ptr[0] == 0, FEMGPU_ERR_INVALID_ARG, "csr_create: row_ptr[0] != 0");
        const uint64_t nnz = row_ptr[nrows];
        for (uint64_t r = 0; r < nrows; ++r)
            FG_REQUIRE(row_ptr[r] <= row_ptr[r + 1], FEMGPU_ERR_INVALID_ARG,
                       "csr_create: row_ptr not monotone at row " + std::to_string(r));
        FG_REQUIRE(nnz == 0 || (col_idx && values), FEMGPU_ERR_INVALID_ARG,
                   "csr_create: NULL col_idx/values");
        for (uint64_t k = 0; k < nnz; ++k)
            FG_REQUIRE(col_idx[k] < ncols, FEMGPU_ERR_INVALID_ARG,
                       "csr_create: col_idx[" + std::to_string(k) + "] out of range");
        FG_REQUIRE(nrows == ncols, FEMGPU_ERR_DIM_MISMATCH,
                   "csr_create: operator must be square (" + std::to_string(nrows) + "x" +
                       std::to_string(ncols) + ")");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        auto op = new CsrOp();
        try {
            op->ctx = c; op->dtype = dtype; op->n = nrows; op->ncols = ncols; op->nnz = nnz;
            const size_t es = op->esize();
            op->rp.alloc((nrows + 1) * 8);
            op->ci.alloc(nnz * 4);
            op->vals.alloc(nnz * es);
            FG_HIP(hipMemcpyAsync(op->rp.ptr, row_ptr, (nrows + 1) * 8, hipMemcpyHostToDevice, c->stream));
            if (nnz) {
                FG_HIP(hipMemcpyAsync(op->ci.ptr, col_idx, nnz * 4, hipMemcpyHostToDevice, c->stream));
                FG_HIP(hipMemcpyAsync(op->vals.ptr, values, nnz * es, hipMemcpyHostToDevice, c->stream));
            }
            FG_HIP(hipStreamSynchronize(c->stream));
            op->sub = csr_lanes(nrows, nnz);
            op->stage = csr_stage(nrows);
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
        return 0;
    });
}

int femgpu_csr_spmv_add(femgpu_op *op_, double alpha, const void *x, double beta, void *y,
                        femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<CsrOp *>(op_);
        FG_REQUIRE(op && x && y, FEMGPU_ERR_INVALID_ARG, "csr_spmv_add: not a CSR operator");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        const size_t bytes = op->n * op->esize();
        const void *xd = x;
        void *yd = y;
        if (mem == FEMGPU_HOST) {
            op->stage_x.ensure(bytes); op->stage_y.ensure(bytes);
            FG_HIP(hipMemcpyAsync(op->stage_x.ptr, x, bytes, hipMemcpyHostToDevice, c->stream));
            FG_HIP(hipMemcpyAsync(op->stage_y.ptr, y, bytes, hipMemcpyHostToDevice, c->stream));
            xd = op->stage_x.ptr; yd = op->stage_y.ptr;
        }
        if (op->dtype == FEMGPU_F64)
            op->launch<double>((const double *)xd, (double *)yd, 2, alpha, beta, nullptr);
        else
            op->launch<float>((const float *)xd, (float *)yd, 2, (float)alpha, (float)beta, nullptr);
        FG_HIP(hipGetLastError());
        if (mem == FEMGPU_HOST)
            FG_HIP(hipMemcpyAsync(y, yd, bytes, hipMemcpyDeviceToHost, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

// apply_dirichlet (space/src/constraints.rs:31-42) on a device CSR matrix.
int femgpu_csr_apply_dirichlet(femgpu_op *op_, uint64_t n_constrained, const uint32_t *dofs,
                               const double *values, void *rhs, femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<CsrOp *>(op_);
        FG_REQUIRE(op, FEMGPU_ERR_INVALID_ARG, "csr_apply_dirichlet: not a CSR operator");
        FG_REQUIRE(n_constrained == 0 || (dofs && values), FEMGPU_ERR_INVALID_ARG,
                   "csr_apply_dirichlet: NULL dofs or values");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        // last value wins for repeated DOFs, as in the reference's loop
        std::vector<int64_t> last(op->n, -1);
        for (uint64_t i = 0; i < n_constrained; ++i) {
            FG_REQUIRE(dofs[i] < op->n, FEMGPU_ERR_DIM_MISMATCH,
                       "csr_apply_dirichlet: DOF " + std::to_string(dofs[i]) + " out of range");
            last[dofs[i]] = (int64_t)i;
        }
        std::vector<uint32_t> ud;
        std::vector<double> uv;
        for (uint64_t i = 0; i < n_constrained; ++i)
            if (last[dofs[i]] == (int64_t)i) { ud.push_back(dofs[i]); uv.push_back(values[i]); }
        const size_t m = ud.size();
        hipStream_t s = c->stream;
        const size_t bytes = op->n * op->esize();
        DevBuf dd(std::max<size_t>(m, 1) * 4), dv(std::max<size_t>(m, 1) * 8);
        void *rd = rhs;
        if (rhs && mem == FEMGPU_HOST) {
            op->stage_y.ensure(bytes);
            FG_HIP(hipMemcpyAsync(op->stage_y.ptr, rhs, bytes, hipMemcpyHostToDevice, s));
            rd = op->stage_y.ptr;
        }
        if (m) {
            FG_HIP(hipMemcpyAsync(dd.ptr, ud.data(), m * 4, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemcpyAsync(dv.ptr, uv.data(), m * 8, hipMemcpyHostToDevice, s));
            const dim3 grid((unsigned)((m + 255) / 256));
            if (op->dtype == FEMGPU_F64)
                hipLaunchKernelGGL(k_csr_dirichlet<double>, grid, dim3(256), 0, s, m, (const uint32_t *)dd.ptr,
                                   (const double *)dv.ptr, op->rp.as<uint64_t>(), op->ci.as<uint32_t>(),
                                   op->vals.as<double>(), (double *)rd);
            else
                hipLaunchKernelGGL(k_csr_dirichlet<float>, grid, dim3(256), 0, s, m, (const uint32_t *)dd.ptr,
                                   (const double *)dv.ptr, op->rp.as<uint64_t>(), op->ci.as<uint32_t>(),
                                   op->vals.as<float>(), (float *)rd);
            FG_HIP(hipGetLastError());
        }
        op->diag_ready = false;  // the cached diagonal no longer holds
        if (rhs && mem == FEMGPU_HOST) FG_HIP(hipMemcpyAsync(rhs, rd, bytes, hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

// Copy of the matrix values (nnz entries of the operator's dtype).
int femgpu_csr_get_values(femgpu_op *op_, void *values, femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<CsrOp *>(op_);
        FG_REQUIRE(op && values, FEMGPU_ERR_INVALID_ARG, "csr_get_values: not a CSR operator");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        FG_HIP(hipMemcpyAsync(values, op->vals.ptr, op->nnz * op->esize(),
                              mem == FEMGPU_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

int femgpu_csr_get_pattern(femgpu_op *op_, uint64_t *nnz, uint64_t *row_ptr, uint32_t *col_idx,
                           femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<CsrOp *>(op_);
        FG_REQUIRE(op && nnz, FEMGPU_ERR_INVALID_ARG, "csr_get_pattern: not a CSR operator");
        *nnz = op->nnz;
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        const hipMemcpyKind k = mem == FEMGPU_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        if (row_ptr) FG_HIP(hipMemcpyAsync(row_ptr, op->rp.ptr, (op->n + 1) * 8, k, c->stream));
        if (col_idx && op->nnz) FG_HIP(hipMemcpyAsync(col_idx, op->ci.ptr, op->nnz * 4, k, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

}  // extern "C"
