// Vector kernels (BLAS-1) and deterministic dot products.
//
// These are the device forms of the reference's Vector/ParVector helpers
// (linalg/src/vector.rs:39-70, parallel/src/par_vector.rs:182-218,
// parallel/src/par_solver.rs:15-78).  Dots accumulate in double for both f64
// and f32 vectors.
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include "reduce.hpp"

namespace femgpu {

static inline unsigned grid_for(size_t n, int block = 256) {
    size_t g = (n + block - 1) / block;
    if (g > 65535u * 16u) g = 65535u * 16u;
    return (unsigned)(g ? g : 1);
}

template <class T>
__global__ void k_fill(T *__restrict__ y, T v, size_t n, const int *skip) {
    if (skip && *skip) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = v;
}

template <class T>
__global__ void k_copy(T *__restrict__ y, const T *__restrict__ x, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = x[i];
}

template <class T>
__global__ void k_axpby(T *__restrict__ y, double a, const T *__restrict__ x, double b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

template <class T>
__global__ void k_inv_diag(T *__restrict__ inv, const T *__restrict__ d, size_t n, double tiny) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const T v = d[i];
        inv[i] = (fabs((double)v) > tiny) ? (T)(1.0 / (double)v) : (T)1;
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_dot_partial(const T *__restrict__ a,
                                                     const T *__restrict__ b, size_t n,
                                                     double *__restrict__ partials) {
    double v[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        v[0] += (double)a[i] * (double)b[i];
    block_sum<1, 256>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(1024) void k_finalize1(const double *__restrict__ partials, int nb,
                                                    double *__restrict__ out) {
    double v[1] = {threadIdx.x < (unsigned)nb ? partials[threadIdx.x] : 0.0};
    block_sum<1, 1024>(v);
    if (threadIdx.x == 0) out[0] = v[0];
}

__global__ void k_f64_to_f32(float *__restrict__ y, const double *__restrict__ x, size_t n,
                             double scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = (float)(x[i] * scale);
}

__global__ void k_f32_axpy_f64(double *__restrict__ y, const float *__restrict__ x, size_t n,
                               double scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] += scale * (double)x[i];
}

template <class T> void fill(hipStream_t s, T *y, T v, size_t n, const int *skip) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(256), 0, s, y, v, n, skip);
}
template <class T> void copy(hipStream_t s, T *y, const T *x, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(n)), dim3(256), 0, s, y, x, n);
}
template <class T> void axpby(hipStream_t s, T *y, double a, const T *x, double b, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_axpby<T>, dim3(grid_for(n)), dim3(256), 0, s, y, a, x, b, n);
}
template <class T> void inv_diag(hipStream_t s, T *inv, const T *d, size_t n, double tiny) {
    if (!n) return;
    hipLaunchKernelGGL(k_inv_diag<T>, dim3(grid_for(n)), dim3(256), 0, s, inv, d, n, tiny);
}
template <class T> void dot(femgpu_ctx *c, const T *a, const T *b, size_t n, double *out) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(k_dot_partial<T>, dim3(nb), dim3(256), 0, c->stream, a, b, n,
                       c->partials.as<double>());
    hipLaunchKernelGGL(k_finalize1, dim3(1), dim3(1024), 0, c->stream,
                       (const double *)c->partials.as<double>(), nb, out);
}

double dot_host(femgpu_ctx *c, const void *a, const void *b, size_t n, femgpu_dtype dt) {
    double *out = c->scalars.as<double>() + 63;
    if (dt == FEMGPU_F64) dot<double>(c, (const double *)a, (const double *)b, n, out);
    else dot<float>(c, (const float *)a, (const float *)b, n, out);
    FG_HIP(hipMemcpyAsync(c->pinned + 63, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    return c->pinned[63];
}

// ---- multi-vector kernels (flexible CG orthogonalisation) ----
// partials[j][block] of a . v[j], j < m <= 8, in one pass over a
template <class T>
__global__ __launch_bounds__(256) void k_multidot(const T *__restrict__ a, VecPack<T> vs, size_t n,
                                                  double *__restrict__ partials) {
    double v[kPack];
#pragma unroll
    for (int j = 0; j < kPack; ++j) v[j] = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double ai = (double)a[i];
#pragma unroll
        for (int j = 0; j < kPack; ++j)
            if (j < vs.m) v[j] += ai * (double)vs.v[j][i];
    }
    block_sum<kPack, 256>(v);
    if (threadIdx.x == 0)
        for (int j = 0; j < vs.m; ++j) partials[(size_t)j * kRedBlocks + blockIdx.x] = v[j];
}

__global__ __launch_bounds__(1024) void k_finalize_m(const double *__restrict__ partials, int nb,
                                                     int m, double *__restrict__ out) {
    for (int j = 0; j < m; ++j) {
        double v[1] = {threadIdx.x < (unsigned)nb ? partials[(size_t)j * kRedBlocks + threadIdx.x] : 0.0};
        block_sum<1, 1024>(v);
        if (threadIdx.x == 0) out[j] = v[0];
        __syncthreads();
    }
}

// y = x0 + sum_{j<m} c[j] v[j] (m <= 8); y may be x0
template <class T>
__global__ void k_multiaxpy(T *y, const T *x0, VecPack<T> vs, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        double s = (double)x0[i];
#pragma unroll
        for (int j = 0; j < kPack; ++j)
            if (j < vs.m) s += vs.c[j] * (double)vs.v[j][i];
        y[i] = (T)s;
    }
}

template <class T>
void multidot_host(femgpu_ctx *c, const T *a, const T *const *v, int m, size_t n, double *out) {
    FG_REQUIRE(m >= 0 && m <= kMultiDotMax, FEMGPU_ERR_INVALID_ARG, "multidot: too many vectors");
    if (m == 0) return;
    const int nb = red_blocks(n);
    double *o = c->scalars.as<double>() + kMultiDotBase;
    for (int j0 = 0; j0 < m; j0 += kPack) {
        VecPack<T> p{};
        p.m = std::min(kPack, m - j0);
        for (int j = 0; j < p.m; ++j) p.v[j] = v[j0 + j];
        hipLaunchKernelGGL(k_multidot<T>, dim3(nb), dim3(256), 0, c->stream, a, p, n, c->partials.as<double>());
        hipLaunchKernelGGL(k_finalize_m, dim3(1), dim3(1024), 0, c->stream,
                           (const double *)c->partials.as<double>(), nb, p.m, o + j0);
    }
    FG_HIP(hipMemcpyAsync(c->pinned + kMultiDotBase, o, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->pinned + kMultiDotBase, sizeof(double) * m);
}

template <class T>
void multiaxpy(hipStream_t s, T *y, const T *x0, const T *const *v, const double *coef, int m, size_t n) {
    if (m == 0) {
        if (y != x0) copy<T>(s, y, x0, n);
        return;
    }
    for (int j0 = 0; j0 < m; j0 += kPack) {
        VecPack<T> p{};
        p.m = std::min(kPack, m - j0);
        for (int j = 0; j < p.m; ++j) { p.v[j] = v[j0 + j]; p.c[j] = coef[j0 + j]; }
        hipLaunchKernelGGL(k_multiaxpy<T>, dim3(grid_for(n)), dim3(256), 0, s, y, j0 == 0 ? x0 : (const T *)y, p, n);
    }
}

void dots_host(femgpu_ctx *c, int m, const void *const *a, const void *const *b, size_t n,
               femgpu_dtype dt, double *out) {
    FG_REQUIRE(m >= 0 && m <= kMultiDotMax, FEMGPU_ERR_INVALID_ARG, "dots_host: too many dot products");
    if (m == 0) return;
    double *o = c->scalars.as<double>() + kMultiDotBase;
    for (int j = 0; j < m; ++j) {
        if (dt == FEMGPU_F64) dot<double>(c, (const double *)a[j], (const double *)b[j], n, o + j);
        else dot<float>(c, (const float *)a[j], (const float *)b[j], n, o + j);
    }
    FG_HIP(hipMemcpyAsync(c->pinned + kMultiDotBase, o, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->pinned + kMultiDotBase, sizeof(double) * m);
}

void convert_f64_to_f32(hipStream_t s, float *y, const double *x, size_t n, double scale) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n)), dim3(256), 0, s, y, x, n, scale);
}
void convert_f32_to_f64_axpy(hipStream_t s, double *y, const float *x, size_t n, double scale) {
    hipLaunchKernelGGL(k_f32_axpy_f64, dim3(grid_for(n)), dim3(256), 0, s, y, x, n, scale);
}

#define INST(T)                                                                          \
    template void fill<T>(hipStream_t, T *, T, size_t, const int *);                     \
    template void copy<T>(hipStream_t, T *, const T *, size_t);                          \
    template void axpby<T>(hipStream_t, T *, double, const T *, double, size_t);         \
    template void inv_diag<T>(hipStream_t, T *, const T *, size_t, double);              \
    template void dot<T>(femgpu_ctx *, const T *, const T *, size_t, double *);          \
    template void multidot_host<T>(femgpu_ctx *, const T *, const T *const *, int, size_t, double *); \
    template void multiaxpy<T>(hipStream_t, T *, const T *, const T *const *, const double *, int, size_t);
INST(double)
INST(float)
#undef INST

}  // namespace femgpu

using namespace femgpu;

extern "C" {

int femgpu_vec_axpby(femgpu_ctx *c, femgpu_dtype dtype, uint64_t n, double alpha, const void *x, double beta,
                     void *y) {
    return guard([&] {
        FG_REQUIRE(c && (n == 0 || (x && y)), FEMGPU_ERR_INVALID_ARG, "vec_axpby: NULL argument");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        if (n == 0) return 0;
        if (dtype == FEMGPU_F64) axpby<double>(c->stream, (double *)y, alpha, (const double *)x, beta, n);
        else axpby<float>(c->stream, (float *)y, alpha, (const float *)x, beta, n);
        FG_HIP(hipGetLastError());
        return 0;
    });
}

int femgpu_vec_convert(femgpu_ctx *c, uint64_t n, femgpu_dtype src_dtype, const void *src, femgpu_dtype dst_dtype,
                       void *dst) {
    return guard([&] {
        FG_REQUIRE(c && (n == 0 || (src && dst)), FEMGPU_ERR_INVALID_ARG, "vec_convert: NULL argument");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        if (n == 0) return 0;
        if (src_dtype == dst_dtype) {
            FG_HIP(hipMemcpyAsync(dst, src, n * (src_dtype == FEMGPU_F64 ? 8 : 4), hipMemcpyDeviceToDevice,
                                  c->stream));
        } else if (src_dtype == FEMGPU_F64) {
            convert_f64_to_f32(c->stream, (float *)dst, (const double *)src, n, 1.0);
        } else {
            fill<double>(c->stream, (double *)dst, 0.0, n);
            convert_f32_to_f64_axpy(c->stream, (double *)dst, (const float *)src, n, 1.0);
        }
        FG_HIP(hipGetLastError());
        return 0;
    });
}

}  // extern "C"
