// One CSR row times a vector in the reference's summation order, shared by
// the SpMV kernels and the fused PCG-Jacobi iteration.
//  - f64: blocks of four, s += v0 x0 + v1 x1 + v2 x2 + v3 x3, then the tail
//    one by one (csr_row_dot_f64, linalg/src/csr.rs:38-64);
//  - f32: one by one (spmv_serial, linalg/src/csr.rs:195-205).
// FMA contraction is off, so the sums are bitwise the reference's.  xat(j)
// returns x_j (a plain load, or a value formed on the fly).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace femgpu {

struct CsrView {
    uint64_t n = 0;
    const uint64_t *rp = nullptr;
    const uint32_t *ci = nullptr;
    const void *vals = nullptr;
};

// Entries s .. e-1 of a row; vat(k) / cat(k) return value and column k.
template <class T, class V, class Cl, class X>
__device__ __forceinline__ T row_sum_ref(uint64_t s, uint64_t e, V vat, Cl cat, X xat) {
#pragma clang fp contract(off)
    T sum = 0;
    uint64_t k = s;
    if constexpr (sizeof(T) == 8) {
        const uint64_t e4 = s + (e - s) / 4 * 4;
        for (; k < e4; k += 4)
            sum += vat(k) * xat(cat(k)) + vat(k + 1) * xat(cat(k + 1)) + vat(k + 2) * xat(cat(k + 2)) +
                   vat(k + 3) * xat(cat(k + 3));
    }
    for (; k < e; ++k) sum += vat(k) * xat(cat(k));
    return sum;
}

template <class T, class X>
__device__ __forceinline__ T row_dot_ref(const uint64_t *__restrict__ rp,
                                         const uint32_t *__restrict__ ci, const T *__restrict__ v,
                                         X xat, size_t row) {
    return row_sum_ref<T>(rp[row], rp[row + 1], [v](uint64_t k) { return v[k]; },
                          [ci](uint64_t k) { return ci[k]; }, xat);
}

// Row kernels with one thread per row, 256 rows per workgroup: the
// workgroup's rows hold one contiguous range of entries.  When it fits
// kRowStage entries (8 per row on average), values and columns are first
// copied to LDS with coalesced loads, and each lane then sums its row from
// there in row_dot_ref's order, so a lane's loads no longer wait on one
// another.  Every thread of the workgroup calls stage_rows (it holds a
// barrier); row_staged then returns row r's sum (r < n).
constexpr int kRowStage = 2048;

template <class T>
struct RowStage {
    T v[kRowStage];
    uint32_t c[kRowStage];
    uint64_t k0;
    bool on;
};

template <class T>
__device__ __forceinline__ void stage_rows(RowStage<T> &st, const uint64_t *__restrict__ rp,
                                           const uint32_t *__restrict__ ci, const T *__restrict__ v,
                                           size_t n, size_t r0) {
    const size_t r1 = r0 + 256 < n ? r0 + 256 : n;
    const uint64_t k0 = rp[r0], k1 = rp[r1];
    const bool on = k1 - k0 <= (uint64_t)kRowStage;  // uniform over the workgroup
    if (on)
        for (uint64_t k = k0 + threadIdx.x; k < k1; k += 256) {
            st.v[k - k0] = __builtin_nontemporal_load(v + k);
            st.c[k - k0] = __builtin_nontemporal_load(ci + k);
        }
    if (threadIdx.x == 0) {
        st.k0 = k0;
        st.on = on;
    }
    __syncthreads();
}

template <class T, class X>
__device__ __forceinline__ T row_staged(const RowStage<T> &st, const uint64_t *__restrict__ rp,
                                        const uint32_t *__restrict__ ci, const T *__restrict__ v,
                                        X xat, size_t r) {
    if (!st.on) return row_dot_ref<T>(rp, ci, v, xat, r);
    const uint64_t k0 = st.k0;
    return row_sum_ref<T>(rp[r] - k0, rp[r + 1] - k0, [&st](uint64_t k) { return st.v[k]; },
                          [&st](uint64_t k) { return st.c[k]; }, xat);
}

}  // namespace femgpu
