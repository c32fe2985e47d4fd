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

}  // namespace femgpu
