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

template <class T, class X>
__device__ __forceinline__ T row_dot_ref(const uint64_t *__restrict__ rp,
                                         const uint32_t *__restrict__ ci, const T *__restrict__ v,
                                         X xat, size_t row) {
#pragma clang fp contract(off)
    const uint64_t s = rp[row], e = rp[row + 1];
    T sum = 0;
    uint64_t k = s;
    if constexpr (sizeof(T) == 8) {
        const uint64_t e4 = s + (e - s) / 4 * 4;
        for (; k < e4; k += 4)
            sum += v[k] * xat(ci[k]) + v[k + 1] * xat(ci[k + 1]) + v[k + 2] * xat(ci[k + 2]) +
                   v[k + 3] * xat(ci[k + 3]);
    }
    for (; k < e; ++k) sum += v[k] * xat(ci[k]);
    return sum;
}

}  // namespace femgpu
