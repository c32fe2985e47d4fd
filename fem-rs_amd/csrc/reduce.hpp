// Device-side reduction helpers: 64-lane wavefront shuffles, then LDS across
// the waves of a block.  Block sums are written to a fixed-size partials
// array and reduced by one block in a fixed order, so every dot product and
// norm is bitwise reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>

namespace femgpu {

// The first Chebyshev iterate z0 = d0 = (1/theta) (inv r), as the kernels that
// store it compute it (two rounded products); kept out of FMA contraction so
// that a kernel recomputing it gets the stored value bit for bit.
template <class T> __device__ __forceinline__ T cheb_first(double inv_theta, T inv, T r) {
#pragma clang fp contract(off)
    return (T)inv_theta * (inv * r);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sums NV values per thread across the block; result valid in thread 0.
template <int NV, int BLOCK>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    static_assert(BLOCK % 64 == 0, "block must be whole wavefronts");
    constexpr int NW = BLOCK / 64;
    __shared__ double red[NV][NW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k][wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < NW; ++w) s += red[k][w];
            v[k] = s;
        }
    }
}

// Number of partial blocks used for a vector of length n (fixed per n).
inline int red_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (int)b;
}

}  // namespace femgpu
