// Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw, SC'11), usable
// on host and device so generators give the same numbers on both.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace femgpu {

__host__ __device__ inline double philox_u01(uint64_t seed, uint32_t stream, uint64_t index) {
    uint32_t c0 = (uint32_t)index, c1 = (uint32_t)(index >> 32), c2 = stream, c3 = 0u;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
    }
    const uint64_t bits = (((uint64_t)c0 << 32) | c1) >> 11;
    return (double)bits * (1.0 / 9007199254740992.0);
}

}  // namespace femgpu
