// Learned preconditioner: a 3-level 3-D convolutional U-Net on the structured
// DOF grid, BF16 activations and weights on MFMA with FP32 accumulation, with
// hand-written forward, backward and Adam kernels.  No reference counterpart.
//
// Network (grid levels D0 = DOF grid, D1 = ceil(D0/2), D2 = ceil(D1/2);
// channels 16 / 32 / 64; every conv is 3x3x3 with zero padding):
//   e0 = relu(conv_{4->16}(in))            @D0   in = (r/|r|, aux, 0, 0)
//   e1 = relu(conv_{16->32}(pool(e0)))     @D1
//   e2 = relu(conv_{32->64}(pool(e1)))     @D2
//   d1 = relu(up(conv_{64->32}(e2)) + e1)  @D1
//   d0 = relu(up(conv_{32->16}(d1)) + e0)  @D0
//   out = conv_{16->1}(d0)                 @D0
// pool = 2x2x2 average (ceil), up = nearest.  The decoder convolves at the
// coarse level before upsampling, so the full-resolution work is two convs.
//
// Activations are channels-last, v = x + nx (y + ny z).  A conv is an implicit
// GEMM: M = voxels, N = output channels, K = 27 taps x input channels with
// k = tap * Cin + c, tap = (dz+1)*9 + (dy+1)*3 + (dx+1).  Weights are stored
// [Cout][Kpad] (Kpad = 27*Cin rounded up to 16) so one lane's four K values are
// one 8-byte load.  MFMA v_mfma_f32_16x16x16_bf16: lane l holds A[m = l%16]
// [k = 4(l/16)..+3], B[k = 4(l/16)..+3][n = l%16] and D[m = 4(l/16)+i][n = l%16].
#include "internal.hpp"
#include "philox.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>

namespace femgpu {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Convolutions with CIN >= 16 take two taps per MFMA with gfx950's
// v_mfma_f32_16x16x32_bf16 (lane l holds A[l % 16][8 (l / 16) + i] and
// B[8 (l / 16) + i][l % 16], i < 8; D as for 16x16x16), instead of one tap per
// v_mfma_f32_16x16x16bf16_1k (DK = false; FEMGPU_CONV_K16=1 selects it, for
// comparison).

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Vol {
    int nx, ny, nz;
    __host__ __device__ long long n() const { return (long long)nx * ny * nz; }
};

static inline Vol half(Vol d) { return Vol{(d.nx + 1) / 2, (d.ny + 1) / 2, (d.nz + 1) / 2}; }

static constexpr int kpad_of(int cin) { return (27 * cin + 15) / 16 * 16; }

static unsigned g1(long long n, int b = 256) {
    long long g = (n + b - 1) / b;
    if (g > (1ll << 30)) g = 1ll << 30;
    return (unsigned)(g < 1 ? 1 : g);
}

// Output layer weights w[432] (f32) -> B matrix [16][432] (bf16): row 0 the
// high parts bf16(w), row 1 the low parts bf16(w - high), rows 2..15 zero.
__global__ void k_split_out_weights(const float *__restrict__ w, uint16_t *__restrict__ B) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 16 * 432) return;
    const int n = id / 432, k = id - n * 432;
    uint16_t v = 0;
    if (n < 2) {
        const uint16_t hi = f2bf(w[k]);
        v = n == 0 ? hi : f2bf(w[k] - bf2f(hi));
    }
    B[id] = v;
}

// ------------------------------------------------------------ conv fwd --
// out[v][n] = act(sum_k A[v][k] W[n][k] + bias[n] (+ skip[v][n]))
// OUTF: 0 -> bf16 output with bias/skip/relu; 1 -> f32 output, plain sum;
//       2 -> the 1-channel output layer: W holds the f32 weights split into
//            bf16 high and low parts (B columns 0 and 1, the rest zero;
//            k_split_out_weights), so the product keeps the f32 weights to
//            about 2^-16; out[v] = bias[0] + (column 0 + column 1) (f32 [V]).
// Implicit GEMM on v_mfma_f32_16x16x16bf16_1k, K = 27 CIN in steps of 16.
// A workgroup owns an output tile of 16 x 4 x 4 voxels (x, y, z) and NTB x 16
// output channels; it stages the tile's input halo brick (18 x 6 x 6 voxels,
// zero outside the grid) in LDS 16 channels at a time (20.7 KB whatever CIN
// is, so the coarse levels keep several workgroups per CU), and the 27 taps
// read LDS instead of re-reading HBM / L2.  Wave w computes z = z0 + w, its
// four 16-voxel M tiles are the rows y = y0..y0+3.  For CIN >= 16 a k-step is
// one tap and the 16 staged channels (tap, offset and channel base are
// wave-uniform); CIN = 4 packs four taps per k-step (lane group g = lane / 16
// takes tap 4 s + g).
constexpr int kBX = 16, kBY = 4, kBZ = 4;
constexpr int kHX = kBX + 2, kHY = kBY + 2, kHZ = kBZ + 2;

template <int CIN, int COUT, int NTB, int OUTF, bool DK>
__global__ __launch_bounds__(256) void k_conv_fwd(Vol d, int kpad, const uint16_t *__restrict__ in,
                                                  const uint16_t *__restrict__ W,
                                                  const float *__restrict__ bias,
                                                  const uint16_t *__restrict__ skip,
                                                  void *__restrict__ outp, int relu) {
    static_assert(CIN == 4 || CIN % 16 == 0, "CIN must be 4 or a multiple of 16");
    static_assert(OUTF != 2 || (CIN == 16 && COUT == 16 && NTB == 1), "output layer is 16 -> 1");
    static_assert(COUT % (16 * NTB) == 0, "NTB must divide the output tiles");
    constexpr int NBRICK = kHX * kHY * kHZ;
    constexpr int CG = CIN >= 16 ? 16 : CIN;  // channels staged per pass
    constexpr int NPASS = CIN / CG;
    constexpr int NCH = NTB * 16;           // output channels of this workgroup
    constexpr int ESZ = OUTF == 0 ? 2 : 4;  // output element bytes
    constexpr int SIN_B = NBRICK * CG * 2, SOUT_B = 256 * NCH * ESZ;
    __shared__ __align__(16) char smem[SIN_B > SOUT_B ? SIN_B : SOUT_B];
    uint16_t *sIn = reinterpret_cast<uint16_t *>(smem);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntx = (d.nx + kBX - 1) / kBX;
    const int csplit = COUT / (16 * NTB);  // output-channel groups per tile
    const int bx = blockIdx.x % ntx, cg = (blockIdx.x / ntx) % csplit;
    const int x0 = bx * kBX, y0 = blockIdx.y * kBY, z0 = blockIdx.z * kBZ;
    f32x4 acc[kBY][NTB];
#pragma unroll
    for (int mt = 0; mt < kBY; ++mt)
#pragma unroll
        for (int nt = 0; nt < NTB; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int ksub = (lane >> 4) * 4;
    const int xl = lane & 15;
    const uint16_t *wrow[NTB];
#pragma unroll
    for (int nt = 0; nt < NTB; ++nt)
        wrow[nt] = W + (size_t)((cg * NTB + nt) * 16 + (lane & 15)) * kpad + ksub;
    // staging chunks: 16 bytes (8 channels) when 16 channels are staged, else 8
    constexpr int CK = CG >= 8 ? 8 : 4, NCK = CG / CK;
    for (int pass = 0; pass < NPASS; ++pass) {
        // ---- stage the halo brick's channels [pass CG, pass CG + CG)
        if (pass) __syncthreads();  // the previous pass's reads are done
        for (int idx = threadIdx.x; idx < NBRICK * NCK; idx += 256) {
            const int vi = idx / NCK, ck = idx - vi * NCK;
            const int xx = vi % kHX, yy = (vi / kHX) % kHY, zz = vi / (kHX * kHY);
            const int gx = x0 - 1 + xx, gy = y0 - 1 + yy, gz = z0 - 1 + zz;
            const bool inb = gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz;
            const uint16_t *src = in + ((size_t)gx + (size_t)d.nx * ((size_t)gy + (size_t)d.ny * gz)) * CIN + pass * CG + ck * CK;
            if constexpr (CK == 8) {
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (inb) v = *reinterpret_cast<const uint4 *>(src);
                *reinterpret_cast<uint4 *>(sIn + (size_t)vi * CG + ck * CK) = v;
            } else {
                uint2 v = make_uint2(0u, 0u);
                if (inb) v = *reinterpret_cast<const uint2 *>(src);
                *reinterpret_cast<uint2 *>(sIn + (size_t)vi * CG + ck * CK) = v;
            }
        }
        __syncthreads();
        // ---- MFMA loop over the staged channels, one dz plane (9 taps) at a
        // time; the plane's weights are loaded together before its MFMAs, so
        // their latency is paid 3 times per pass rather than 27
        if constexpr (CIN >= 16 && DK) {
            // K = 32 per MFMA: lane group g = lane / 16 takes tap 2 pp + g / 2 of
            // the plane and staged channels 8 (g % 2) .. + 7; a plane's 9 taps
            // are 5 pairs, the last one padded with zeros
            const int g4 = lane >> 4, half = (g4 & 1) * 8, tsel = g4 >> 1;
            const u32x4 zero4 = {0u, 0u, 0u, 0u};
            for (int pz = 0; pz < 3; ++pz) {
                bf16x8 b[5][NTB];
#pragma unroll
                for (int pp = 0; pp < 5; ++pp) {
                    const int t9 = 2 * pp + tsel;
#pragma unroll
                    for (int nt = 0; nt < NTB; ++nt) {
                        u32x4 w = zero4;
                        if (t9 < 9)
                            w = *reinterpret_cast<const u32x4 *>(wrow[nt] - ksub + half + (pz * 9 + t9) * CIN + pass * 16);
                        b[pp][nt] = __builtin_bit_cast(bf16x8, w);
                    }
                }
#pragma unroll
                for (int pp = 0; pp < 5; ++pp) {
                    const int t9 = 2 * pp + tsel;
                    const bool tv = t9 < 9;
                    const int dx = t9 % 3 - 1, dy = t9 / 3 - 1, dz = pz - 1;
                    const int base = ((wave + 1 + dz) * kHY + 1 + dy) * kHX + xl + 1 + dx;
#pragma unroll
                    for (int mt = 0; mt < kBY; ++mt) {
                        u32x4 av = zero4;
                        if (tv) av = *reinterpret_cast<const u32x4 *>(sIn + (size_t)(base + mt * kHX) * CG + half);
                        const bf16x8 a = __builtin_bit_cast(bf16x8, av);
#pragma unroll
                        for (int nt = 0; nt < NTB; ++nt)
                            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[pp][nt], acc[mt][nt], 0, 0, 0);
                    }
                }
            }
        } else if constexpr (CIN >= 16) {
            for (int pz = 0; pz < 3; ++pz) {
                bf16x4 b[9][NTB];
#pragma unroll
                for (int t9 = 0; t9 < 9; ++t9)
#pragma unroll
                    for (int nt = 0; nt < NTB; ++nt)
                        b[t9][nt] = *reinterpret_cast<const bf16x4 *>(wrow[nt] + (pz * 9 + t9) * CIN + pass * 16);
#pragma unroll
                for (int t9 = 0; t9 < 9; ++t9) {
                    const int dx = t9 % 3 - 1, dy = t9 / 3 - 1, dz = pz - 1;
                    const int base = ((wave + 1 + dz) * kHY + 1 + dy) * kHX + xl + 1 + dx;
#pragma unroll
                    for (int mt = 0; mt < kBY; ++mt) {
                        const bf16x4 a = *reinterpret_cast<const bf16x4 *>(sIn + (size_t)(base + mt * kHX) * CG + ksub);
#pragma unroll
                        for (int nt = 0; nt < NTB; ++nt)
                            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t9][nt], acc[mt][nt], 0, 0, 0);
                    }
                }
            }
        } else {
            static_assert(CIN != 4 || kpad_of(4) == 112, "CIN = 4 takes 7 k-steps");
            bf16x4 bw[7][NTB];
#pragma unroll
            for (int ks = 0; ks < 7; ++ks)
#pragma unroll
                for (int nt = 0; nt < NTB; ++nt) bw[ks][nt] = *reinterpret_cast<const bf16x4 *>(wrow[nt] + ks * 16);
#pragma unroll
            for (int ks = 0; ks < 7; ++ks) {
                const int k0 = ks * 16;
                bf16x4 b[NTB];
#pragma unroll
                for (int nt = 0; nt < NTB; ++nt) b[nt] = bw[ks][nt];
                const int tap = k0 / 4 + (lane >> 4);
                const bool tv = tap < 27;
                const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
                const int base = ((wave + 1 + dz) * kHY + 1 + dy) * kHX + xl + 1 + dx;
#pragma unroll
                for (int mt = 0; mt < kBY; ++mt) {
                    bf16x4 a = bf16x4{0, 0, 0, 0};
                    if (tv) a = *reinterpret_cast<const bf16x4 *>(sIn + (size_t)(base + mt * kHX) * CG);
#pragma unroll
                    for (int nt = 0; nt < NTB; ++nt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nt], acc[mt][nt], 0, 0, 0);
                }
            }
        }
    }
    // ---- epilogue: D[m = voxel x (4 (lane/16) + i)][n = lane % 16]
    if constexpr (OUTF == 2) {
        const int z = z0 + wave;
        const float b0 = bias[0];
#pragma unroll
        for (int mt = 0; mt < kBY; ++mt) {
            const int y = y0 + mt;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float lo = __shfl_down(acc[mt][0][i], 1, 64);  // column 1 of the same voxel
                const int x = x0 + (lane >> 4) * 4 + i;
                if ((lane & 15) || x >= d.nx || y >= d.ny || z >= d.nz) continue;
                reinterpret_cast<float *>(outp)[(size_t)x + (size_t)d.nx * ((size_t)y + (size_t)d.ny * z)] =
                    b0 + (acc[mt][0][i] + lo);
            }
        }
        return;
    } else {
        // Through LDS, so that global stores are 16-byte chunks: the tile's
        // voxels [z][y][x] each hold NCH contiguous output channels.
        __syncthreads();  // every wave is done reading the brick
#pragma unroll
        for (int mt = 0; mt < kBY; ++mt)
#pragma unroll
            for (int nt = 0; nt < NTB; ++nt) {
                const int nl = nt * 16 + (lane & 15), n = cg * NCH + nl;
                const float bn = (OUTF == 0 && bias) ? bias[n] : 0.f;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int vl = (wave * kBY + mt) * kBX + (lane >> 4) * 4 + i;
                    float val = acc[mt][nt][i];
                    if constexpr (OUTF == 1) {
                        reinterpret_cast<float *>(smem)[vl * NCH + nl] = val;
                    } else {
                        val += bn;
                        if (skip) {
                            const int x = x0 + (lane >> 4) * 4 + i, y = y0 + mt, z = z0 + wave;
                            if (x < d.nx && y < d.ny && z < d.nz)
                                val += bf2f(skip[((size_t)x + (size_t)d.nx * ((size_t)y + (size_t)d.ny * z)) * COUT + n]);
                        }
                        if (relu) val = fmaxf(val, 0.f);
                        reinterpret_cast<uint16_t *>(smem)[vl * NCH + nl] = f2bf(val);
                    }
                }
            }
        __syncthreads();
        constexpr int CPV = NCH * ESZ / 16;  // 16-byte chunks per voxel
        for (int id = threadIdx.x; id < 256 * CPV; id += 256) {
            const int vl = id / CPV, part = id - vl * CPV;
            const int x = x0 + (vl & 15), y = y0 + ((vl >> 4) & 3), z = z0 + (vl >> 6);
            if (x >= d.nx || y >= d.ny || z >= d.nz) continue;
            const size_t v = (size_t)x + (size_t)d.nx * ((size_t)y + (size_t)d.ny * z);
            *reinterpret_cast<uint4 *>(reinterpret_cast<char *>(outp) + (v * COUT + (size_t)cg * NCH) * ESZ + part * 16) =
                *reinterpret_cast<const uint4 *>(smem + (size_t)vl * NCH * ESZ + part * 16);
        }
    }
}

// ----------------------------------------------------- conv weight grad --
// part[tap][chunk][n][c] = sum_{v in chunk} dy[v][n] x[v + off(tap)][c],
// n < cout (padded to 16s), c < cin (padded to 16s).  Grid (chunks, 27).
template <int NT, int CT>
__global__ __launch_bounds__(256) void k_conv_wgrad(Vol d, int cin, int cout,
                                                    const uint16_t *__restrict__ x,
                                                    const uint16_t *__restrict__ dy,
                                                    float *__restrict__ part, long long chunk) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tap = blockIdx.y;
    const int dx = tap % 3 - 1, dyo = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    const long long V = d.n();
    const long long v0 = (long long)blockIdx.x * chunk, v1 = v0 + chunk < V ? v0 + chunk : V;
    f32x4 acc[NT][CT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[nt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long vs = v0 + wave * 16; vs < v1; vs += 64) {
        bf16x4 a[NT], b[CT];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long v = vs + (lane >> 4) * 4 + i;
            const bool ok = v < v1;
            // A[m = n][k = v]
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int n = nt * 16 + (lane & 15);
                a[nt][i] = (ok && n < cout) ? (short)dy[(size_t)v * cout + n] : (short)0;
            }
            // B[k = v][n = c]
            size_t nb = 0;
            bool in = false;
            if (ok) {
                const int xx = (int)(v % d.nx);
                const long long r = v / d.nx;
                const int yy = (int)(r % d.ny), zz = (int)(r / d.ny);
                const int X = xx + dx, Y = yy + dyo, Z = zz + dz;
                in = X >= 0 && X < d.nx && Y >= 0 && Y < d.ny && Z >= 0 && Z < d.nz;
                nb = (size_t)X + (size_t)d.nx * ((size_t)Y + (size_t)d.ny * Z);
            }
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int c = ct * 16 + (lane & 15);
                b[ct][i] = (in && c < cin) ? (short)x[nb * cin + c] : (short)0;
            }
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                acc[nt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[nt], b[ct], acc[nt][ct], 0, 0, 0);
    }
    // block reduction of the 4 waves through LDS, then one partial per block
    __shared__ float red[4][NT * CT * 256];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = (lane >> 4) * 4 + i, c = lane & 15;
                red[wave][(nt * CT + ct) * 256 + m * 16 + c] = acc[nt][ct][i];
            }
    __syncthreads();
    const int ncp = CT * 16, nnp = NT * 16;
    float *dst = part + ((size_t)tap * gridDim.x + blockIdx.x) * (size_t)(nnp * ncp);
    for (int t = threadIdx.x; t < NT * CT * 256; t += 256) {
        const int tile = t / 256, w = t % 256, nt = tile / CT, ct = tile % CT;
        const int m = w / 16, c = w % 16;
        const float s = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        dst[(size_t)(nt * 16 + m) * ncp + ct * 16 + c] = s;
    }
}

// dW[n][tap*cin + c] = sum over chunks (fixed order) of part
__global__ void k_wgrad_reduce(const float *__restrict__ part, int nchunk, int cin, int cout,
                               int nnp, int ncp, int kpad, float *__restrict__ dW) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 27 * cout * cin) return;
    const int c = id % cin, n = (id / cin) % cout, tap = id / (cin * cout);
    double s = 0.0;
    for (int ch = 0; ch < nchunk; ++ch)
        s += part[((size_t)tap * nchunk + ch) * (size_t)(nnp * ncp) + (size_t)n * ncp + c];
    dW[(size_t)n * kpad + tap * cin + c] = (float)s;
}

// ----------------------------------------------- 1-channel output layer --
// The forward pass is k_conv_fwd<16, 16, 1, 2> (MFMA, f32 weights split into
// bf16 high and low parts).
// dIn[u][c] = sum_tap W[tap*16+c] dout[u - off(tap)]   (f32 out)
__global__ __launch_bounds__(256) void k_out_bwd_data(Vol d, const float *__restrict__ dout,
                                                      const float *__restrict__ W,
                                                      float *__restrict__ din) {
    __shared__ float sW[432];
    for (int i = threadIdx.x; i < 432; i += 256) sW[i] = W[i];
    __syncthreads();
    const long long V = d.n();
    const long long u = blockIdx.x * 256ll + threadIdx.x;
    if (u >= V) return;
    const int x = (int)(u % d.nx);
    const long long r = u / d.nx;
    const int y = (int)(r % d.ny), z = (int)(r / d.ny);
    float acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = 0.f;
    for (int tap = 0; tap < 27; ++tap) {
        const int X = x - (tap % 3 - 1), Y = y - ((tap / 3) % 3 - 1), Z = z - (tap / 9 - 1);
        if (X < 0 || X >= d.nx || Y < 0 || Y >= d.ny || Z < 0 || Z >= d.nz) continue;
        const float g = dout[(size_t)X + (size_t)d.nx * ((size_t)Y + (size_t)d.ny * Z)];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] += sW[tap * 16 + c] * g;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) din[(size_t)u * 16 + c] = acc[c];
}

// ----------------------------------------------------- element-wise ops --
// avg pool 2x2x2 (ceil): out[V1][C] from in[V0][C]  (bf16).  One thread per
// coarse voxel and 8-channel chunk (16-byte loads and stores); each channel
// sums its children in (dz, dy, dx) order.
__global__ void k_pool(Vol f, Vol c, int C, const uint16_t *__restrict__ in, uint16_t *__restrict__ out) {
    const int C8 = C / 8;
    const long long total = c.n() * C8;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(id % C8);
        const long long v = id / C8;
        const int x = (int)(v % c.nx);
        const long long r = v / c.nx;
        const int y = (int)(r % c.ny), z = (int)(r / c.ny);
        float s[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = 0.f;
        int cnt = 0;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int X = 2 * x + dx, Y = 2 * y + dy, Z = 2 * z + dz;
                    if (X >= f.nx || Y >= f.ny || Z >= f.nz) continue;
                    const uint4 u = *reinterpret_cast<const uint4 *>(
                        in + ((size_t)X + (size_t)f.nx * ((size_t)Y + (size_t)f.ny * Z)) * C + q * 8);
                    const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        s[2 * k] += bf2f((uint16_t)(w[k] & 0xffffu));
                        s[2 * k + 1] += bf2f((uint16_t)(w[k] >> 16));
                    }
                    ++cnt;
                }
        unsigned o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            o[k] = (unsigned)f2bf(s[2 * k] / cnt) | ((unsigned)f2bf(s[2 * k + 1] / cnt) << 16);
        *reinterpret_cast<uint4 *>(out + v * C + q * 8) = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// unpool (pool backward): dfine[v][ch] += dcoarse[parent][ch] / count(parent)   (f32)
__global__ void k_unpool_add(Vol f, Vol c, int C, const float *__restrict__ dc, float *__restrict__ df) {
    const long long total = f.n() * C;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(id % C);
        const long long v = id / C;
        const int x = (int)(v % f.nx);
        const long long r = v / f.nx;
        const int y = (int)(r % f.ny), z = (int)(r / f.ny);
        const int X = x / 2, Y = y / 2, Z = z / 2;
        const int cnt = (2 * X + 1 < f.nx ? 2 : 1) * (2 * Y + 1 < f.ny ? 2 : 1) * (2 * Z + 1 < f.nz ? 2 : 1);
        df[id] += dc[((size_t)X + (size_t)c.nx * ((size_t)Y + (size_t)c.ny * Z)) * C + ch] / cnt;
    }
}

// d = relu(up(t) + e): t coarse f32 [Vc][C], e fine bf16, d fine bf16.  One
// thread per fine voxel and 8-channel chunk (16-byte loads and stores).
__global__ void k_up_add_relu(Vol f, Vol c, int C, const float *__restrict__ t,
                              const uint16_t *__restrict__ e, uint16_t *__restrict__ d) {
    const int C8 = C / 8;
    const long long total = f.n() * C8;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(id % C8);
        const long long v = id / C8;
        const int x = (int)(v % f.nx);
        const long long r = v / f.nx;
        const int y = (int)(r % f.ny), z = (int)(r / f.ny);
        const float4 *tp = reinterpret_cast<const float4 *>(
            t + ((size_t)(x / 2) + (size_t)c.nx * ((size_t)(y / 2) + (size_t)c.ny * (z / 2))) * C + q * 8);
        const float4 t0 = tp[0], t1 = tp[1];
        const float tv[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
        const uint4 u = *reinterpret_cast<const uint4 *>(e + v * C + q * 8);
        const unsigned w[4] = {u.x, u.y, u.z, u.w};
        unsigned o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float lo = fmaxf(tv[2 * k] + bf2f((uint16_t)(w[k] & 0xffffu)), 0.f);
            const float hi = fmaxf(tv[2 * k + 1] + bf2f((uint16_t)(w[k] >> 16)), 0.f);
            o[k] = (unsigned)f2bf(lo) | ((unsigned)f2bf(hi) << 16);
        }
        *reinterpret_cast<uint4 *>(d + v * C + q * 8) = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// backward of d = relu(up(t) + e): g = dd * (d > 0); de = g (or +=); dt[coarse] = sum children g
__global__ void k_up_add_relu_bwd_fine(long long n, const float *__restrict__ dd,
                                       const uint16_t *__restrict__ d, float *__restrict__ g) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        g[i] = bf2f(d[i]) > 0.f ? dd[i] : 0.f;
}
__global__ void k_sum_children(Vol f, Vol c, int C, const float *__restrict__ g, float *__restrict__ dt) {
    const long long total = c.n() * C;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(id % C);
        const long long v = id / C;
        const int x = (int)(v % c.nx);
        const long long r = v / c.nx;
        const int y = (int)(r % c.ny), z = (int)(r / c.ny);
        float s = 0.f;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int X = 2 * x + dx, Y = 2 * y + dy, Z = 2 * z + dz;
                    if (X >= f.nx || Y >= f.ny || Z >= f.nz) continue;
                    s += g[((size_t)X + (size_t)f.nx * ((size_t)Y + (size_t)f.ny * Z)) * C + ch];
                }
        dt[id] = s;
    }
}

// y_bf16 = x_f32 * (mask ? (m > 0) : 1)
__global__ void k_to_bf16_mask(long long n, const float *__restrict__ x, const uint16_t *__restrict__ m,
                               uint16_t *__restrict__ y) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = f2bf((m && !(bf2f(m[i]) > 0.f)) ? 0.f : x[i]);
}

__global__ void k_add_bias_f32(long long V, int C, const float *__restrict__ b, float *__restrict__ t) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < V * C;
         i += (long long)gridDim.x * blockDim.x)
        t[i] += b[i % C];
}

__global__ void k_add_f32(long long n, const float *__restrict__ a, float *__restrict__ b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] += a[i];
}

// per-channel sums of dy (bias gradient): partials per block then host-side-free final kernel
__global__ __launch_bounds__(256) void k_chan_sum_partial(long long V, int C, const uint16_t *__restrict__ dy,
                                                          const float *__restrict__ dyf,
                                                          double *__restrict__ part) {
    // thread handles channel (threadIdx.x % C) over voxels strided by (256 / C) * gridDim
    const int per = 256 / C;
    const int ch = threadIdx.x % C, lane_v = threadIdx.x / C;
    double s = 0.0;
    if (lane_v < per)
        for (long long v = (long long)blockIdx.x * per + lane_v; v < V; v += (long long)gridDim.x * per)
            s += dyf ? (double)dyf[v * C + ch] : (double)bf2f(dy[v * C + ch]);
    __shared__ double red[256];
    red[threadIdx.x] = (lane_v < per) ? s : 0.0;
    __syncthreads();
    if (threadIdx.x < C) {
        double t = 0.0;
        for (int k = 0; k < per; ++k) t += red[k * C + threadIdx.x];
        part[(size_t)blockIdx.x * C + threadIdx.x] = t;
    }
}
__global__ void k_chan_sum_final(int nb, int C, const double *__restrict__ part, float *__restrict__ out) {
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch >= C) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * C + ch];
    out[ch] = (float)s;
}

// Wt[c][tap' * cout + n] = W[n][tap * cin + c], tap' = 26 - tap (data-gradient weights)
__global__ void k_flip_weights(int cin, int cout, int kpad, int kpadT, const uint16_t *__restrict__ W,
                               uint16_t *__restrict__ Wt) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= cin * kpadT) return;
    const int c = id / kpadT, kk = id % kpadT;
    uint16_t v = 0;
    if (kk < 27 * cout) {
        const int tapT = kk / cout, n = kk % cout, tap = 26 - tapT;
        v = W[(size_t)n * kpad + tap * cin + c];
    }
    Wt[id] = v;
}

// network input: in[v] = (r[v] * scale, aux[v], 0, 0)
template <class T>
__global__ void k_make_input(long long V, const T *__restrict__ r, double scale,
                             const float *__restrict__ aux, uint16_t *__restrict__ in,
                             const double *nrm2 = nullptr) {
    // nrm2: |r|^2 on the device; the scale is then sqrt(V / |r|^2) (0 when unusable)
    if (nrm2) {
        const double q = *nrm2;
        scale = q > 0.0 && isfinite(q) ? sqrt((double)V / q) : 0.0;
    }
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < V;
         v += (long long)gridDim.x * blockDim.x) {
        in[v * 4 + 0] = f2bf((float)((double)r[v] * scale));
        in[v * 4 + 1] = f2bf(aux ? aux[v] : 0.f);
        in[v * 4 + 2] = 0;
        in[v * 4 + 3] = 0;
    }
}

// loss 0.5 |o - t|^2 / tt and dout = (o - t) / tt; partials of the loss
__global__ __launch_bounds__(256) void k_loss(long long V, const float *__restrict__ o,
                                              const float *__restrict__ t, double inv_tt,
                                              float *__restrict__ dout, double *__restrict__ part) {
    double v[1] = {0.0};
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < V; i += (long long)gridDim.x * 256) {
        const double e = (double)o[i] - (double)t[i];
        dout[i] = (float)(e * inv_tt);
        v[0] += 0.5 * e * e * inv_tt;
    }
    block_sum<1, 256>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// Adam on FP32 master weights; padding entries (mask == 0) stay zero.
__global__ void k_adam(long long n, float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2, const uint8_t *__restrict__ mask) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        if (mask && !mask[i]) continue;
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void k_f32_to_bf16(long long n, const float *__restrict__ a, uint16_t *__restrict__ b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] = f2bf(a[i]);
}

template <class T>
__global__ void k_out_scale(long long V, const float *__restrict__ o, double s, T *__restrict__ z,
                            int accumulate) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < V;
         v += (long long)gridDim.x * blockDim.x) {
        const T val = (T)((double)o[v] * s);
        if (accumulate) z[v] += val;
        else z[v] = val;
    }
}

// Energy-norm loss: partials of a = n.r, c = n.(A n), E = e.r with n the
// network output, r and e scaled by sc (A e = r).
template <class T>
__global__ __launch_bounds__(256) void k_energy_dots(long long V, const float *__restrict__ o,
                                                     const float *__restrict__ r,
                                                     const float *__restrict__ e,
                                                     const T *__restrict__ An, double sc,
                                                     double *__restrict__ part) {
    double v[3] = {0, 0, 0};
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < V; i += (long long)gridDim.x * 256) {
        const double ri = (double)r[i] * sc;
        v[0] += (double)o[i] * ri;
        v[1] += (double)o[i] * (double)An[i];
        v[2] += (double)e[i] * sc * ri;
    }
    block_sum<3, 256>(v);
    if (threadIdx.x == 0) {
        part[0 * kRedBlocks + blockIdx.x] = v[0];
        part[1 * kRedBlocks + blockIdx.x] = v[1];
        part[2 * kRedBlocks + blockIdx.x] = v[2];
    }
}

// d loss / d n = coef (r - gamma A n) with coef = -2 a / (c E), gamma = a / c.
template <class T>
__global__ void k_energy_grad(long long V, const float *__restrict__ r, double sc,
                              const T *__restrict__ An, double coef, double gamma,
                              float *__restrict__ dout) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < V;
         i += (long long)gridDim.x * blockDim.x)
        dout[i] = (float)(coef * ((double)r[i] * sc - gamma * (double)An[i]));
}

// ------------------------------------------------- coarse placement --
// With a coarsening factor cf > 1 the network runs on the grid of every cf-th
// DOF node (Nc = (Nf - 1) / cf + 1 per direction).  Its output reaches the DOF
// grid by trilinear interpolation P (fine node g takes the coarse nodes
// around g / cf with weights prod_d (1 - |g_d / cf - i_d|)), and residuals go
// to the coarse grid by the transpose R = P^T, applied one axis at a time: on
// an array laid out [outer][n][inner], out[o][i][s] = sum_t (1 - |t| / cf)
// in[o][cf i + t][s] over |t| < cf.
template <class TI>
__global__ void k_restrict_axis(long long outer, int nf, int nc, long long inner, int cf,
                                const TI *__restrict__ in, float *__restrict__ out) {
    const long long total = outer * nc * inner;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const long long sI = id % inner, t2 = id / inner;
        const int i = (int)(t2 % nc);
        const long long o = t2 / nc;
        // weights (cf - |t|) / cf: integer taps, one scaling per output
        const int t0 = max(-(cf - 1), -cf * i), t1 = min(cf - 1, nf - 1 - cf * i);
        double acc = 0.0;
        for (int t = t0; t <= t1; ++t)
            acc += (double)(cf - abs(t)) * (double)in[(o * nf + cf * i + t) * inner + sI];
        out[id] = (float)(acc * (1.0 / cf));
    }
}

// The x pass of R: out[line][i] = sum_t (1 - |t|/cf) in[line][cf i + t]
// (|t| < cf, inside the line).  A workgroup stages `lines` consecutive
// x-lines in LDS with coalesced loads, then each thread sums one output's
// window from LDS.  A staged line is stored de-interleaved, value g at
// (g mod cf) nc + g / cf, so that one tap's reads by consecutive outputs are
// consecutive words (in line order they were cf words apart: for cf = 8, 16
// lanes per bank).
template <class TI>
__global__ __launch_bounds__(256) void k_restrict_x(long long nlines, int nf, int nc, int cf, int lines,
                                                    const TI *__restrict__ in, float *__restrict__ out) {
    extern __shared__ double sline[];  // lines * cf * nc values
    const long long l0 = (long long)blockIdx.x * lines;
    const int nl = (int)min((long long)lines, nlines - l0);
    const TI *src = in + l0 * nf;
    const int stride = cf * nc;
    // k / nf and g / cf as floor((k + 1/2) / d) in float: exact for k <= 4096,
    // any d (the relative error, ~2^-23, times (k + 1/2) / d stays below the
    // distance 0.5 / d from the nearest integer)
    const float icf = 1.0f / (float)cf, inf = 1.0f / (float)nf;
    // every load of the thread is issued before any is used (lines nf <= 4096,
    // at most 16 per thread): in a loop they would wait one after another
    const int cnt = nl * nf;
    TI v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int k = threadIdx.x + 256 * u;
        v[u] = k < cnt ? src[k] : TI(0);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int k = threadIdx.x + 256 * u;
        if (k < cnt) {
            const int l = (int)(((float)k + 0.5f) * inf), g = k - l * nf;
            const int q = (int)(((float)g + 0.5f) * icf), r = g - q * cf;
            sline[l * stride + r * nc + q] = (double)v[u];
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nl * nc; k += 256) {
        const int l = k / nc, i = k - l * nc;
        const double *line = sline + l * stride;
        // weights (cf - |t|) / cf: integer taps over the window inside the
        // line, one scaling per output; g = cf i + t sits at (g mod cf) nc + g / cf
        const int t0 = max(-(cf - 1), -cf * i), t1 = min(cf - 1, nf - 1 - cf * i);
        double acc = 0.0;
        for (int t = t0; t <= t1; ++t)
            acc += (double)(cf - abs(t)) * line[t >= 0 ? t * nc + i : (cf + t) * nc + i - 1];
        out[(l0 + l) * nc + i] = (float)(acc * (1.0 / cf));
    }
}

// z[g] = scale * (P oc)[g] (trilinear interpolation of the coarse values)
// (bmask: 0 on the boundary of the DOF grid, where every DOF is essential).
// A workgroup takes lp consecutive fine x-lines (gy, gz) (prolong_lines: up
// to 8, as LDS allows): each line's coarse row, interpolated in y and z, is
// formed once in LDS; each fine node then interpolates along x.
template <class T>
__global__ __launch_bounds__(256) void k_prolong(Vol f, Vol c, int cf, int lp, const float *__restrict__ oc,
                                                 double scale, int bmask, T *__restrict__ z,
                                                 const double *nrm2 = nullptr) {
    // nrm2: |R r|^2 on the device; the scale is then sqrt(|R r|^2 / Vc)
    if (nrm2) {
        const double q = *nrm2;
        scale = q > 0.0 && isfinite(q) ? sqrt(q / (double)c.n()) : 0.0;
    }
    extern __shared__ float brow[];  // lp * c.nx values
    const unsigned nlines = (unsigned)f.ny * (unsigned)f.nz;
    const unsigned line0 = blockIdx.x * (unsigned)lp;
    const float icf = 1.0f / (float)cf, inx = 1.0f / (float)c.nx;
    // the lp c.nx coarse-row entries, in rounds of 8 per thread whose 4 loads
    // each are issued before any is used
    const int cnt = lp * c.nx;
    for (int base = 0; base < cnt; base += 8 * 256) {
        float a0[8], a1[8], a2[8], a3[8], tyv[8], tzv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = base + threadIdx.x + 256 * u;
            a0[u] = a1[u] = a2[u] = a3[u] = tyv[u] = tzv[u] = 0.0f;
            if (k >= cnt) continue;
            // l = k / c.nx in float: exact for k <= 8192 (as for gx / cf below)
            const int l = (int)(((float)k + 0.5f) * inx), i = k - l * c.nx;
            const unsigned line = line0 + l;
            if (line >= nlines) continue;
            const int gy = (int)(line % (unsigned)f.ny), gz = (int)(line / (unsigned)f.ny);
            // lower coarse node of the cell holding g (the last cell for g = Nf - 1);
            // the upper one is clamped for a one-node axis, where t = 0
            const int iy = min(gy / cf, max(c.ny - 2, 0)), iz = min(gz / cf, max(c.nz - 2, 0));
            const int iy1 = min(iy + 1, c.ny - 1), iz1 = min(iz + 1, c.nz - 1);
            tyv[u] = (float)(gy - cf * iy) * icf;
            tzv[u] = (float)(gz - cf * iz) * icf;
            a0[u] = oc[((size_t)iz * c.ny + iy) * c.nx + i];
            a1[u] = oc[((size_t)iz * c.ny + iy1) * c.nx + i];
            a2[u] = oc[((size_t)iz1 * c.ny + iy) * c.nx + i];
            a3[u] = oc[((size_t)iz1 * c.ny + iy1) * c.nx + i];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = base + threadIdx.x + 256 * u;
            if (k < cnt) {
                const float b0 = a0[u] + tyv[u] * (a1[u] - a0[u]), b1 = a2[u] + tyv[u] * (a3[u] - a2[u]);
                brow[k] = b0 + tzv[u] * (b1 - b0);
            }
        }
    }
    __syncthreads();
    for (int l = 0; l < lp; ++l) {
        const unsigned line = line0 + l;
        if (line >= nlines) break;
        const int gy = (int)(line % (unsigned)f.ny), gz = (int)(line / (unsigned)f.ny);
        T *zl = z + (size_t)line * f.nx;
        const float *br = brow + l * c.nx;
        const bool zero_line = bmask && (gy == 0 || gz == 0 || gy == f.ny - 1 || gz == f.nz - 1);
        for (int gx = threadIdx.x; gx < f.nx; gx += 256) {
            if (zero_line || (bmask && (gx == 0 || gx == f.nx - 1))) {
                zl[gx] = T(0);
                continue;
            }
            // gx / cf as floor((gx + 1/2) / cf) in float: exact for gx <= 4096, any cf
            // (the value's relative error, ~2^-23, times (gx + 1/2) / cf stays below
            // its distance 0.5 / cf from the nearest integer)
            const int ix = min((int)(((float)gx + 0.5f) * icf), max(c.nx - 2, 0)), ix1 = min(ix + 1, c.nx - 1);
            const float tx = (float)(gx - cf * ix) * icf;
            zl[gx] = (T)((double)(br[ix] + tx * (br[ix1] - br[ix])) * scale);
        }
    }
}

// fine lines per k_prolong workgroup: up to 8, with at most 32 KB of LDS
static inline int prolong_lines(const Vol &c) { return std::max(1, std::min(8, 8192 / std::max(c.nx, 1))); }
static inline unsigned prolong_blocks(const Vol &f, const Vol &c) {
    const int lp = prolong_lines(c);
    return (unsigned)(((long long)f.ny * f.nz + lp - 1) / lp);
}

// ac[i] = af[cf i] (the coarse nodes are DOF nodes)
__global__ void k_inject(Vol f, Vol c, int cf, const float *__restrict__ af, float *__restrict__ ac) {
    const long long V = c.n();
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < V;
         id += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(id % c.nx);
        const long long r = id / c.nx;
        const int iy = (int)(r % c.ny), iz = (int)(r / c.ny);
        ac[id] = af[((long long)(cf * iz) * f.ny + cf * iy) * f.nx + cf * ix];
    }
}

// ------------------------------------------------------------ network --
struct Layer {
    int cin, cout, kpad;
    Vol d;  // grid the conv runs on
    size_t woff, boff, wn;  // offsets of W [cout][kpad] and bias [cout] in the parameter array
};

}  // namespace femgpu

// Opaque C-ABI handle.
struct femgpu_unet {
    femgpu_ctx *ctx = nullptr;
    femgpu::Vol d0{}, d1{}, d2{};
    femgpu::Layer L[6];
    size_t nparam = 0;
    femgpu::DevBuf params, grads, adam_m, adam_v, mask, wbf, wflip, aux;
    femgpu::DevBuf wout;  // output layer as a 16-column bf16 B matrix: rows 0/1 = high/low parts
    // activations (bf16) and gradients (f32)
    femgpu::DevBuf in, e0, p0, e1, p1, e2, t1, d1b, t0, d0b, out;
    femgpu::DevBuf gout, gd0, ge0, gt0, gd1, ge1, gt1, ge2, gp1, gp0, gtmp, gbf, part, dpart;
    long long step = 0;
    uint64_t seed = 0;
    // coarse placement (cf > 1): the DOF grid, its restriction temporaries,
    // the coarse residual and the prolongated output (f32, DOF grid)
    int cf = 1;
    int bmask = 0;  // zero the correction on the DOF grid's boundary (essential DOFs)
    femgpu::Vol fine{};
    femgpu::DevBuf rx, rxy, rc, nf;
};

namespace femgpu {

static void unet_layout(femgpu_unet *u) {
    const int cins[6] = {4, 16, 32, 64, 32, 16}, couts[6] = {16, 32, 64, 32, 16, 1};
    const Vol ds[6] = {u->d0, u->d1, u->d2, u->d2, u->d1, u->d0};
    size_t off = 0;
    for (int l = 0; l < 6; ++l) {
        Layer &L = u->L[l];
        L.cin = cins[l]; L.cout = couts[l]; L.d = ds[l];
        L.kpad = l == 5 ? 432 : kpad_of(L.cin);
        L.wn = (size_t)L.cout * L.kpad;
        L.woff = off; off += L.wn;
        L.boff = off; off += L.cout;
    }
    u->nparam = off;
}

static void unet_alloc(femgpu_unet *u) {
    const long long V0 = u->d0.n(), V1 = u->d1.n(), V2 = u->d2.n();
    auto bf = [](DevBuf &b, long long n) { b.alloc((size_t)n * 2); };
    auto f = [](DevBuf &b, long long n) { b.alloc((size_t)n * 4); };
    f(u->params, u->nparam); f(u->grads, u->nparam); f(u->adam_m, u->nparam); f(u->adam_v, u->nparam);
    u->mask.alloc(u->nparam);
    bf(u->wbf, u->nparam);
    bf(u->wout, 16 * 432);
    size_t maxflip = 0;
    for (int l = 1; l < 5; ++l) maxflip = std::max(maxflip, (size_t)u->L[l].cin * kpad_of(u->L[l].cout));
    bf(u->wflip, maxflip);
    f(u->aux, V0);
    bf(u->in, V0 * 4); bf(u->e0, V0 * 16); bf(u->p0, V1 * 16); bf(u->e1, V1 * 32); bf(u->p1, V2 * 32);
    bf(u->e2, V2 * 64); bf(u->d1b, V1 * 32); bf(u->d0b, V0 * 16);
    f(u->t1, V2 * 32); f(u->t0, V1 * 16); f(u->out, V0);
    f(u->gout, V0); f(u->gd0, V0 * 16); f(u->ge0, V0 * 16); f(u->gt0, V1 * 16); f(u->gd1, V1 * 32);
    f(u->ge1, V1 * 32); f(u->gt1, V2 * 32); f(u->ge2, V2 * 64); f(u->gp1, V2 * 32); f(u->gp0, V1 * 16);
    f(u->gtmp, std::max(V0 * 16, V1 * 32));
    bf(u->gbf, std::max({V0 * 16, V1 * 32, V2 * 64}));
    u->dpart.alloc(sizeof(double) * 1024 * 64);
    u->part.alloc(sizeof(float) * 27 * 256 * 64 * 64);
}

static int wgrad_chunks(long long V) {
    long long c = (V + 4095) / 4096;
    if (c > 256) c = 256;
    return (int)(c < 1 ? 1 : c);
}

// false when FEMGPU_CONV_K16=1 (read once): one tap per MFMA, for comparison
static bool conv_double_k() {
    static const bool dk = !(std::getenv("FEMGPU_CONV_K16") && std::getenv("FEMGPU_CONV_K16")[0] == '1');
    return dk;
}

static void conv_fwd(femgpu_unet *u, int l, const uint16_t *in, const uint16_t *W, const float *bias,
                     const uint16_t *skip, void *out, bool relu, bool outf32, int cout_override = 0) {
    const Layer &L = u->L[l];
    hipStream_t s = u->ctx->stream;
    const int cout = cout_override ? cout_override : L.cout;
    const int cin = cout_override ? L.cout : L.cin;  // data-gradient: channels swap
    const int kpad = cout_override ? kpad_of(L.cout) : L.kpad;
    const long long V = L.d.n();
    const int ntx = (L.d.nx + kBX - 1) / kBX;
    const int r = relu ? 1 : 0;
    // split the output channels over workgroups at the coarse levels (few tiles)
    const long long tiles = (long long)ntx * ((L.d.ny + kBY - 1) / kBY) * ((L.d.nz + kBZ - 1) / kBZ);
    const bool split = tiles < 2048;
    const bool dk = conv_double_k();
    (void)V;
#define CF(CI, CO, NB, OF)                                                                                  \
    do {                                                                                                    \
        const dim3 grid((unsigned)(ntx * (CO / (16 * NB))), (unsigned)((L.d.ny + kBY - 1) / kBY),           \
                        (unsigned)((L.d.nz + kBZ - 1) / kBZ));                                              \
        if (dk) hipLaunchKernelGGL((k_conv_fwd<CI, CO, NB, OF, true>), grid, dim3(256), 0, s, L.d, kpad, in, W, bias, skip, out, r); \
        else hipLaunchKernelGGL((k_conv_fwd<CI, CO, NB, OF, false>), grid, dim3(256), 0, s, L.d, kpad, in, W, bias, skip, out, r); \
    } while (0)
#define CM(CI, CO, OF)                        \
    do {                                      \
        if (split) CF(CI, CO, 1, OF);         \
        else CF(CI, CO, CO / 16, OF);         \
    } while (0)
    if (!outf32 && cin == 4 && cout == 16) CM(4, 16, 0);
    else if (!outf32 && cin == 16 && cout == 32) CM(16, 32, 0);
    else if (!outf32 && cin == 32 && cout == 64) CM(32, 64, 0);
    else if (outf32 && cin == 64 && cout == 32) CM(64, 32, 1);
    else if (outf32 && cin == 32 && cout == 16) CM(32, 16, 1);
    else if (outf32 && cin == 16 && cout == 32) CM(16, 32, 1);
    else if (outf32 && cin == 32 && cout == 64) CM(32, 64, 1);
    else throw Error(FEMGPU_ERR_UNSUPPORTED, "conv_fwd: unsupported channel pair");
#undef CM
#undef CF
}

// dW (f32, into grads at layer offset) from input x (bf16 [V][cin]) and dy (bf16 [V][cout])
static void conv_wgrad(femgpu_unet *u, int l, const uint16_t *x, const uint16_t *dy) {
    const Layer &L = u->L[l];
    hipStream_t s = u->ctx->stream;
    const long long V = L.d.n();
    const int nch = wgrad_chunks(V);
    const long long chunk = (V + nch - 1) / nch;
    const int NT = (L.cout + 15) / 16, CT = (L.cin + 15) / 16;
    float *part = u->part.as<float>();
    const dim3 grid((unsigned)nch, 27), block(256);
#define WG(A, B) hipLaunchKernelGGL((k_conv_wgrad<A, B>), grid, block, 0, s, L.d, L.cin, L.cout, x, dy, part, chunk)
    if (NT == 1 && CT == 1) WG(1, 1);
    else if (NT == 2 && CT == 1) WG(2, 1);
    else if (NT == 4 && CT == 2) WG(4, 2);
    else if (NT == 2 && CT == 4) WG(2, 4);
    else if (NT == 1 && CT == 2) WG(1, 2);
    else throw Error(FEMGPU_ERR_UNSUPPORTED, "conv_wgrad: unsupported channel tiling");
#undef WG
    float *dW = u->grads.as<float>() + L.woff;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(g1(27ll * L.cout * L.cin)), dim3(256), 0, s, part, nch, L.cin,
                       L.cout, NT * 16, CT * 16, L.kpad, dW);
}

static void chan_sum(femgpu_unet *u, long long V, int C, const uint16_t *dy, const float *dyf, float *out) {
    hipStream_t s = u->ctx->stream;
    const int nb = 512;
    hipLaunchKernelGGL(k_chan_sum_partial, dim3(nb), dim3(256), 0, s, V, C, dy, dyf, u->dpart.as<double>());
    hipLaunchKernelGGL(k_chan_sum_final, dim3(1), dim3(64), 0, s, nb, C, (const double *)u->dpart.as<double>(), out);
}

static void sync_bf16(femgpu_unet *u) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(g1((long long)u->nparam)), dim3(256), 0, u->ctx->stream,
                       (long long)u->nparam, (const float *)u->params.as<float>(), u->wbf.as<uint16_t>());
    hipLaunchKernelGGL(k_split_out_weights, dim3(g1(16 * 432)), dim3(256), 0, u->ctx->stream,
                       (const float *)(u->params.as<float>() + u->L[5].woff), u->wout.as<uint16_t>());
}

static const uint16_t *Wb(femgpu_unet *u, int l) { return u->wbf.as<uint16_t>() + u->L[l].woff; }
static const float *Bf(femgpu_unet *u, int l) { return u->params.as<float>() + u->L[l].boff; }

// forward from u->in; leaves every activation for the backward pass; out f32 [V0]
static void forward(femgpu_unet *u) {
    hipStream_t s = u->ctx->stream;
    const Vol D0 = u->d0, D1 = u->d1, D2 = u->d2;
    conv_fwd(u, 0, u->in.as<uint16_t>(), Wb(u, 0), Bf(u, 0), nullptr, u->e0.ptr, true, false);
    hipLaunchKernelGGL(k_pool, dim3(g1(D1.n() * 2)), dim3(256), 0, s, D0, D1, 16, (const uint16_t *)u->e0.as<uint16_t>(), u->p0.as<uint16_t>());
    conv_fwd(u, 1, u->p0.as<uint16_t>(), Wb(u, 1), Bf(u, 1), nullptr, u->e1.ptr, true, false);
    hipLaunchKernelGGL(k_pool, dim3(g1(D2.n() * 4)), dim3(256), 0, s, D1, D2, 32, (const uint16_t *)u->e1.as<uint16_t>(), u->p1.as<uint16_t>());
    conv_fwd(u, 2, u->p1.as<uint16_t>(), Wb(u, 2), Bf(u, 2), nullptr, u->e2.ptr, true, false);
    // decoder: t1 = conv(e2) + b (f32), d1 = relu(up(t1) + e1)
    conv_fwd(u, 3, u->e2.as<uint16_t>(), Wb(u, 3), nullptr, nullptr, u->t1.ptr, false, true);
    hipLaunchKernelGGL(k_add_bias_f32, dim3(g1(D2.n() * 32)), dim3(256), 0, s, D2.n(), 32, Bf(u, 3), u->t1.as<float>());
    hipLaunchKernelGGL(k_up_add_relu, dim3(g1(D1.n() * 4)), dim3(256), 0, s, D1, D2, 32, (const float *)u->t1.as<float>(),
                       (const uint16_t *)u->e1.as<uint16_t>(), u->d1b.as<uint16_t>());
    conv_fwd(u, 4, u->d1b.as<uint16_t>(), Wb(u, 4), nullptr, nullptr, u->t0.ptr, false, true);
    hipLaunchKernelGGL(k_add_bias_f32, dim3(g1(D1.n() * 16)), dim3(256), 0, s, D1.n(), 16, Bf(u, 4), u->t0.as<float>());
    hipLaunchKernelGGL(k_up_add_relu, dim3(g1(D0.n() * 2)), dim3(256), 0, s, D0, D1, 16, (const float *)u->t0.as<float>(),
                       (const uint16_t *)u->e0.as<uint16_t>(), u->d0b.as<uint16_t>());
    const dim3 og((unsigned)((D0.nx + kBX - 1) / kBX), (unsigned)((D0.ny + kBY - 1) / kBY),
                  (unsigned)((D0.nz + kBZ - 1) / kBZ));
    if (conv_double_k())
        hipLaunchKernelGGL((k_conv_fwd<16, 16, 1, 2, true>), og, dim3(256), 0, s, D0, 432,
                           (const uint16_t *)u->d0b.as<uint16_t>(), (const uint16_t *)u->wout.as<uint16_t>(), Bf(u, 5),
                           (const uint16_t *)nullptr, (void *)u->out.as<float>(), 0);
    else
        hipLaunchKernelGGL((k_conv_fwd<16, 16, 1, 2, false>), og, dim3(256), 0, s, D0, 432,
                           (const uint16_t *)u->d0b.as<uint16_t>(), (const uint16_t *)u->wout.as<uint16_t>(), Bf(u, 5),
                           (const uint16_t *)nullptr, (void *)u->out.as<float>(), 0);
}

// backward from u->gout (dL/dout, f32 [V0]); fills u->grads
static void backward(femgpu_unet *u) {
    hipStream_t s = u->ctx->stream;
    const Vol D0 = u->d0, D1 = u->d1, D2 = u->d2;
    const long long V0 = D0.n(), V1 = D1.n(), V2 = D2.n();
    float *G = u->grads.as<float>();
    uint16_t *gbf = u->gbf.as<uint16_t>();
    FG_HIP(hipMemsetAsync(G, 0, u->nparam * 4, s));
    // L5 (16 -> 1, VALU): dW5 via the MFMA weight-gradient kernel on bf16 dout
    hipLaunchKernelGGL(k_to_bf16_mask, dim3(g1(V0)), dim3(256), 0, s, V0, (const float *)u->gout.as<float>(), nullptr, gbf);
    conv_wgrad(u, 5, u->d0b.as<uint16_t>(), gbf);
    chan_sum(u, V0, 1, nullptr, u->gout.as<float>(), G + u->L[5].boff);
    hipLaunchKernelGGL(k_out_bwd_data, dim3(g1(V0)), dim3(256), 0, s, D0, (const float *)u->gout.as<float>(),
                       (const float *)(u->params.as<float>() + u->L[5].woff), u->gd0.as<float>());
    // d0 = relu(up(t0) + e0): g = gd0 * (d0 > 0) -> ge0 (skip) and gt0 (sum of children)
    hipLaunchKernelGGL(k_up_add_relu_bwd_fine, dim3(g1(V0 * 16)), dim3(256), 0, s, V0 * 16, (const float *)u->gd0.as<float>(),
                       (const uint16_t *)u->d0b.as<uint16_t>(), u->ge0.as<float>());
    hipLaunchKernelGGL(k_sum_children, dim3(g1(V1 * 16)), dim3(256), 0, s, D0, D1, 16, (const float *)u->ge0.as<float>(), u->gt0.as<float>());
    // L4: t0 = conv_{32->16}(d1) + b4
    hipLaunchKernelGGL(k_to_bf16_mask, dim3(g1(V1 * 16)), dim3(256), 0, s, V1 * 16, (const float *)u->gt0.as<float>(), nullptr, gbf);
    conv_wgrad(u, 4, u->d1b.as<uint16_t>(), gbf);
    chan_sum(u, V1, 16, gbf, nullptr, G + u->L[4].boff);
    hipLaunchKernelGGL(k_flip_weights, dim3(g1(32ll * kpad_of(16))), dim3(256), 0, s, 32, 16, u->L[4].kpad, kpad_of(16), Wb(u, 4), u->wflip.as<uint16_t>());
    conv_fwd(u, 4, gbf, u->wflip.as<uint16_t>(), nullptr, nullptr, u->gd1.ptr, false, true, 32);
    // d1 = relu(up(t1) + e1)
    hipLaunchKernelGGL(k_up_add_relu_bwd_fine, dim3(g1(V1 * 32)), dim3(256), 0, s, V1 * 32, (const float *)u->gd1.as<float>(),
                       (const uint16_t *)u->d1b.as<uint16_t>(), u->ge1.as<float>());
    hipLaunchKernelGGL(k_sum_children, dim3(g1(V2 * 32)), dim3(256), 0, s, D1, D2, 32, (const float *)u->ge1.as<float>(), u->gt1.as<float>());
    // L3: t1 = conv_{64->32}(e2) + b3
    hipLaunchKernelGGL(k_to_bf16_mask, dim3(g1(V2 * 32)), dim3(256), 0, s, V2 * 32, (const float *)u->gt1.as<float>(), nullptr, gbf);
    conv_wgrad(u, 3, u->e2.as<uint16_t>(), gbf);
    chan_sum(u, V2, 32, gbf, nullptr, G + u->L[3].boff);
    hipLaunchKernelGGL(k_flip_weights, dim3(g1(64ll * kpad_of(32))), dim3(256), 0, s, 64, 32, u->L[3].kpad, kpad_of(32), Wb(u, 3), u->wflip.as<uint16_t>());
    conv_fwd(u, 3, gbf, u->wflip.as<uint16_t>(), nullptr, nullptr, u->ge2.ptr, false, true, 64);
    // L2: e2 = relu(conv_{32->64}(p1) + b2)
    hipLaunchKernelGGL(k_to_bf16_mask, dim3(g1(V2 * 64)), dim3(256), 0, s, V2 * 64, (const float *)u->ge2.as<float>(),
                       (const uint16_t *)u->e2.as<uint16_t>(), gbf);
    conv_wgrad(u, 2, u->p1.as<uint16_t>(), gbf);
    chan_sum(u, V2, 64, gbf, nullptr, G + u->L[2].boff);
    hipLaunchKernelGGL(k_flip_weights, dim3(g1(32ll * kpad_of(64))), dim3(256), 0, s, 32, 64, u->L[2].kpad, kpad_of(64), Wb(u, 2), u->wflip.as<uint16_t>());
    conv_fwd(u, 2, gbf, u->wflip.as<uint16_t>(), nullptr, nullptr, u->gp1.ptr, false, true, 32);
    hipLaunchKernelGGL(k_unpool_add, dim3(g1(V1 * 32)), dim3(256), 0, s, D1, D2, 32, (const float *)u->gp1.as<float>(), u->ge1.as<float>());
    // L1: e1 = relu(conv_{16->32}(p0) + b1)
    hipLaunchKernelGGL(k_to_bf16_mask, dim3(g1(V1 * 32)), dim3(256), 0, s, V1 * 32, (const float *)u->ge1.as<float>(),
                       (const uint16_t *)u->e1.as<uint16_t>(), gbf);
    conv_wgrad(u, 1, u->p0.as<uint16_t>(), gbf);
    chan_sum(u, V1, 32, gbf, nullptr, G + u->L[1].boff);
    hipLaunchKernelGGL(k_flip_weights, dim3(g1(16ll * kpad_of(32))), dim3(256), 0, s, 16, 32, u->L[1].kpad, kpad_of(32), Wb(u, 1), u->wflip.as<uint16_t>());
    conv_fwd(u, 1, gbf, u->wflip.as<uint16_t>(), nullptr, nullptr, u->gp0.ptr, false, true, 16);
    hipLaunchKernelGGL(k_unpool_add, dim3(g1(V0 * 16)), dim3(256), 0, s, D0, D1, 16, (const float *)u->gp0.as<float>(), u->ge0.as<float>());
    // L0: e0 = relu(conv_{4->16}(in) + b0)
    hipLaunchKernelGGL(k_to_bf16_mask, dim3(g1(V0 * 16)), dim3(256), 0, s, V0 * 16, (const float *)u->ge0.as<float>(),
                       (const uint16_t *)u->e0.as<uint16_t>(), gbf);
    conv_wgrad(u, 0, u->in.as<uint16_t>(), gbf);
    chan_sum(u, V0, 16, gbf, nullptr, G + u->L[0].boff);
}

}  // namespace femgpu

namespace femgpu {

// R v (v on the DOF grid, type T) into u->rc (f32, network grid): three
// one-axis passes x, y, z.
template <class T>
static void restrict_to_coarse(femgpu_unet *u, const T *v) {
    hipStream_t s = u->ctx->stream;
    const Vol f = u->fine, c = u->d0;
    const long long o1 = (long long)f.nz * f.ny;
    const int lines = std::max(1, std::min(4, 4096 / f.nx));  // at most 32 KB of LDS (f.nx <= 4096)
    FG_REQUIRE(f.nx <= 4096, FEMGPU_ERR_UNSUPPORTED, "unet: DOF grid rows longer than 4096");
    hipLaunchKernelGGL(k_restrict_x<T>, dim3((unsigned)((o1 + lines - 1) / lines)), dim3(256),
                       sizeof(double) * lines * u->cf * c.nx, s, o1, f.nx, c.nx, u->cf, lines, v,
                       u->rx.as<float>());
    const long long n2 = (long long)f.nz * c.ny * c.nx;
    hipLaunchKernelGGL(k_restrict_axis<float>, dim3(g1(n2)), dim3(256), 0, s, (long long)f.nz, f.ny, c.ny,
                       (long long)c.nx, u->cf, (const float *)u->rx.as<float>(), u->rxy.as<float>());
    hipLaunchKernelGGL(k_restrict_axis<float>, dim3(g1(c.n())), dim3(256), 0, s, 1ll, f.nz, c.nz,
                       (long long)c.ny * c.nx, u->cf, (const float *)u->rxy.as<float>(), u->rc.as<float>());
}

// z = |r|/sqrt(V) N(r sqrt(V)/|r|) written as type T (accumulate: z +=).  Used by
// the flexible-CG preconditioner; returns false when |r| == 0.
template <class T>
static bool unet_apply_t(femgpu_unet *u, const T *r, T *z, double rnorm) {
    hipStream_t s = u->ctx->stream;
    const long long V = u->d0.n();
    if (!(rnorm > 0.0) || !std::isfinite(rnorm)) return false;
    if (u->cf > 1) {
        // z = |Rr|/sqrt(Vc) P N(Rr sqrt(Vc)/|Rr|)
        restrict_to_coarse<T>(u, r);
        const double rc = std::sqrt(dot_host(u->ctx, u->rc.ptr, u->rc.ptr, V, FEMGPU_F32));
        if (!(rc > 0.0) || !std::isfinite(rc)) return false;
        const double scc = std::sqrt((double)V) / rc;
        hipLaunchKernelGGL(k_make_input<float>, dim3(g1(V)), dim3(256), 0, s, V, (const float *)u->rc.as<float>(),
                           scc, (const float *)u->aux.as<float>(), u->in.as<uint16_t>());
        forward(u);
        hipLaunchKernelGGL(k_prolong<T>, dim3(prolong_blocks(u->fine, u->d0)), dim3(256),
                           sizeof(float) * prolong_lines(u->d0) * u->d0.nx, s, u->fine, u->d0, u->cf, prolong_lines(u->d0), (const float *)u->out.as<float>(),
                           1.0 / scc, u->bmask, z);
        return true;
    }
    const double sc = std::sqrt((double)V) / rnorm;
    hipLaunchKernelGGL(k_make_input<T>, dim3(g1(V)), dim3(256), 0, s, V, r, sc, (const float *)u->aux.as<float>(),
                       u->in.as<uint16_t>());
    forward(u);
    hipLaunchKernelGGL(k_out_scale<T>, dim3(g1(V)), dim3(256), 0, s, V, (const float *)u->out.as<float>(), 1.0 / sc, z, 0);
    return true;
}

// The coarse placement's correction with its scale on the device: *nrm2 =
// |R r|^2 is computed here and read by the input and prolongation kernels, so
// nothing is read back (the flexible CG's per-iteration path).  An unusable
// |R r| gives z = 0.
template <class T>
static void unet_apply_coarse_t(femgpu_unet *u, const T *r, T *z, double *nrm2) {
    hipStream_t s = u->ctx->stream;
    const long long V = u->d0.n();
    restrict_to_coarse<T>(u, r);
    dot<float>(u->ctx, (const float *)u->rc.as<float>(), (const float *)u->rc.as<float>(), (size_t)V, nrm2);
    hipLaunchKernelGGL(k_make_input<float>, dim3(g1(V)), dim3(256), 0, s, V, (const float *)u->rc.as<float>(), 0.0,
                       (const float *)u->aux.as<float>(), u->in.as<uint16_t>(), (const double *)nrm2);
    forward(u);
    hipLaunchKernelGGL(k_prolong<T>, dim3(prolong_blocks(u->fine, u->d0)), dim3(256),
                       sizeof(float) * prolong_lines(u->d0) * u->d0.nx, s, u->fine, u->d0, u->cf, prolong_lines(u->d0), (const float *)u->out.as<float>(), 0.0,
                       u->bmask, z, (const double *)nrm2);
}

void unet_apply_coarse_dev(femgpu_unet *u, const void *r, void *z, femgpu_dtype dt, double *nrm2) {
    FG_REQUIRE(u->cf > 1, FEMGPU_ERR_INVALID_ARG, "unet: device-scaled apply needs the coarse placement");
    if (dt == FEMGPU_F64) unet_apply_coarse_t<double>(u, (const double *)r, (double *)z, nrm2);
    else unet_apply_coarse_t<float>(u, (const float *)r, (float *)z, nrm2);
}

bool unet_apply_dev(femgpu_unet *u, const void *r, void *z, femgpu_dtype dt, double rnorm) {
    if (dt == FEMGPU_F64) return unet_apply_t<double>(u, (const double *)r, (double *)z, rnorm);
    return unet_apply_t<float>(u, (const float *)r, (float *)z, rnorm);
}

// voxels of the grid the network's input and output live on (the DOF grid)
long long unet_voxels(const femgpu_unet *u) { return u->cf > 1 ? u->fine.n() : u->d0.n(); }
bool unet_coarse(const femgpu_unet *u) { return u->cf > 1; }
femgpu_ctx *unet_ctx(const femgpu_unet *u) { return u->ctx; }

}  // namespace femgpu

using namespace femgpu;

extern "C" {

static int unet_create(femgpu_ctx *c, int nx, int ny, int nz, int cf, uint64_t seed, femgpu_unet **out) {
    return guard([&] {
        FG_REQUIRE(c && out && nx > 0 && ny > 0 && nz > 0 && cf >= 1, FEMGPU_ERR_INVALID_ARG,
                   "unet_create: bad arguments");
        FG_REQUIRE((nx - 1) % cf == 0 && (ny - 1) % cf == 0 && (nz - 1) % cf == 0, FEMGPU_ERR_INVALID_ARG,
                   "unet_create: grid size - 1 must be a multiple of the coarsening factor");
        FG_REQUIRE((long long)nx * ny * nz < (1ll << 31), FEMGPU_ERR_INVALID_ARG, "unet_create: grid too large");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        auto *u = new femgpu_unet();
        try {
            u->ctx = c;
            u->seed = seed;
            u->cf = cf;
            u->fine = Vol{nx, ny, nz};
            u->d0 = cf > 1 ? Vol{(nx - 1) / cf + 1, (ny - 1) / cf + 1, (nz - 1) / cf + 1} : Vol{nx, ny, nz};
            if (cf > 1) {
                u->rx.alloc((size_t)nz * ny * u->d0.nx * 4);
                u->rxy.alloc((size_t)nz * u->d0.ny * u->d0.nx * 4);
                u->rc.alloc((size_t)u->d0.n() * 4);
                u->nf.alloc((size_t)u->fine.n() * 4);
            }
            u->d1 = half(u->d0);
            u->d2 = half(u->d1);
            unet_layout(u);
            unet_alloc(u);
            // He initialisation (normal, variance 2 / fan_in) from Philox; biases 0;
            // the output layer starts small.
            std::vector<float> p(u->nparam, 0.f);
            std::vector<uint8_t> mk(u->nparam, 0);
            uint64_t idx = 0;
            for (int l = 0; l < 6; ++l) {
                const Layer &L = u->L[l];
                const int real_cin = l == 0 ? 2 : L.cin;
                const double sd = std::sqrt(2.0 / (27.0 * real_cin)) * (l == 5 ? 0.1 : 1.0);
                for (int n = 0; n < L.cout; ++n)
                    for (int k = 0; k < 27 * L.cin; ++k) {
                        const size_t at = L.woff + (size_t)n * L.kpad + k;
                        mk[at] = 1;
                        const double u1 = philox_u01(seed, 30, idx), u2 = philox_u01(seed, 31, idx);
                        ++idx;
                        p[at] = (float)(sd * std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2));
                    }
                for (int n = 0; n < L.cout; ++n) mk[L.boff + n] = 1;
            }
            hipStream_t s = c->stream;
            FG_HIP(hipMemcpyAsync(u->params.ptr, p.data(), u->nparam * 4, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemcpyAsync(u->mask.ptr, mk.data(), u->nparam, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemsetAsync(u->adam_m.ptr, 0, u->nparam * 4, s));
            FG_HIP(hipMemsetAsync(u->adam_v.ptr, 0, u->nparam * 4, s));
            FG_HIP(hipMemsetAsync(u->aux.ptr, 0, (size_t)u->d0.n() * 4, s));
            sync_bf16(u);
            FG_HIP(hipStreamSynchronize(s));
        } catch (...) {
            delete u;
            throw;
        }
        *out = u;
        return 0;
    });
}

int femgpu_unet_create(femgpu_ctx *c, int nx, int ny, int nz, uint64_t seed, femgpu_unet **out) {
    return unet_create(c, nx, ny, nz, 1, seed, out);
}

int femgpu_unet_create_coarse(femgpu_ctx *c, int nx, int ny, int nz, int coarsen, uint64_t seed,
                              femgpu_unet **out) {
    return unet_create(c, nx, ny, nz, coarsen, seed, out);
}

int femgpu_unet_grid(const femgpu_unet *u, int *nx, int *ny, int *nz) {
    return guard([&] {
        FG_REQUIRE(u && nx && ny && nz, FEMGPU_ERR_INVALID_ARG, "unet_grid: NULL argument");
        *nx = u->d0.nx; *ny = u->d0.ny; *nz = u->d0.nz;
        return 0;
    });
}

int femgpu_unet_destroy(femgpu_unet *u) {
    if (!u) return 0;
    (void)hipSetDevice(u->ctx->device);
    (void)hipStreamSynchronize(u->ctx->stream);
    delete u;
    return 0;
}

int femgpu_unet_n_params(const femgpu_unet *u, uint64_t *n) {
    return guard([&] {
        FG_REQUIRE(u && n, FEMGPU_ERR_INVALID_ARG, "unet_n_params: NULL");
        *n = u->nparam;
        return 0;
    });
}

int femgpu_unet_layer_info(const femgpu_unet *u, int layer, int *cin, int *cout, int *kpad,
                           uint64_t *woff, uint64_t *boff) {
    return guard([&] {
        FG_REQUIRE(u && layer >= 0 && layer < 6, FEMGPU_ERR_INVALID_ARG, "unet_layer_info: bad layer");
        const Layer &L = u->L[layer];
        if (cin) *cin = L.cin;
        if (cout) *cout = L.cout;
        if (kpad) *kpad = L.kpad;
        if (woff) *woff = L.woff;
        if (boff) *boff = L.boff;
        return 0;
    });
}

int femgpu_unet_get_params(femgpu_unet *u, float *host) {
    return guard([&] {
        FG_REQUIRE(u && host, FEMGPU_ERR_INVALID_ARG, "unet_get_params: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipMemcpyAsync(host, u->params.ptr, u->nparam * 4, hipMemcpyDeviceToHost, u->ctx->stream));
        FG_HIP(hipStreamSynchronize(u->ctx->stream));
        return 0;
    });
}

int femgpu_unet_set_params(femgpu_unet *u, const float *host) {
    return guard([&] {
        FG_REQUIRE(u && host, FEMGPU_ERR_INVALID_ARG, "unet_set_params: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipMemcpyAsync(u->params.ptr, host, u->nparam * 4, hipMemcpyHostToDevice, u->ctx->stream));
        sync_bf16(u);
        FG_HIP(hipStreamSynchronize(u->ctx->stream));
        return 0;
    });
}

int femgpu_unet_set_aux(femgpu_unet *u, const float *aux, femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(u && aux, FEMGPU_ERR_INVALID_ARG, "unet_set_aux: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipMemcpyAsync(u->aux.ptr, aux, (size_t)u->d0.n() * 4,
                              mem == FEMGPU_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, u->ctx->stream));
        FG_HIP(hipStreamSynchronize(u->ctx->stream));
        return 0;
    });
}

// Raw network: out = N(in) with in channel 0 = x0 (as given, no scaling) and
// channel 1 = the aux field.  Keeps the activations for femgpu_unet_backward.
int femgpu_unet_forward(femgpu_unet *u, const float *x0, float *out, femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(u && x0 && out, FEMGPU_ERR_INVALID_ARG, "unet_forward: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipSetDevice(u->ctx->device));
        hipStream_t s = u->ctx->stream;
        const long long V = u->d0.n();
        const float *xd = x0;
        DevBuf st;
        if (mem == FEMGPU_HOST) {
            st.alloc(V * 4);
            FG_HIP(hipMemcpyAsync(st.ptr, x0, V * 4, hipMemcpyHostToDevice, s));
            xd = st.as<float>();
        }
        hipLaunchKernelGGL(k_make_input<float>, dim3(g1(V)), dim3(256), 0, s, V, xd, 1.0,
                           (const float *)u->aux.as<float>(), u->in.as<uint16_t>());
        forward(u);
        FG_HIP(hipMemcpyAsync(out, u->out.ptr, V * 4, mem == FEMGPU_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
        FG_HIP(hipGetLastError());
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

// Parameter gradient of sum(dout * out) at the last forward input.
int femgpu_unet_backward(femgpu_unet *u, const float *dout, float *grads, femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(u && dout && grads, FEMGPU_ERR_INVALID_ARG, "unet_backward: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipSetDevice(u->ctx->device));
        hipStream_t s = u->ctx->stream;
        const long long V = u->d0.n();
        FG_HIP(hipMemcpyAsync(u->gout.ptr, dout, V * 4, mem == FEMGPU_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
        backward(u);
        FG_HIP(hipMemcpyAsync(grads, u->grads.ptr, u->nparam * 4,
                              mem == FEMGPU_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
        FG_HIP(hipGetLastError());
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

// One Adam step on the pair (r, e): input r sqrt(V)/|r|, target e sqrt(V)/|e|,
// loss 0.5 |N - t|^2 / |t|^2.
int femgpu_unet_train_step(femgpu_unet *u, const float *r, const float *e, femgpu_mem mem, float lr,
                           double *loss) {
    return guard([&] {
        FG_REQUIRE(u && r && e, FEMGPU_ERR_INVALID_ARG, "unet_train_step: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipSetDevice(u->ctx->device));
        hipStream_t s = u->ctx->stream;
        femgpu_ctx *c = u->ctx;
        const long long V = u->d0.n();
        const float *rd = r, *ed = e;
        DevBuf sr, se;
        if (mem == FEMGPU_HOST) {
            sr.alloc(V * 4); se.alloc(V * 4);
            FG_HIP(hipMemcpyAsync(sr.ptr, r, V * 4, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemcpyAsync(se.ptr, e, V * 4, hipMemcpyHostToDevice, s));
            rd = sr.as<float>(); ed = se.as<float>();
        }
        const double rn = std::sqrt(dot_host(c, rd, rd, V, FEMGPU_F32));
        const double en = std::sqrt(dot_host(c, ed, ed, V, FEMGPU_F32));
        FG_REQUIRE(rn > 0 && en > 0 && std::isfinite(rn) && std::isfinite(en), FEMGPU_ERR_INVALID_ARG,
                   "unet_train_step: zero or non-finite pair");
        hipLaunchKernelGGL(k_make_input<float>, dim3(g1(V)), dim3(256), 0, s, V, rd, std::sqrt((double)V) / rn,
                           (const float *)u->aux.as<float>(), u->in.as<uint16_t>());
        forward(u);
        // target t = e sqrt(V)/|e| into gtmp (f32), |t|^2 = V
        float *t = u->gtmp.as<float>();
        hipLaunchKernelGGL(k_out_scale<float>, dim3(g1(V)), dim3(256), 0, s, V, ed, std::sqrt((double)V) / en, t, 0);
        const int nb = red_blocks(V);
        hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, V, (const float *)u->out.as<float>(), (const float *)t,
                           1.0 / (double)V, u->gout.as<float>(), c->partials.as<double>());
        std::vector<double> hp(nb);
        FG_HIP(hipMemcpyAsync(hp.data(), c->partials.ptr, nb * 8, hipMemcpyDeviceToHost, s));
        backward(u);
        ++u->step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
        hipLaunchKernelGGL(k_adam, dim3(g1((long long)u->nparam)), dim3(256), 0, s, (long long)u->nparam,
                           u->params.as<float>(), (const float *)u->grads.as<float>(), u->adam_m.as<float>(),
                           u->adam_v.as<float>(), lr, b1, b2, 1e-8f, c1, c2, (const uint8_t *)u->mask.as<uint8_t>());
        sync_bf16(u);
        FG_HIP(hipGetLastError());
        FG_HIP(hipStreamSynchronize(s));
        double L = 0.0;
        for (double v : hp) L += v;
        if (loss) *loss = L;
        return 0;
    });
}

// One Adam step on the energy-norm loss 1 - (n.r)^2 / ((n.A n)(e.r)) of the
// Galerkin-scaled network correction n = N(r sqrt(V)/|r|).
int femgpu_unet_train_step_energy(femgpu_unet *u, femgpu_op *op, const float *r, const float *e,
                                  femgpu_mem mem, float lr, double *loss) {
    return guard([&] {
        FG_REQUIRE(u && op && r && e, FEMGPU_ERR_INVALID_ARG, "unet_train_step_energy: NULL");
        FG_REQUIRE(op->ctx == u->ctx, FEMGPU_ERR_INVALID_ARG, "unet_train_step_energy: different contexts");
        FG_REQUIRE((long long)op->n == unet_voxels(u), FEMGPU_ERR_DIM_MISMATCH,
                   "unet_train_step_energy: network grid and operator size differ");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipSetDevice(u->ctx->device));
        hipStream_t s = u->ctx->stream;
        femgpu_ctx *c = u->ctx;
        // V: the DOF grid (the operator's vectors); Vn: the network's grid
        const long long V = unet_voxels(u), Vn = u->d0.n();
        const bool coarse = u->cf > 1;
        const float *rd = r, *ed = e;
        DevBuf sr, se;
        if (mem == FEMGPU_HOST) {
            sr.alloc(V * 4); se.alloc(V * 4);
            FG_HIP(hipMemcpyAsync(sr.ptr, r, V * 4, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemcpyAsync(se.ptr, e, V * 4, hipMemcpyHostToDevice, s));
            rd = sr.as<float>(); ed = se.as<float>();
        }
        const double rn = std::sqrt(dot_host(c, rd, rd, V, FEMGPU_F32));
        FG_REQUIRE(rn > 0 && std::isfinite(rn), FEMGPU_ERR_INVALID_ARG,
                   "unet_train_step_energy: zero or non-finite residual");
        // the network input, normalised as in unet_apply_t
        const float *rin = rd;
        double sc = std::sqrt((double)V) / rn;
        if (coarse) {
            restrict_to_coarse<float>(u, rd);
            const double rc = std::sqrt(dot_host(c, u->rc.ptr, u->rc.ptr, Vn, FEMGPU_F32));
            FG_REQUIRE(rc > 0 && std::isfinite(rc), FEMGPU_ERR_INVALID_ARG,
                       "unet_train_step_energy: residual vanishes on the network's grid");
            rin = u->rc.as<float>();
            sc = std::sqrt((double)Vn) / rc;
        }
        hipLaunchKernelGGL(k_make_input<float>, dim3(g1(Vn)), dim3(256), 0, s, Vn, rin, sc,
                           (const float *)u->aux.as<float>(), u->in.as<uint16_t>());
        forward(u);
        // n on the DOF grid: the output itself, or its prolongation
        const float *nout = u->out.as<float>();
        if (coarse) {
            hipLaunchKernelGGL(k_prolong<float>, dim3(prolong_blocks(u->fine, u->d0)), dim3(256),
                               sizeof(float) * prolong_lines(u->d0) * u->d0.nx, s, u->fine, u->d0, u->cf, prolong_lines(u->d0), (const float *)u->out.as<float>(),
                               1.0, u->bmask, u->nf.as<float>());
            nout = u->nf.as<float>();
        }
        const int nb = red_blocks(V);
        double a = 0, cc = 0, E = 0;
        auto energy = [&](auto tag) {
            using T = decltype(tag);
            T *nT = (T *)op->wsvec(kUnetOpSlot, V * sizeof(T));
            T *AnT = (T *)op->wsvec(kUnetOpSlot + 1, V * sizeof(T));
            hipLaunchKernelGGL(k_out_scale<T>, dim3(g1(V)), dim3(256), 0, s, V, nout, 1.0, nT, 0);
            op->apply_dev(nT, AnT, false, nullptr);
            hipLaunchKernelGGL(k_energy_dots<T>, dim3(nb), dim3(256), 0, s, V, nout, rd, ed, (const T *)AnT, sc,
                               c->partials.as<double>());
            std::vector<double> hp(3 * (size_t)kRedBlocks);
            FG_HIP(hipMemcpyAsync(hp.data(), c->partials.ptr, hp.size() * 8, hipMemcpyDeviceToHost, s));
            FG_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < nb; ++i) { a += hp[i]; cc += hp[kRedBlocks + i]; E += hp[2 * kRedBlocks + i]; }
            FG_REQUIRE(cc > 0 && E > 0 && std::isfinite(a) && std::isfinite(cc) && std::isfinite(E),
                       FEMGPU_ERR_INVALID_ARG,
                       "unet_train_step_energy: network output or pair gives no positive energy");
            const double gamma = a / cc, coef = -2.0 * a / (cc * E);
            // d loss / d n on the DOF grid; with the coarse placement it goes to the
            // network's output through P^T = R
            float *gn = coarse ? u->nf.as<float>() : u->gout.as<float>();
            hipLaunchKernelGGL(k_energy_grad<T>, dim3(g1(V)), dim3(256), 0, s, V, rd, sc, (const T *)AnT, coef,
                               gamma, gn);
            if (coarse) {
                restrict_to_coarse<float>(u, gn);
                FG_HIP(hipMemcpyAsync(u->gout.ptr, u->rc.ptr, (size_t)Vn * 4, hipMemcpyDeviceToDevice, s));
            }
        };
        if (op->dtype == FEMGPU_F64) energy(double{});
        else energy(float{});
        backward(u);
        ++u->step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
        hipLaunchKernelGGL(k_adam, dim3(g1((long long)u->nparam)), dim3(256), 0, s, (long long)u->nparam,
                           u->params.as<float>(), (const float *)u->grads.as<float>(), u->adam_m.as<float>(),
                           u->adam_v.as<float>(), lr, b1, b2, 1e-8f, c1, c2, (const uint8_t *)u->mask.as<uint8_t>());
        sync_bf16(u);
        FG_HIP(hipGetLastError());
        FG_HIP(hipStreamSynchronize(s));
        if (loss) *loss = 1.0 - a * a / (cc * E);
        return 0;
    });
}

// z = preconditioner correction of r (see unet_apply_t); r, z of type dtype.
int femgpu_unet_apply(femgpu_unet *u, const void *r, void *z, femgpu_dtype dt, femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(u && r && z, FEMGPU_ERR_INVALID_ARG, "unet_apply: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        FG_HIP(hipSetDevice(u->ctx->device));
        femgpu_ctx *c = u->ctx;
        hipStream_t s = c->stream;
        const long long V = unet_voxels(u);
        const size_t es = dt == FEMGPU_F64 ? 8 : 4;
        const void *rd = r;
        void *zd = z;
        DevBuf sr, sz;
        if (mem == FEMGPU_HOST) {
            sr.alloc(V * es); sz.alloc(V * es);
            FG_HIP(hipMemcpyAsync(sr.ptr, r, V * es, hipMemcpyHostToDevice, s));
            rd = sr.ptr; zd = sz.ptr;
        }
        const double rn = std::sqrt(dot_host(c, rd, rd, V, dt));
        if (!unet_apply_dev(u, rd, zd, dt, rn)) FG_HIP(hipMemsetAsync(zd, 0, V * es, s));
        if (mem == FEMGPU_HOST) FG_HIP(hipMemcpyAsync(z, zd, V * es, hipMemcpyDeviceToHost, s));
        FG_HIP(hipGetLastError());
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

// Checkpoint: "FGUNET01", grid (3 x int32), nparam (u64), adam step (i64),
// params, adam m, adam v (f32), FNV-1a 64 checksum of everything before it.
static uint64_t fnv1a(const void *p, size_t n, uint64_t h) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

int femgpu_unet_save(femgpu_unet *u, const char *path) {
    return guard([&] {
        FG_REQUIRE(u && path, FEMGPU_ERR_INVALID_ARG, "unet_save: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        std::vector<float> p(u->nparam), m(u->nparam), v(u->nparam);
        hipStream_t s = u->ctx->stream;
        FG_HIP(hipMemcpyAsync(p.data(), u->params.ptr, u->nparam * 4, hipMemcpyDeviceToHost, s));
        FG_HIP(hipMemcpyAsync(m.data(), u->adam_m.ptr, u->nparam * 4, hipMemcpyDeviceToHost, s));
        FG_HIP(hipMemcpyAsync(v.data(), u->adam_v.ptr, u->nparam * 4, hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
        std::ofstream f(path, std::ios::binary);
        FG_REQUIRE(f.good(), FEMGPU_ERR_INVALID_ARG, std::string("unet_save: cannot open ") + path);
        const char magic[8] = {'F', 'G', 'U', 'N', 'E', 'T', '0', '1'};
        const int32_t dims[3] = {u->d0.nx, u->d0.ny, u->d0.nz};
        const uint64_t np = u->nparam;
        const int64_t step = u->step;
        uint64_t h = 1469598103934665603ull;
        auto put = [&](const void *d, size_t n) { f.write((const char *)d, n); h = fnv1a(d, n, h); };
        put(magic, 8); put(dims, 12); put(&np, 8); put(&step, 8);
        put(p.data(), np * 4); put(m.data(), np * 4); put(v.data(), np * 4);
        f.write((const char *)&h, 8);
        FG_REQUIRE(f.good(), FEMGPU_ERR_INVALID_ARG, "unet_save: write failed");
        return 0;
    });
}

int femgpu_unet_load(femgpu_unet *u, const char *path) {
    return guard([&] {
        FG_REQUIRE(u && path, FEMGPU_ERR_INVALID_ARG, "unet_load: NULL");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        std::ifstream f(path, std::ios::binary);
        FG_REQUIRE(f.good(), FEMGPU_ERR_INVALID_ARG, std::string("unet_load: cannot open ") + path);
        char magic[8];
        int32_t dims[3];
        uint64_t np = 0, hs = 0;
        int64_t step = 0;
        uint64_t h = 1469598103934665603ull;
        auto get = [&](void *d, size_t n) { f.read((char *)d, n); h = fnv1a(d, n, h); };
        get(magic, 8);
        FG_REQUIRE(f.good() && std::memcmp(magic, "FGUNET01", 8) == 0, FEMGPU_ERR_INVALID_ARG, "unet_load: not a checkpoint");
        get(dims, 12); get(&np, 8); get(&step, 8);
        FG_REQUIRE(np == u->nparam, FEMGPU_ERR_DIM_MISMATCH, "unet_load: parameter count differs");
        std::vector<float> p(np), m(np), v(np);
        get(p.data(), np * 4); get(m.data(), np * 4); get(v.data(), np * 4);
        f.read((char *)&hs, 8);
        FG_REQUIRE(f.good() && hs == h, FEMGPU_ERR_INVALID_ARG, "unet_load: checksum mismatch");
        hipStream_t s = u->ctx->stream;
        FG_HIP(hipMemcpyAsync(u->params.ptr, p.data(), np * 4, hipMemcpyHostToDevice, s));
        FG_HIP(hipMemcpyAsync(u->adam_m.ptr, m.data(), np * 4, hipMemcpyHostToDevice, s));
        FG_HIP(hipMemcpyAsync(u->adam_v.ptr, v.data(), np * 4, hipMemcpyHostToDevice, s));
        u->step = step;
        sync_bf16(u);
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

}  // extern "C"

extern "C" int femgpu_unet_aux_from_hex(femgpu_unet *u, femgpu_op *hex) {
    return guard([&] {
        FG_REQUIRE(u && hex, FEMGPU_ERR_INVALID_ARG, "unet_aux_from_hex: NULL");
        int Nx = 0, Ny = 0, Nz = 0;
        FG_REQUIRE(hex_grid_dims(hex, &Nx, &Ny, &Nz), FEMGPU_ERR_INVALID_ARG, "unet_aux_from_hex: not a hex operator");
        const Vol g = u->cf > 1 ? u->fine : u->d0;
        FG_REQUIRE(Nx == g.nx && Ny == g.ny && Nz == g.nz, FEMGPU_ERR_DIM_MISMATCH,
                   "unet_aux_from_hex: grid of the network and the operator differ");
        FG_REQUIRE(hex->ctx == u->ctx, FEMGPU_ERR_INVALID_ARG, "unet_aux_from_hex: different contexts");
        std::lock_guard<std::recursive_mutex> lk(u->ctx->mu);
        if (u->cf > 1) {
            // log kappa at the DOF nodes, then at the coarse nodes (a subset of them)
            hex_log_kappa_nodes(hex, u->nf.as<float>());
            hipLaunchKernelGGL(k_inject, dim3(g1(u->d0.n())), dim3(256), 0, u->ctx->stream, u->fine, u->d0, u->cf,
                               (const float *)u->nf.as<float>(), u->aux.as<float>());
            u->bmask = hex_dirichlet(hex) ? 1 : 0;
        } else {
            hex_log_kappa_nodes(hex, u->aux.as<float>());
        }
        FG_HIP(hipStreamSynchronize(u->ctx->stream));
        return 0;
    });
}
