// Sum-factorised partial-assembly (PA) mass / diffusion apply on structured
// hexahedral Q_p meshes for gfx950.
//
// Reference behaviour: PADiffusionOperator / PAMassOperator::apply
// (assembly/src/partial.rs:85-130, 154-208), y_e += sum_q w|J| kappa grad phi_i .
// grad u (+ w|J| rho phi_i u), which the reference evaluates element by element
// on simplices only.  Here the element operator is factored through 1-D
// tables: with B = l_i(x_q), G = l_i'(x_q)
//   grad_ref u = (B(x)B(x)G, B(x)G(x)B, G(x)B(x)B) u,  v = D grad_ref u,
//   y_e = (B^T B^T G^T, B^T G^T B^T, G^T B^T B^T) v + (B^T)^3 (M u)
// where D = w kappa |J| J^-1 J^-T and M = w rho |J| are the quadrature data
// ("qdata") computed once at set-up from the trilinear (HexQ1) geometry.
//
// An apply is two kernels: the element kernel writes an element vector (the
// x-merged layout below: x-neighbours inside a group summed, no atomics), and
// k_hex_gather sums it per DOF in a fixed order, so applies are bitwise
// reproducible.
//
// Workgroup element kernel k_pa_hex (one workgroup = NE elements x Q*Q lanes,
// 64-wide waves):
//   X  : lane (j,k) reads its x-line of the element from HBM and contracts it
//        in registers -> LDS
//   Y  : lane (qx,k) contracts along y in registers -> LDS
//   Z  : lane (qx,qy) contracts along z, streams its qdata column from HBM,
//        applies D / M and immediately contracts back along z, all in
//        registers -> LDS (in place, no hazard: each lane owns its column)
//   Y^T, X^T as Y, X; X^T writes the element's rows of the element vector.
// Every contraction of a line is done in registers, so LDS carries ~10 values
// per element point while the 1-D tables come from the scalar cache
// (compile-time-indexed __constant__ memory).  Q2 diffusion with Cartesian
// qdata instead runs k_pa_hex_tpe, one element per thread with no LDS.
//
// qdata is stored tiled by workgroup and plane-major inside a tile (qd_index),
// so the Z phase of a wavefront reads contiguous, coalesced runs;
// k_pa_hex_tpe's layout (qd_index_tpe) is contiguous across its 64 lanes.
//
// On an axis-aligned box mesh (vertex x depends on i only, y on j, z on k) J
// is diagonal and constant per element, D = (w kappa |J|) diag(1/J_xx^2,
// 1/J_yy^2, 1/J_zz^2) with the three factors depending on ex, ey, ez alone.
// There ("Cartesian" qdata) one value per point, s = w kappa |J|, is stored
// with per-axis factor arrays cf = [1/J_xx^2 (nx) | 1/J_yy^2 (ny) | 1/J_zz^2
// (nz)], a sixth of the general diffusion qdata.
#include "basis.hpp"
#include "internal.hpp"
#include "philox.hpp"
#include "reduce.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace femgpu {

constexpr int kTab = 64;
constexpr int combo(int n1, int q) { return (n1 - 2) * 3 + (q - n1); }
constexpr int kCombos = 15;  // n1 = 2..6, q = n1..n1+2

__constant__ double c_tab64[kCombos][2][kTab];
__constant__ float c_tab32[kCombos][2][kTab];
// Gauss-Legendre points (on [-1, 1]) and weights for q = 0..7 points: [q][0] points, [q][1] weights
constexpr int kGaussMax = 8;
__constant__ double c_gauss64[kGaussMax][2][kGaussMax];
__constant__ float c_gauss32[kGaussMax][2][kGaussMax];

template <class T> struct Tab;
template <> struct Tab<double> {
    static __device__ __forceinline__ const double *B(int c) { return c_tab64[c][0]; }
    static __device__ __forceinline__ const double *G(int c) { return c_tab64[c][1]; }
    static __device__ __forceinline__ const double *X(int q) { return c_gauss64[q][0]; }
    static __device__ __forceinline__ const double *W(int q) { return c_gauss64[q][1]; }
};
template <> struct Tab<float> {
    static __device__ __forceinline__ const float *B(int c) { return c_tab32[c][0]; }
    static __device__ __forceinline__ const float *G(int c) { return c_tab32[c][1]; }
    static __device__ __forceinline__ const float *X(int q) { return c_gauss32[q][0]; }
    static __device__ __forceinline__ const float *W(int q) { return c_gauss32[q][1]; }
};

// Even-odd form of the 1-D tables.  With nodes and Gauss points symmetric
// about the centre, an (R x Cc) table M (B or G, or their transposes) obeys
// M[R-1-r][Cc-1-c] = s M[r][c] with s = +1 (B) or -1 (G).  With the mirrored
// sums e_c = u_c + u_{Cc-1-c} and differences o_c = u_c - u_{Cc-1-c} (c <
// Cc/2) and, for odd Cc, the middle input m, the products are
//   Se_r = sum_c P[r][c] e_c + Mid[r] m,   So_r = sum_c Mm[r][c] o_c
//   (M u)_r = Se_r + So_r,   (M u)_{R-1-r} = s (Se_r - So_r)
// for r < (R+1)/2, with P = (M[r][c] + M[r][Cc-1-c]) / 2, Mm = (M[r][c] -
// M[r][Cc-1-c]) / 2 and Mid[r] = M[r][Cc/2] (a middle output row keeps only
// Se for s = +1 and only So for s = -1).  A 5 x 5 line product takes 13
// multiply-adds and 8 additions instead of 25 multiply-adds (Kronbichler and
// Kormann's even-odd decomposition).  Table [combo][m]: m = 0 B (Q x N1), 1
// G, 2 B^T (N1 x Q), 3 G^T; P at [r (Cc/2) + c], Mid at [16 + r], Mm at
// [24 + r (Cc/2) + c].
__constant__ double c_eo64[kCombos][4][kTab];
__constant__ float c_eo32[kCombos][4][kTab];
template <class T> __device__ __forceinline__ const T *eo_tab(int c, int m);
template <> __device__ __forceinline__ const double *eo_tab<double>(int c, int m) { return c_eo64[c][m]; }
template <> __device__ __forceinline__ const float *eo_tab<float>(int c, int m) { return c_eo32[c][m]; }

// mirrored sums / differences of an NI-vector and its middle entry
template <int NI, class T> struct EOs {
    T e[NI / 2 > 0 ? NI / 2 : 1], o[NI / 2 > 0 ? NI / 2 : 1];
    T m;
};
template <int NI, class T> __device__ __forceinline__ EOs<NI, T> eo_split(const T (&u)[NI]) {
    EOs<NI, T> r;
#pragma unroll
    for (int c = 0; c < NI / 2; ++c) {
        r.e[c] = u[c] + u[NI - 1 - c];
        r.o[c] = u[c] - u[NI - 1 - c];
    }
    r.m = (NI & 1) ? u[NI / 2] : T(0);
    return r;
}
// Se_r (EVEN) or So_r (odd part) for r < (R+1)/2, added to acc; the middle
// row (odd R) only when MIDROW
template <int R, int NI, bool EVEN, bool MIDROW, class T>
__device__ __forceinline__ void eo_part(const T *tab, const EOs<NI, T> &u, T (&acc)[(R + 1) / 2]) {
    constexpr int HC = NI / 2, HR = R / 2;
#pragma unroll
    for (int r = 0; r < HR + ((R & 1) && MIDROW ? 1 : 0); ++r) {
        T a = acc[r];
        if constexpr (EVEN) {
            if constexpr (NI & 1) a += tab[16 + r] * u.m;
#pragma unroll
            for (int c = 0; c < HC; ++c) a += tab[r * HC + c] * u.e[c];
        } else {
#pragma unroll
            for (int c = 0; c < HC; ++c) a += tab[24 + r * HC + c] * u.o[c];
        }
        acc[r] = a;
    }
}
// out = M u for a table of symmetry SIG
template <int R, int NI, int SIG, class T>
__device__ __forceinline__ void eo_contract(const T *tab, const EOs<NI, T> &u, T (&out)[R]) {
    constexpr int HR = R / 2, HR1 = (R + 1) / 2;
    T se[HR1], so[HR1];
#pragma unroll
    for (int r = 0; r < HR1; ++r) se[r] = so[r] = T(0);
    eo_part<R, NI, true, SIG == 1>(tab, u, se);
    eo_part<R, NI, false, SIG == -1>(tab, u, so);
#pragma unroll
    for (int r = 0; r < HR; ++r) {
        out[r] = se[r] + so[r];
        out[R - 1 - r] = SIG == 1 ? se[r] - so[r] : so[r] - se[r];
    }
    if constexpr (R & 1) out[HR] = SIG == 1 ? se[HR] : so[HR];
}
// out = M1 u1 + M2 u2 for tables of symmetry +1 (M1) and -1 (M2): A = Se1 +
// So2 and Bv = So1 + Se2 give out_r = A + Bv, out_{R-1-r} = A - Bv
template <int R, int NI, class T>
__device__ __forceinline__ void eo_contract2(const T *tab1, const EOs<NI, T> &u1, const T *tab2,
                                             const EOs<NI, T> &u2, T (&out)[R]) {
    constexpr int HR = R / 2, HR1 = (R + 1) / 2;
    T a[HR1], b[HR1];
#pragma unroll
    for (int r = 0; r < HR1; ++r) a[r] = b[r] = T(0);
    eo_part<R, NI, true, true>(tab1, u1, a);
    eo_part<R, NI, false, true>(tab2, u2, a);
    eo_part<R, NI, false, false>(tab1, u1, b);
    eo_part<R, NI, true, false>(tab2, u2, b);
#pragma unroll
    for (int r = 0; r < HR; ++r) {
        out[r] = a[r] + b[r];
        out[R - 1 - r] = a[r] - b[r];
    }
    if constexpr (R & 1) out[HR] = a[HR];
}

// the even-odd tables of an R x Cc table M (row-major, or its transpose when
// TR: M[r][c] = src[c * R + r])
static void eo_fill(const double *src, int R, int Cc, bool tr, double *dst) {
    auto M = [&](int r, int c) { return tr ? src[c * R + r] : src[r * Cc + c]; };
    const int HC = Cc / 2, HR1 = (R + 1) / 2;
    for (int r = 0; r < HR1; ++r) {
        for (int c = 0; c < HC; ++c) {
            dst[r * HC + c] = 0.5 * (M(r, c) + M(r, Cc - 1 - c));
            dst[24 + r * HC + c] = 0.5 * (M(r, c) - M(r, Cc - 1 - c));
        }
        dst[16 + r] = (Cc & 1) ? M(r, HC) : 0.0;
    }
}

static std::mutex g_tab_mu;
static bool g_tab_done[64];

// Fill the 1-D basis tables for every supported (n1, q) on the current device.
static void ensure_tables(int device) {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    if (device < 64 && g_tab_done[device]) return;
    static double h64[kCombos][2][kTab];
    static float h32[kCombos][2][kTab];
    static double e64[kCombos][4][kTab];
    static float e32[kCombos][4][kTab];
    std::memset(h64, 0, sizeof h64);
    std::memset(h32, 0, sizeof h32);
    std::memset(e64, 0, sizeof e64);
    std::memset(e32, 0, sizeof e32);
    for (int n1 = 2; n1 <= 6; ++n1)
        for (int q = n1; q <= n1 + 2 && q <= 6; ++q) {
            double x[8], w[8], B[64], G[64];
            gauss_legendre(q, x, w);
            lagrange_1d(n1 - 1, q, x, B, G);
            const int c = combo(n1, q);
            for (int i = 0; i < q * n1; ++i) {
                h64[c][0][i] = B[i]; h64[c][1][i] = G[i];
                h32[c][0][i] = (float)B[i]; h32[c][1][i] = (float)G[i];
            }
            eo_fill(B, q, n1, false, e64[c][0]);
            eo_fill(G, q, n1, false, e64[c][1]);
            eo_fill(B, n1, q, true, e64[c][2]);
            eo_fill(G, n1, q, true, e64[c][3]);
            for (int m = 0; m < 4; ++m)
                for (int i = 0; i < kTab; ++i) e32[c][m][i] = (float)e64[c][m][i];
        }
    FG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tab64), h64, sizeof h64));
    FG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tab32), h32, sizeof h32));
    FG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_eo64), e64, sizeof e64));
    FG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_eo32), e32, sizeof e32));
    static double g64[kGaussMax][2][kGaussMax];
    static float g32[kGaussMax][2][kGaussMax];
    std::memset(g64, 0, sizeof g64);
    std::memset(g32, 0, sizeof g32);
    for (int q = 1; q <= 6; ++q) {  // the element kernels' point counts
        double x[8], w[8];
        gauss_legendre(q, x, w);
        for (int i = 0; i < q; ++i) {
            g64[q][0][i] = x[i]; g64[q][1][i] = w[i];
            g32[q][0][i] = (float)x[i]; g32[q][1][i] = (float)w[i];
        }
    }
    FG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gauss64), g64, sizeof g64));
    FG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gauss32), g32, sizeof g32));
    if (device < 64) g_tab_done[device] = true;
}

__host__ __device__ constexpr int ne_of(int q) { return 256 / (q * q) > 0 ? 256 / (q * q) : 1; }

struct HexGrid {
    int nx, ny, nz;
    long long E;
    long long Nx, Ny, Nz;
    int ne;         // elements per element-kernel workgroup, ne_of(Q)
    long long Nxp;  // row stride of the x-merged element vector: Nx rounded up to even
    int ymerge;     // k_pa_hex_tpe with nx = 64 or 128: y-faces merged inside 256-element blocks
    unsigned mx, my;  // n / nx and n / ny for n < 2^31 as udiv_m(n, mx, shx) (div_magic)
    int shx, shy;
};

// Division by a mesh dimension d >= 1 without the division sequence: for
// n < 2^31, n / d = umulhi(n, M) >> (s - 1) with s = ceil(log2 d) and M =
// ceil(2^(31+s) / d) < 2^32 (the error n e / 2^(31+s), e = M d - 2^(31+s) <
// d <= 2^s, stays below 1/d); M = 0 stands for d = 1.
static inline void div_magic(long long d, unsigned &m, int &sh) {
    if (d <= 1) { m = 0; sh = 0; return; }
    int s = 0;
    while ((1ll << s) < d) ++s;
    m = (unsigned)(((1ull << (31 + s)) + (unsigned long long)d - 1) / (unsigned long long)d);
    sh = s - 1;
}
__device__ __forceinline__ unsigned udiv_m(unsigned n, unsigned m, int sh) {
    return m ? __umulhi(n, m) >> sh : n;
}


__device__ __forceinline__ bool on_bdr(const HexGrid &g, long long gx, long long gy, long long gz) {
    return gx == 0 || gx == g.Nx - 1 || gy == 0 || gy == g.Ny - 1 || gz == 0 || gz == g.Nz - 1;
}

// Element-vector layout ("x-merged"): for each element row (ey, ez) and local
// (k, j) a row of Nx values (stride Nxp, even, so that even-P segments are
// 16-byte aligned), one per DOF position gx along x.  Where two
// elements of one workgroup share an x-face DOF, the row holds their sum
// (lower element first); everywhere else it holds the single contribution.
// The first element of a workgroup (e % ne == 0) with ex > 0 stores its
// i = 0 value in the seam area instead, at seam_index(e / ne, j, k), and the
// gather adds it.  This keeps Nx instead of nx (P + 1) values per row.
__host__ __device__ __forceinline__ size_t mrow_index(const HexGrid &g, int N1, int ey, int ez, int j,
                                                      int k) {
    return ((((size_t)ez * g.ny + ey) * N1 + k) * N1 + j) * (size_t)g.Nxp;
}
__host__ __device__ __forceinline__ size_t seam_base(const HexGrid &g, int N1) {
    return (size_t)g.nz * g.ny * N1 * N1 * (size_t)g.Nxp;
}
__host__ __device__ __forceinline__ size_t seam_index(const HexGrid &g, int N1, long long blk, int j,
                                                      int k) {
    return seam_base(g, N1) + ((size_t)blk * N1 + k) * N1 + j;
}
// entries of the element vector
static inline size_t evec_count(const HexGrid &g, int N1) {
    return seam_base(g, N1) + (size_t)((g.E + g.ne - 1) / g.ne) * N1 * N1;
}

// Quadrature-data layout.  Elements are grouped in blocks of NE = ne_of(Q)
// consecutive elements (one element-kernel workgroup each); inside a block the
// data are plane-major over (component, qz), then element-major over
// (element, qx + Q qy):
//   index = ((blk * ncomp + comp) * Q + qz) * NE Q^2 + le Q^2 + (qx + Q qy)
// Each workgroup streams one contiguous range, and every (component, qz) load
// of a workgroup is one contiguous run of NE Q^2 values.
__host__ __device__ __forceinline__ size_t qd_index(int Q, int ncomp, long long e, int comp, int qz,
                                                    int t) {
    const int NE = ne_of(Q), T2 = Q * Q;
    const long long blk = e / NE, le = e - blk * NE;
    return (((size_t)blk * ncomp + comp) * Q + qz) * (size_t)(NE * T2) + (size_t)le * T2 + t;
}
// Thread-per-element layout (k_pa_hex_tpe): groups of 64 consecutive elements
// (one wavefront), point-major inside a group and element-minor, so every
// (component, point) load of a wavefront is one contiguous run of 64 values:
//   index = ((grp * ncomp + comp) * Q^3 + qp) * 64 + (e mod 64),  qp = qx + Q (qy + Q qz)
__host__ __device__ __forceinline__ size_t qd_index_tpe(int Q, int ncomp, long long e, int comp, int qp) {
    const long long grp = e >> 6;
    return (((size_t)grp * ncomp + comp) * (size_t)(Q * Q * Q) + qp) * 64 + (size_t)(e & 63);
}
// values per component in either layout (E rounded up to whole blocks)
static inline size_t qd_count_any(long long E, int Q) {
    const long long NE = ne_of(Q);
    const size_t a = (size_t)((E + NE - 1) / NE) * NE, b = (size_t)((E + 63) / 64) * 64;
    return (a > b ? a : b) * Q * Q * Q;
}
// values per component (E rounded up to whole blocks)
static inline size_t qd_count(long long E, int Q) {
    const long long NE = ne_of(Q);
    return (size_t)((E + NE - 1) / NE) * NE * Q * Q * Q;
}

// LDS layout of the element kernel's per-element arrays: index order (outer to
// inner) of A(k, j, a) and B(k, b, a), and the per-element strides SA, SB
// (>= the array sizes; the extra is padding).  Chosen per (N1, Q, element
// size) by counting, for each phase's lane -> address map, the bank
// multiplicity of every 32-lane (8-byte) or 64-lane (4-byte) group on 64
// banks of 4 bytes, without lowering the workgroups per CU.  Q2 FP64:
// SQ_LDS_BANK_CONFLICT 6.45M -> 4.33M cycles per apply.
struct LdsLayout { int a0, a1, a2, b0, b1, b2, SA, SB; };
constexpr LdsLayout lds_layout(int n1, int q, int es) {
    if (es == 8) {
        if (n1 == 2 && q == 3) return {2, 1, 0, 2, 0, 1, 13, 18};
        if (n1 == 3 && q == 3) return {0, 1, 2, 0, 2, 1, 27, 27};
        if (n1 == 3 && q == 4) return {0, 2, 1, 1, 0, 2, 36, 48};
        if (n1 == 4 && q == 4) return {1, 0, 2, 0, 1, 2, 64, 64};
        if (n1 == 4 && q == 5) return {1, 0, 2, 0, 2, 1, 85, 100};
        if (n1 == 5 && q == 5) return {0, 1, 2, 0, 2, 1, 125, 125};
        if (n1 == 5 && q == 6) return {2, 0, 1, 1, 2, 0, 153, 180};
    } else {
        if (n1 == 2 && q == 3) return {0, 1, 2, 0, 2, 1, 12, 18};
        if (n1 == 3 && q == 3) return {1, 0, 2, 0, 2, 1, 27, 27};
        if (n1 == 3 && q == 4) return {0, 2, 1, 1, 0, 2, 36, 48};
        if (n1 == 4 && q == 5) return {1, 0, 2, 0, 2, 1, 85, 100};
        if (n1 == 5 && q == 5) return {0, 1, 2, 0, 2, 1, 125, 125};
        if (n1 == 5 && q == 6) return {2, 0, 1, 0, 2, 1, 155, 181};
    }
    return {0, 1, 2, 0, 1, 2, n1 * n1 * q, n1 * q * q};
}
// offset of (i0, i1, i2) with sizes (S0, S1, S2) stored in dimension order P0, P1, P2
template <int P0, int P1, int P2, int S0, int S1, int S2>
__device__ __forceinline__ int off3(int i0, int i1, int i2) {
    const int idx[3] = {i0, i1, i2};
    constexpr int sz[3] = {S0, S1, S2};
    return (idx[P0] * sz[P1] + idx[P1]) * sz[P2] + idx[P2];
}

// 1 / d for d > 0: the hardware reciprocal refined by Newton steps (to about
// an ulp in either precision), in place of the longer IEEE division sequence.
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float recip(float d) {
    float r = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}

// QM, the quadrature-data mode: 0 general (6 diffusion values per point), 1
// Cartesian (one value per point and per-axis factors cf), 2 geometry on the
// fly (qd holds w kappa per point, cf the mesh vertices; J is rebuilt from the
// element's 8 vertices at every point, and the mass weight is rho w |J|).
// Waves per SIMD the compiler must allow (Q4 geometry kernels).  FP32 is
// held to 72 VGPRs, 7 waves instead of 6 at its own 80, without spills: 1.22
// -> 1.18 ms on 128^3; a cap of 8 waves spills and is slower.  FP64 is held
// to 128 VGPRs (4 waves): left free, the compiler takes 146 (3 waves, 2.14 ->
// 2.39 ms); 5 waves would spill 45 registers.
template <int N1, int Q, class T, int QM> constexpr int pa_min_waves() {
    return QM == 2 && N1 == 5 && Q == 5 ? (sizeof(T) == 4 ? 7 : 4) : 1;
}
template <int N1, int Q, class T, int FORM, int QM>
__global__ __launch_bounds__(ne_of(Q) * Q * Q, (pa_min_waves<N1, Q, T, QM>())) void k_pa_hex(
    const HexGrid g, const T *__restrict__ qd, const T *__restrict__ qm, const T *__restrict__ cf,
    const T *__restrict__ x, T *__restrict__ y, const int dirichlet, const int *skip, const T rho) {
    constexpr bool CART = QM == 1;
    constexpr bool GEO = QM == 2;
    static_assert(!GEO || (FORM & 2), "geometry on the fly serves the diffusion forms");
    constexpr int P = N1 - 1;
    constexpr int T2 = Q * Q;
    constexpr int NE = ne_of(Q);
    constexpr int C = combo(N1, Q);
    constexpr bool DIFF = (FORM & 2) != 0;
    constexpr bool MASS = (FORM & 1) != 0;
    constexpr int ND = DIFF ? NE : 1;
    // LDS: per element, A arrays hold (k, j, a) and B arrays (k, b, a) in the
    // index order and element stride of lds_layout (chosen against bank
    // conflicts of the five phases' access patterns).
    constexpr LdsLayout LL = lds_layout(N1, Q, (int)sizeof(T));
    // Geometry on the fly (measured on Q4 128^3, FP64 apply 3.21 -> 2.84 ms):
    // - ALIAS: the X-phase arrays (sA, sAg) and the Y-phase arrays (sBB, sBG,
    //   sGB) share one buffer; the Y and Y^T phases read all their inputs
    //   into registers and pass a barrier before they write, and the X^T phase
    //   hands its face values over in sF.  33.7 instead of 51.7 KB per
    //   workgroup at Q4 FP64, so 4 workgroups per CU instead of 3;
    // - the Z phase in three passes (below), 126 instead of 150 VGPRs, so 4
    //   waves per SIMD;
    // - EO: the contractions in the even-odd form (eo_contract), 947 instead
    //   of 1073 vector instructions per lane.
    constexpr bool ALIAS = GEO;
    // EO: the contractions in the even-odd form (eo_contract)
    constexpr bool EO = GEO;
    const T *E0 = eo_tab<T>(C, 0), *E1 = eo_tab<T>(C, 1), *E2 = eo_tab<T>(C, 2), *E3 = eo_tab<T>(C, 3);
    constexpr int SZX = NE * LL.SA + ND * LL.SA, SZY = NE * LL.SB + 2 * ND * LL.SB;
    __shared__ T sBuf[ALIAS ? (SZX > SZY ? SZX : SZY) : SZX + SZY];
    // X^T's face values: in the part of sBuf beyond the X arrays (unused by
    // X^T) when it fits
    constexpr int SZF = NE * N1 * N1;
    constexpr bool FIN = ALIAS && SZY >= SZX + SZF;
    __shared__ T sF_[ALIAS && !FIN ? SZF : 1];
    T *const sF = FIN ? sBuf + SZX : sF_;
    T *const sA = sBuf;                    // X out (B path), later F_B
    T *const sAg = sBuf + NE * LL.SA;      // X out (G path), later F_G
    T *const sBB = sBuf + (ALIAS ? 0 : SZX);  // Y out BB, later C_zeta+mass
    T *const sBG = sBB + NE * LL.SB;       // Y out BG, later C_xi
    T *const sGB = sBG + ND * LL.SB;       // Y out GB, later C_eta
    __shared__ T sC[GEO ? NE * 21 : 1];  // GEO: the map's monomial coefficients, component * 7 + m - 1
    auto iA = [](int e_, int k_, int j_, int a_) {
        return e_ * LL.SA + off3<LL.a0, LL.a1, LL.a2, N1, N1, Q>(k_, j_, a_);
    };
    auto iB = [](int e_, int k_, int b_, int a_) {
        return e_ * LL.SB + off3<LL.b0, LL.b1, LL.b2, N1, Q, Q>(k_, b_, a_);
    };

    if (skip && *skip) return;
    // Element group of this workgroup.  Geometry on the fly: XCD-contiguous
    // order (workgroup b runs on XCD b mod 8, and XCD k takes the groups
    // [k C, (k + 1) C) in order, C = gridDim.x / 8, the grid a multiple of 8):
    // neighbouring groups share x in one XCD's L2.  Q4 128^3, bitwise the
    // same output: FP64 2.12 -> 2.10-2.12 ms, FP32 1.18 -> 1.13 ms.  With
    // stored qdata the hardware order was faster (Q2 64^3: 76 against 81.5 us).
    const unsigned lb = GEO ? (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const T *Bt = Tab<T>::B(C);
    const T *Gt = Tab<T>::G(C);
    const int le = threadIdx.x / T2;
    const int t = threadIdx.x - le * T2;
    // the lane's coordinates in the (Q, Q) and (N1, N1) phase layouts
    const int tdq = t / Q, tq = t - tdq * Q;
    const int tdn = t / N1, tn = t - tdn * N1;
    // E < 2^31 (checked at creation), so 32-bit unsigned division is exact
    // and avoids the 64-bit division sequence
    const unsigned eu = lb * (unsigned)NE + (unsigned)le;
    const long long e = (long long)eu;
    const bool valid = e < g.E;
    int ex = 0, ey = 0, ez = 0;
    if (valid) {
        const unsigned r = udiv_m(eu, g.mx, g.shx);
        ex = (int)(eu - r * (unsigned)g.nx);
        const unsigned r2 = udiv_m(r, g.my, g.shy);
        ey = (int)(r - r2 * (unsigned)g.ny);
        ez = (int)r2;
    }

    // One stored value per point (Cartesian: w kappa |J|; geometry on the fly:
    // w kappa): the lane's Q values along its z-column, and (geometry) its
    // point coordinates and weights, are loaded here so that their latency
    // overlaps the X and Y phases instead of stalling the Z phase.
    constexpr size_t plane = (size_t)NE * T2;
    T sqv[(CART || GEO) && DIFF ? Q : 1];
    T xa = 0, yb = 0, wa = 0, wb = 0;
    if constexpr ((CART || GEO) && DIFF) {
        if (valid) {
            const size_t off1 = (size_t)lb * Q * plane + (size_t)le * T2 + t;
#pragma unroll
            for (int c = 0; c < Q; ++c) sqv[c] = qd[(size_t)c * plane + off1];
            if constexpr (GEO) {
                const int a = tq, b = tdq;
                xa = Tab<T>::X(Q)[a];
                yb = Tab<T>::X(Q)[b];
                wa = Tab<T>::W(Q)[a];
                wb = Tab<T>::W(Q)[b];
            }
        }
    }

    if constexpr (GEO) {
        // The trilinear map as a polynomial, X = sum_m c_m xi^m0 eta^m1 zeta^m2
        // (m0, m1, m2 the bits of m): c_m = 1/8 sum_v s(v)^m V_v over the
        // corners v = i + 2j + 4k with signs s(v) = (2i - 1, 2j - 1, 2k - 1),
        // a Walsh-Hadamard transform of the 8 vertices; one lane per
        // component.
        if (valid && t < 3) {
            const size_t sy = 3 * (size_t)(g.nx + 1), sz = sy * (size_t)(g.ny + 1);
            const T *p0 = cf + 3 * (size_t)ex + sy * (size_t)ey + sz * (size_t)ez + t;
            const T *p1 = p0 + sz;
            T h[8] = {p0[0], p0[3], p0[sy], p0[sy + 3], p1[0], p1[3], p1[sy], p1[sy + 3]};
#pragma unroll
            for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
                for (int v = 0; v < 8; ++v)
                    if (!(v & bit)) {
                        const T lo = h[v], hi = h[v | bit];
                        h[v] = hi + lo;
                        h[v | bit] = hi - lo;
                    }
#pragma unroll
            for (int m = 1; m < 8; ++m) sC[le * 21 + 7 * t + m - 1] = T(0.125) * h[m];
        }
    }

    // ---- X: lane (j,k) owns an x-line
    if (valid && t < N1 * N1) {
        const int j = tn, k = tdn;
        // 32-bit row index gy + Ny gz (Ny Nz < 2^32, checked at creation)
        const unsigned gy = (unsigned)(ey * P + j), gz = (unsigned)(ez * P + k);
        const unsigned Nyu = (unsigned)g.Ny, Nzu = (unsigned)g.Nz;
        const T *xr = x + (size_t)(gy + Nyu * gz) * (size_t)g.Nx + (size_t)(ex * P);
        const bool rowb = dirichlet && (gy == 0 || gy == Nyu - 1 || gz == 0 || gz == Nzu - 1);
        const bool lo = dirichlet && ex == 0, hi = dirichlet && ex == g.nx - 1;
        T u[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
            // every x entry exists, so the load is unconditional and the
            // essential DOFs are zeroed by a select
            const T v = xr[i];
            const bool b = rowb || (i == 0 && lo) || (i == P && hi);
            u[i] = b ? T(0) : v;
        }
        if constexpr (EO) {
            const EOs<N1, T> su = eo_split(u);
            T s[Q], sg[Q];
            eo_contract<Q, N1, 1>(E0, su, s);
            if constexpr (DIFF) eo_contract<Q, N1, -1>(E1, su, sg);
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                sA[iA(le, k, j, a)] = s[a];
                if constexpr (DIFF) sAg[iA(le, k, j, a)] = sg[a];
            }
        } else
#pragma unroll
        for (int a = 0; a < Q; ++a) {
            T s = 0, sg = 0;
#pragma unroll
            for (int i = 0; i < N1; ++i) {
                s += Bt[a * N1 + i] * u[i];
                if constexpr (DIFF) sg += Gt[a * N1 + i] * u[i];
            }
            sA[iA(le, k, j, a)] = s;
            if constexpr (DIFF) sAg[iA(le, k, j, a)] = sg;
        }
    }
    __syncthreads();

    // ---- Y: lane (qx,k)
    const bool ylane = valid && t < Q * N1;
    T A_[N1], Ag_[N1];
    if (ylane) {
        const int a = tq, k = tdq;
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            A_[j] = sA[iA(le, k, j, a)];
            if constexpr (DIFF) Ag_[j] = sAg[iA(le, k, j, a)];
        }
    }
    if constexpr (ALIAS) __syncthreads();
    if (ylane && EO) {
        const int a = tq, k = tdq;
        const EOs<N1, T> sa = eo_split(A_);
        T bb[Q], bg[Q], gb[Q];
        eo_contract<Q, N1, 1>(E0, sa, bb);
        if constexpr (DIFF) {
            const EOs<N1, T> sag = eo_split(Ag_);
            eo_contract<Q, N1, -1>(E1, sa, gb);
            eo_contract<Q, N1, 1>(E0, sag, bg);
        }
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            sBB[iB(le, k, b, a)] = bb[b];
            if constexpr (DIFF) { sBG[iB(le, k, b, a)] = bg[b]; sGB[iB(le, k, b, a)] = gb[b]; }
        }
    } else if (ylane) {
        const int a = tq, k = tdq;
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            T bb = 0, bg = 0, gb = 0;
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                bb += Bt[b * N1 + j] * A_[j];
                if constexpr (DIFF) {
                    bg += Bt[b * N1 + j] * Ag_[j];
                    gb += Gt[b * N1 + j] * A_[j];
                }
            }
            sBB[iB(le, k, b, a)] = bb;
            if constexpr (DIFF) { sBG[iB(le, k, b, a)] = bg; sGB[iB(le, k, b, a)] = gb; }
        }
    }
    __syncthreads();

    // ---- Z + pointwise + Z^T: lane (qx,qy) owns a z-column
    if (valid) {
        const int a = tq, b = tdq;
        T bb[N1], bg[N1], gb[N1];
        T c0[N1], c1[N1], c2[N1];
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            bb[k] = sBB[iB(le, k, b, a)];
            if constexpr (DIFF) { bg[k] = sBG[iB(le, k, b, a)]; gb[k] = sGB[iB(le, k, b, a)]; }
            c0[k] = c1[k] = c2[k] = 0;
        }
        const size_t offd = (size_t)lb * (CART || GEO ? 1 : 6) * Q * plane + (size_t)le * T2 + t;
        const size_t offm = (size_t)lb * Q * plane + (size_t)le * T2 + t;
        T fx = 0, fy = 0, fz = 0;  // Cartesian per-axis factors of this element
        if constexpr (DIFF && CART) {
            fx = cf[ex];
            fy = cf[g.nx + ey];
            fz = cf[g.nx + g.ny + ez];
        }
        // GEO: the Jacobian columns along this lane's z-column (xi_a, eta_b,
        // zeta): dX/dxi = Pa + Qa zeta, dX/deta = Pb + Qb zeta, dX/dzeta = Hc
        // (the trilinear map is linear in each coordinate)
        T Pa[3] = {0, 0, 0}, Qa[3] = {0, 0, 0}, Pb[3] = {0, 0, 0}, Qb[3] = {0, 0, 0}, Hc[3] = {0, 0, 0};
        T wab = 0;
        if constexpr (GEO) {
            // with X = sum_m c_m xi^m0 eta^m1 zeta^m2 (c_m = C[m - 1]):
            // dX/dxi = (c1 + c3 eta) + (c5 + c7 eta) zeta, dX/deta = (c2 + c3 xi)
            // + (c6 + c7 xi) zeta, dX/dzeta = c4 + c5 xi + (c6 + c7 xi) eta
            const T *Cm = sC + le * 21;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const T *c = Cm + 7 * i;
                Pa[i] = c[0] + c[2] * yb;
                Qa[i] = c[4] + c[6] * yb;
                Pb[i] = c[1] + c[2] * xa;
                Qb[i] = c[5] + c[6] * xa;
                Hc[i] = c[3] + c[4] * xa + Qb[i] * yb;
            }
            wab = (rho * wa) * wb;
        }
        if constexpr (GEO) {
            // Three passes over the column, so that the inputs bb, bg, gb and
            // the accumulators c0, c1, c2 are never live together (register
            // pressure sets the waves per SIMD here): the forward contraction
            // at all Q points, the pointwise operator in place, the transposed
            // contraction.  Each sum keeps the per-point loop's order.
            T G0[Q], G1[Q], G2[Q], UU[Q];
            if constexpr (EO) {
                const EOs<N1, T> sbb = eo_split(bb);
                eo_contract<Q, N1, 1>(E0, eo_split(bg), G0);
                eo_contract<Q, N1, 1>(E0, eo_split(gb), G1);
                eo_contract<Q, N1, -1>(E1, sbb, G2);
                if constexpr (MASS) eo_contract<Q, N1, 1>(E0, sbb, UU);
            } else
#pragma unroll
            for (int c = 0; c < Q; ++c) {
                T g0 = 0, g1 = 0, g2 = 0, uu = 0;
#pragma unroll
                for (int k = 0; k < N1; ++k) {
                    g0 += Bt[c * N1 + k] * bg[k];
                    g1 += Bt[c * N1 + k] * gb[k];
                    g2 += Gt[c * N1 + k] * bb[k];
                }
                if constexpr (MASS) {
#pragma unroll
                    for (int k = 0; k < N1; ++k) uu += Bt[c * N1 + k] * bb[k];
                }
                G0[c] = g0;
                G1[c] = g1;
                G2[c] = g2;
                UU[c] = uu;
            }
#pragma unroll
            for (int c = 0; c < Q; ++c) {
                const T zc = Tab<T>::X(Q)[c];
                const T ax = Pa[0] + Qa[0] * zc, ay = Pa[1] + Qa[1] * zc, az = Pa[2] + Qa[2] * zc;
                const T bx = Pb[0] + Qb[0] * zc, by = Pb[1] + Qb[1] * zc, bz = Pb[2] + Qb[2] * zc;
                const T cx = Hc[0], cy = Hc[1], cz = Hc[2];
                const T r0x = by * cz - bz * cy, r0y = bz * cx - bx * cz, r0z = bx * cy - by * cx;
                const T r1x = cy * az - cz * ay, r1y = cz * ax - cx * az, r1z = cx * ay - cy * ax;
                const T r2x = ay * bz - az * by, r2y = az * bx - ax * bz, r2z = ax * by - ay * bx;
                const T det = ax * r0x + ay * r0y + az * r0z;
                const T adet = fabs(det);
                const T f = sqv[c] * recip(adet);
                const T g0 = G0[c], g1 = G1[c], g2 = G2[c];
                const T hx = g0 * r0x + g1 * r1x + g2 * r2x;
                const T hy = g0 * r0y + g1 * r1y + g2 * r2y;
                const T hz = g0 * r0z + g1 * r1z + g2 * r2z;
                G0[c] = f * (r0x * hx + r0y * hy + r0z * hz);
                G1[c] = f * (r1x * hx + r1y * hy + r1z * hz);
                G2[c] = f * (r2x * hx + r2y * hy + r2z * hz);
                if constexpr (MASS) UU[c] = (wab * Tab<T>::W(Q)[c] * adet) * UU[c];
            }
            if constexpr (EO) {
                eo_contract<N1, Q, 1>(E2, eo_split(G0), c0);
                eo_contract<N1, Q, 1>(E2, eo_split(G1), c1);
                if constexpr (MASS) eo_contract2<N1, Q>(E2, eo_split(UU), E3, eo_split(G2), c2);
                else eo_contract<N1, Q, -1>(E3, eo_split(G2), c2);
            } else
#pragma unroll
            for (int c = 0; c < Q; ++c) {
#pragma unroll
                for (int k = 0; k < N1; ++k) {
                    const T bz = Bt[c * N1 + k];
                    c0[k] += bz * G0[c];
                    c1[k] += bz * G1[c];
                    c2[k] += Gt[c * N1 + k] * G2[c];
                    if constexpr (MASS) c2[k] += bz * UU[c];
                }
            }
        } else
#pragma unroll
        for (int c = 0; c < Q; ++c) {
            T v0 = 0, v1 = 0, v2 = 0, vm = 0;
            T adet = 0;  // GEO: |det J| at this point
            if constexpr (DIFF) {
                T g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
                for (int k = 0; k < N1; ++k) {
                    g0 += Bt[c * N1 + k] * bg[k];
                    g1 += Bt[c * N1 + k] * gb[k];
                    g2 += Gt[c * N1 + k] * bb[k];
                }
                if constexpr (CART) {
                    const T sq = sqv[c];
                    v0 = (sq * fx) * g0;
                    v1 = (sq * fy) * g1;
                    v2 = (sq * fz) * g2;
                } else if constexpr (GEO) {
                    // D = w kappa |J| J^-1 J^-T applied as f R (R^T g): the rows r_d
                    // of R = det(J) J^-1 are cross products of J's columns, and
                    // f = w kappa |J| / det^2 = w kappa / |det|
                    const T zc = Tab<T>::X(Q)[c];
                    const T ax = Pa[0] + Qa[0] * zc, ay = Pa[1] + Qa[1] * zc, az = Pa[2] + Qa[2] * zc;
                    const T bx = Pb[0] + Qb[0] * zc, by = Pb[1] + Qb[1] * zc, bz = Pb[2] + Qb[2] * zc;
                    const T cx = Hc[0], cy = Hc[1], cz = Hc[2];
                    const T r0x = by * cz - bz * cy, r0y = bz * cx - bx * cz, r0z = bx * cy - by * cx;
                    const T r1x = cy * az - cz * ay, r1y = cz * ax - cx * az, r1z = cx * ay - cy * ax;
                    const T r2x = ay * bz - az * by, r2y = az * bx - ax * bz, r2z = ax * by - ay * bx;
                    const T det = ax * r0x + ay * r0y + az * r0z;
                    adet = fabs(det);
                    const T f = sqv[c] * recip(adet);
                    const T hx = g0 * r0x + g1 * r1x + g2 * r2x;
                    const T hy = g0 * r0y + g1 * r1y + g2 * r2y;
                    const T hz = g0 * r0z + g1 * r1z + g2 * r2z;
                    v0 = f * (r0x * hx + r0y * hy + r0z * hz);
                    v1 = f * (r1x * hx + r1y * hy + r1z * hz);
                    v2 = f * (r2x * hx + r2y * hy + r2z * hz);
                } else {
                    const T d0 = qd[(size_t)(0 * Q + c) * plane + offd];
                    const T d1 = qd[(size_t)(1 * Q + c) * plane + offd];
                    const T d2 = qd[(size_t)(2 * Q + c) * plane + offd];
                    const T d3 = qd[(size_t)(3 * Q + c) * plane + offd];
                    const T d4 = qd[(size_t)(4 * Q + c) * plane + offd];
                    const T d5 = qd[(size_t)(5 * Q + c) * plane + offd];
                    v0 = d0 * g0 + d1 * g1 + d2 * g2;
                    v1 = d1 * g0 + d3 * g1 + d4 * g2;
                    v2 = d2 * g0 + d4 * g1 + d5 * g2;
                }
            }
            if constexpr (MASS) {
                T uu = 0;
#pragma unroll
                for (int k = 0; k < N1; ++k) uu += Bt[c * N1 + k] * bb[k];
                if constexpr (GEO) vm = (wab * Tab<T>::W(Q)[c] * adet) * uu;
                else vm = qm[(size_t)c * plane + offm] * uu;
            }
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                const T bz = Bt[c * N1 + k];
                if constexpr (DIFF) {
                    c0[k] += bz * v0;
                    c1[k] += bz * v1;
                    c2[k] += Gt[c * N1 + k] * v2;
                }
                if constexpr (MASS) c2[k] += bz * vm;
            }
        }
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            sBB[iB(le, k, b, a)] = c2[k];
            if constexpr (DIFF) { sBG[iB(le, k, b, a)] = c0[k]; sGB[iB(le, k, b, a)] = c1[k]; }
        }
    }
    __syncthreads();

    // ---- Y^T: lane (qx,k)
    T C0[Q], C1[Q], C2[Q];
    if (ylane) {
        const int a = tq, k = tdq;
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            C2[b] = sBB[iB(le, k, b, a)];
            if constexpr (DIFF) { C0[b] = sBG[iB(le, k, b, a)]; C1[b] = sGB[iB(le, k, b, a)]; }
        }
    }
    if constexpr (ALIAS) __syncthreads();
    if (ylane && EO) {
        const int a = tq, k = tdq;
        T fg[N1], fb[N1];
        if constexpr (DIFF) {
            eo_contract<N1, Q, 1>(E2, eo_split(C0), fg);
            eo_contract2<N1, Q>(E2, eo_split(C2), E3, eo_split(C1), fb);
        } else {
            eo_contract<N1, Q, 1>(E2, eo_split(C2), fb);
        }
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            sA[iA(le, k, j, a)] = fb[j];
            if constexpr (DIFF) sAg[iA(le, k, j, a)] = fg[j];
        }
    } else if (ylane) {
        const int a = tq, k = tdq;
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            T fg = 0, fb = 0;
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                fb += Bt[b * N1 + j] * C2[b];
                if constexpr (DIFF) {
                    fg += Bt[b * N1 + j] * C0[b];
                    fb += Gt[b * N1 + j] * C1[b];
                }
            }
            sA[iA(le, k, j, a)] = fb;
            if constexpr (DIFF) sAg[iA(le, k, j, a)] = fg;
        }
    }
    __syncthreads();

    // ---- X^T: lane (j,k) forms its N1 values of the row; the i = 0 value goes
    // through LDS (sBB is free now) to the lower neighbour in the workgroup,
    // which writes the sum of the shared x-face DOF.
    T sv[N1];
    const bool xt = valid && t < N1 * N1;
    const int jx = tn, kx = tdn;
    if (xt) {
        T FG[Q], FB[Q];
#pragma unroll
        for (int a = 0; a < Q; ++a) {
            FB[a] = sA[iA(le, kx, jx, a)];
            if constexpr (DIFF) FG[a] = sAg[iA(le, kx, jx, a)];
        }
        if constexpr (EO) {
            if constexpr (DIFF) eo_contract2<N1, Q>(E2, eo_split(FB), E3, eo_split(FG), sv);
            else eo_contract<N1, Q, 1>(E2, eo_split(FB), sv);
        } else
#pragma unroll
        for (int i = 0; i < N1; ++i) {
            T s = 0;
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                s += Bt[a * N1 + i] * FB[a];
                if constexpr (DIFF) s += Gt[a * N1 + i] * FG[a];
            }
            sv[i] = s;
        }
        if constexpr (ALIAS) sF[le * N1 * N1 + t] = sv[0];
        else sBB[iB(le, kx, jx, 0)] = sv[0];
    }
    __syncthreads();
    if (xt) {
        // 32-bit element-vector row (ny nz N1^2 < 2^32, checked at creation)
        const unsigned mrow = (((unsigned)ez * (unsigned)g.ny + (unsigned)ey) * N1 + kx) * N1 + jx;
        T *dst = y + (size_t)mrow * (size_t)g.Nxp + (size_t)(ex * P);
        if (ex == 0) dst[0] = sv[0];
        else if (le == 0) y[seam_index(g, N1, lb, jx, kx)] = sv[0];
#pragma unroll
        for (int i = 1; i < P; ++i) dst[i] = sv[i];
        // the next element is in this workgroup (and row) unless this one ends
        // the row or the workgroup
        if (ex < g.nx - 1 && le + 1 < NE)
            dst[P] = sv[P] + (ALIAS ? sF[(le + 1) * N1 * N1 + t] : sBB[iB(le + 1, kx, jx, 0)]);
        else dst[P] = sv[P];
    }
}

// Thread-per-element diffusion apply for low order (Q2 with 3^3 points, used
// with Cartesian qdata, where the workgroup kernel is bound by its LDS
// hand-offs rather than by HBM): one
// thread holds its element's N1^3 values in registers through all five
// contraction phases (no LDS, no barriers), reads its qdata in the
// wavefront-contiguous qd_index_tpe layout, and writes the same x-merged
// element vector as k_pa_hex with groups of 64 elements (one wavefront):
// the x-face value shared with the next element of the wavefront and row
// comes from that lane by a shuffle (lower element first), and the first
// element of a wavefront with ex > 0 stores its i = 0 value in the seam
// area, which the gather adds.
template <int N1, int Q, class T, bool CART>
__global__ __launch_bounds__(256) void k_pa_hex_tpe(const HexGrid g, const T *__restrict__ qd,
                                                    const T *__restrict__ cf, const T *__restrict__ x,
                                                    T *__restrict__ y, const int dirichlet, const int *skip) {
    constexpr int P = N1 - 1;
    constexpr int C = combo(N1, Q);
    if (skip && *skip) return;
    const T *Bt = Tab<T>::B(C);
    const T *Gt = Tab<T>::G(C);
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    const bool valid = e < g.E;
    const int lane = (int)(threadIdx.x & 63);
    int ex = 0, ey = 0, ez = 0;
    if (valid) {  // E < 2^31: 32-bit unsigned division is exact
        const unsigned eu = (unsigned)e, nxu = (unsigned)g.nx, nyu = (unsigned)g.ny;
        const unsigned r = eu / nxu;
        ex = (int)(eu - r * nxu);
        const unsigned r2 = r / nyu;
        ey = (int)(r - r2 * nyu);
        ez = (int)r2;
    }
    T yv[N1][N1][N1];  // [k][j][i]
    if (valid) {
        // ---- X: contract the element's x-lines
        T A[N1][N1][Q], Ag[N1][N1][Q];
#pragma unroll
        for (int k = 0; k < N1; ++k)
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                const long long gy = (long long)ey * P + j, gz = (long long)ez * P + k;
                const long long base = g.Nx * (gy + g.Ny * gz) + (long long)ex * P;
                const bool rowb = dirichlet && (gy == 0 || gy == g.Ny - 1 || gz == 0 || gz == g.Nz - 1);
                T u[N1];
#pragma unroll
                for (int i = 0; i < N1; ++i) {
                    const long long gx = (long long)ex * P + i;
                    const bool b = rowb || (dirichlet && (gx == 0 || gx == g.Nx - 1));
                    u[i] = b ? T(0) : x[base + i];
                }
#pragma unroll
                for (int a = 0; a < Q; ++a) {
                    T s = 0, sg = 0;
#pragma unroll
                    for (int i = 0; i < N1; ++i) {
                        s += Bt[a * N1 + i] * u[i];
                        sg += Gt[a * N1 + i] * u[i];
                    }
                    A[k][j][a] = s;
                    Ag[k][j][a] = sg;
                }
            }
        // ---- Y
        T BB[N1][Q][Q], BG[N1][Q][Q], GB[N1][Q][Q];
#pragma unroll
        for (int k = 0; k < N1; ++k)
#pragma unroll
            for (int a = 0; a < Q; ++a)
#pragma unroll
                for (int b = 0; b < Q; ++b) {
                    T bb = 0, bg = 0, gb = 0;
#pragma unroll
                    for (int j = 0; j < N1; ++j) {
                        bb += Bt[b * N1 + j] * A[k][j][a];
                        bg += Bt[b * N1 + j] * Ag[k][j][a];
                        gb += Gt[b * N1 + j] * A[k][j][a];
                    }
                    BB[k][b][a] = bb;
                    BG[k][b][a] = bg;
                    GB[k][b][a] = gb;
                }
        // ---- Z + pointwise + Z^T, column by column
        T fx = 0, fy = 0, fz = 0;
        if constexpr (CART) {
            fx = cf[ex];
            fy = cf[g.nx + ey];
            fz = cf[g.nx + g.ny + ez];
        }
#pragma unroll
        for (int b = 0; b < Q; ++b)
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                T c0[N1], c1[N1], c2[N1];
#pragma unroll
                for (int k = 0; k < N1; ++k) c0[k] = c1[k] = c2[k] = 0;
#pragma unroll
                for (int c = 0; c < Q; ++c) {
                    T g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
                    for (int k = 0; k < N1; ++k) {
                        g0 += Bt[c * N1 + k] * BG[k][b][a];
                        g1 += Bt[c * N1 + k] * GB[k][b][a];
                        g2 += Gt[c * N1 + k] * BB[k][b][a];
                    }
                    const int qp = a + Q * (b + Q * c);
                    T v0, v1, v2;
                    if constexpr (CART) {
                        const T sq = qd[qd_index_tpe(Q, 1, e, 0, qp)];
                        v0 = (sq * fx) * g0;
                        v1 = (sq * fy) * g1;
                        v2 = (sq * fz) * g2;
                    } else {
                        const T d0 = qd[qd_index_tpe(Q, 6, e, 0, qp)], d1 = qd[qd_index_tpe(Q, 6, e, 1, qp)];
                        const T d2 = qd[qd_index_tpe(Q, 6, e, 2, qp)], d3 = qd[qd_index_tpe(Q, 6, e, 3, qp)];
                        const T d4 = qd[qd_index_tpe(Q, 6, e, 4, qp)], d5 = qd[qd_index_tpe(Q, 6, e, 5, qp)];
                        v0 = d0 * g0 + d1 * g1 + d2 * g2;
                        v1 = d1 * g0 + d3 * g1 + d4 * g2;
                        v2 = d2 * g0 + d4 * g1 + d5 * g2;
                    }
#pragma unroll
                    for (int k = 0; k < N1; ++k) {
                        c0[k] += Bt[c * N1 + k] * v0;
                        c1[k] += Bt[c * N1 + k] * v1;
                        c2[k] += Gt[c * N1 + k] * v2;
                    }
                }
#pragma unroll
                for (int k = 0; k < N1; ++k) {
                    BG[k][b][a] = c0[k];
                    GB[k][b][a] = c1[k];
                    BB[k][b][a] = c2[k];
                }
            }
        // ---- Y^T
        T FB[N1][N1][Q], FG[N1][N1][Q];
#pragma unroll
        for (int k = 0; k < N1; ++k)
#pragma unroll
            for (int a = 0; a < Q; ++a)
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    T fg = 0, fb = 0;
#pragma unroll
                    for (int b = 0; b < Q; ++b) {
                        fb += Bt[b * N1 + j] * BB[k][b][a];
                        fg += Bt[b * N1 + j] * BG[k][b][a];
                        fb += Gt[b * N1 + j] * GB[k][b][a];
                    }
                    FB[k][j][a] = fb;
                    FG[k][j][a] = fg;
                }
        // ---- X^T
#pragma unroll
        for (int k = 0; k < N1; ++k)
#pragma unroll
            for (int j = 0; j < N1; ++j)
#pragma unroll
                for (int i = 0; i < N1; ++i) {
                    T s = 0;
#pragma unroll
                    for (int a = 0; a < Q; ++a) {
                        s += Bt[a * N1 + i] * FB[k][j][a];
                        s += Gt[a * N1 + i] * FG[k][j][a];
                    }
                    yv[k][j][i] = s;
                }
    } else {
#pragma unroll
        for (int k = 0; k < N1; ++k)
#pragma unroll
            for (int j = 0; j < N1; ++j)
#pragma unroll
                for (int i = 0; i < N1; ++i) yv[k][j][i] = 0;
    }
    // ---- y-merge (g.ymerge: nx = 64 or 128, so a wavefront lies in one element
    // row, rows tile the 256-element block, and the y-neighbour e + nx is
    // wavefront w + nx/64 of the same block when that exists): the upper face j = P goes through
    // LDS to the next row's wavefront, which adds it to its j = 0 face (lower
    // element first) and writes it; the lower element writes no j = P rows.
    bool yhi = false;
    if (g.ymerge) {
        __shared__ T sy[4][N1][N1][64];  // [wavefront][k][i][lane], the j = P face
        const int w = (int)(threadIdx.x >> 6), m = g.nx >> 6;
        yhi = valid && w + m < 4 && ey < g.ny - 1;
        const bool ylo = valid && w >= m && ey > 0;
        if (yhi) {
#pragma unroll
            for (int k = 0; k < N1; ++k)
#pragma unroll
                for (int i = 0; i < N1; ++i) sy[w][k][i][lane] = yv[k][P][i];
        }
        __syncthreads();
        if (ylo) {
#pragma unroll
            for (int k = 0; k < N1; ++k)
#pragma unroll
                for (int i = 0; i < N1; ++i) yv[k][0][i] = sy[w - m][k][i][lane] + yv[k][0][i];
        }
    }
    // ---- write the x-merged rows; every lane takes part in the shuffles
    const bool has_next = valid && lane < 63 && ex < g.nx - 1 && e + 1 < g.E;
#pragma unroll
    for (int k = 0; k < N1; ++k)
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            const T next0 = __shfl_down(yv[k][j][0], 1, 64);
            if (valid && !(j == P && yhi)) {
                T *dst = y + mrow_index(g, N1, ey, ez, j, k) + (size_t)ex * P;
                if (ex == 0) dst[0] = yv[k][j][0];
                else if (lane == 0) y[seam_index(g, N1, e >> 6, j, k)] = yv[k][j][0];
#pragma unroll
                for (int i = 1; i < P; ++i) dst[i] = yv[k][j][i];
                dst[P] = has_next ? yv[k][j][P] + next0 : yv[k][j][P];
            }
        }
}

// Value of lane SRC (0..3) of the calling lane's quad (DPP quad_perm broadcast).
template <int SRC, class T>
__device__ __forceinline__ T quad_bcast(T v) {
    constexpr int ctrl = SRC * 0x55;  // quad_perm [SRC, SRC, SRC, SRC]
    if constexpr (sizeof(T) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, ctrl, 0xf, 0xf, false);
        const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), ctrl, 0xf, 0xf, false);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xf, 0xf, false));
    }
}

// sum over the quad's four lanes of w[l] * (lane l's v), lane 0 first
template <class T>
__device__ __forceinline__ T quad_dot(const T (&w)[4], T v) {
    T s = 0;
    s += w[0] * quad_bcast<0>(v);
    s += w[1] * quad_bcast<1>(v);
    s += w[2] * quad_bcast<2>(v);
    s += w[3] * quad_bcast<3>(v);
    return s;
}

// Quad-per-element diffusion apply for Q3 with 4^3 points and Cartesian qdata
// (N1 = Q = 4), where the workgroup kernel is bound by its LDS hand-offs and a
// whole element does not fit one thread's registers: an element is a quad of
// lanes, lane L holding node layer k = L through the x and y contractions and
// their transposes, and quadrature layer c = L through the pointwise step.  The
// z contraction and its transpose exchange the quad's four layers by DPP
// broadcasts (no LDS, no barriers).  It writes the workgroup kernel's x-merged
// element vector (seams every 16 elements, one wavefront); the x-face shared
// with the next element of the wavefront and row comes from that quad by a
// shuffle (lower element first).  qdata in the qd_index_tpe layout.
template <class T>
__global__ __launch_bounds__(256) void k_pa_hex_quad(const HexGrid g, const T *__restrict__ qd,
                                                     const T *__restrict__ cf, const T *__restrict__ x,
                                                     T *__restrict__ y, const int dirichlet, const int *skip) {
    constexpr int N1 = 4, Q = 4, P = 3;
    constexpr int C = combo(N1, Q);
    if (skip && *skip) return;
    const T *Bt = Tab<T>::B(C);
    const T *Gt = Tab<T>::G(C);
    const long long e = (long long)blockIdx.x * 64 + (threadIdx.x >> 2);
    const int L = (int)(threadIdx.x & 3);
    const int lane = (int)(threadIdx.x & 63);
    const bool valid = e < g.E;  // uniform over the quad
    int ex = 0, ey = 0, ez = 0;
    if (valid) {  // E < 2^31: 32-bit unsigned division is exact
        const unsigned eu = (unsigned)e, nxu = (unsigned)g.nx, nyu = (unsigned)g.ny;
        const unsigned r = eu / nxu;
        ex = (int)(eu - r * nxu);
        const unsigned r2 = r / nyu;
        ey = (int)(r - r2 * nyu);
        ez = (int)r2;
    }
    T yv[N1][N1];  // [j][i] of layer k = L
    if (valid) {
        // ---- X on layer k = L
        T A[N1][Q], Ag[N1][Q];
        const long long gz = (long long)ez * P + L;
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            const long long gy = (long long)ey * P + j;
            const long long base = g.Nx * (gy + g.Ny * gz) + (long long)ex * P;
            const bool rowb = dirichlet && (gy == 0 || gy == g.Ny - 1 || gz == 0 || gz == g.Nz - 1);
            T u[N1];
#pragma unroll
            for (int i = 0; i < N1; ++i) {
                const long long gx = (long long)ex * P + i;
                const bool b = rowb || (dirichlet && (gx == 0 || gx == g.Nx - 1));
                u[i] = b ? T(0) : x[base + i];
            }
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                T s = 0, sg = 0;
#pragma unroll
                for (int i = 0; i < N1; ++i) {
                    s += Bt[a * N1 + i] * u[i];
                    sg += Gt[a * N1 + i] * u[i];
                }
                A[j][a] = s;
                Ag[j][a] = sg;
            }
        }
        // ---- Y on layer k = L
        T BB[Q][Q], BG[Q][Q], GB[Q][Q];  // [b][a]
#pragma unroll
        for (int a = 0; a < Q; ++a)
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                T bb = 0, bg = 0, gb = 0;
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    bb += Bt[b * N1 + j] * A[j][a];
                    bg += Bt[b * N1 + j] * Ag[j][a];
                    gb += Gt[b * N1 + j] * A[j][a];
                }
                BB[b][a] = bb;
                BG[b][a] = bg;
                GB[b][a] = gb;
            }
        // ---- Z (row c = L of the tables) + pointwise + Z^T (column k = L)
        T Bc[N1], Gc[N1], Bk[Q], Gk[Q];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            Bc[t] = Bt[L * N1 + t];
            Gc[t] = Gt[L * N1 + t];
            Bk[t] = Bt[t * N1 + L];
            Gk[t] = Gt[t * N1 + L];
        }
        const T fx = cf[ex], fy = cf[g.nx + ey], fz = cf[g.nx + g.ny + ez];
#pragma unroll
        for (int b = 0; b < Q; ++b)
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                const T g0 = quad_dot<T>(Bc, BG[b][a]);
                const T g1 = quad_dot<T>(Bc, GB[b][a]);
                const T g2 = quad_dot<T>(Gc, BB[b][a]);
                const T sq = qd[qd_index_tpe(Q, 1, e, 0, a + Q * (b + Q * L))];
                const T v0 = (sq * fx) * g0, v1 = (sq * fy) * g1, v2 = (sq * fz) * g2;
                BG[b][a] = quad_dot<T>(Bk, v0);
                GB[b][a] = quad_dot<T>(Bk, v1);
                BB[b][a] = quad_dot<T>(Gk, v2);
            }
        // ---- Y^T
        T FB[N1][Q], FG[N1][Q];
#pragma unroll
        for (int a = 0; a < Q; ++a)
#pragma unroll
            for (int j = 0; j < N1; ++j) {
                T fg = 0, fb = 0;
#pragma unroll
                for (int b = 0; b < Q; ++b) {
                    fb += Bt[b * N1 + j] * BB[b][a];
                    fg += Bt[b * N1 + j] * BG[b][a];
                    fb += Gt[b * N1 + j] * GB[b][a];
                }
                FB[j][a] = fb;
                FG[j][a] = fg;
            }
        // ---- X^T
#pragma unroll
        for (int j = 0; j < N1; ++j)
#pragma unroll
            for (int i = 0; i < N1; ++i) {
                T s = 0;
#pragma unroll
                for (int a = 0; a < Q; ++a) {
                    s += Bt[a * N1 + i] * FB[j][a];
                    s += Gt[a * N1 + i] * FG[j][a];
                }
                yv[j][i] = s;
            }
    } else {
#pragma unroll
        for (int j = 0; j < N1; ++j)
#pragma unroll
            for (int i = 0; i < N1; ++i) yv[j][i] = 0;
    }
    // ---- write the x-merged rows of layer L; every lane takes part in the shuffles
    const bool has_next = valid && lane < 60 && ex < g.nx - 1 && e + 1 < g.E;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const T next0 = __shfl_down(yv[j][0], 4, 64);
        if (valid) {
            T *dst = y + mrow_index(g, N1, ey, ez, j, L) + (size_t)ex * P;
            if (ex == 0) dst[0] = yv[j][0];
            else if (lane < 4) y[seam_index(g, N1, e >> 4, j, L)] = yv[j][0];
#pragma unroll
            for (int i = 1; i < P; ++i) dst[i] = yv[j][i];
            dst[P] = has_next ? yv[j][P] + next0 : yv[j][P];
        }
    }
}

template <class T> struct Pair;
template <> struct Pair<double> { using type = double2; };
template <> struct Pair<float> { using type = float2; };

// Element-vector -> DOF values, one x-segment per thread.  Segment (ex, gy,
// gz) holds the DOFs gx = ex P + i, i = 0..P-1, of DOF row (gy, gz), and the
// row's last DOF gx = Nx - 1 too when ex = nx - 1.  Each DOF sums its y/z
// candidate rows (1, 2 or 4) in a fixed order -- z candidate, then y, lower
// element first -- each row already holding the x-neighbours' sum (lower
// first; at a workgroup seam the gather forms that sum from the seam value),
// so applies are bitwise reproducible without atomics.  The row candidates,
// divisions and essential-DOF tests are shared by the segment's DOFs.
// Essential DOFs take x.  Returns the DOF count.
// I: index type of the element vector (uint32_t when it has < 2^31 entries).
// XESS: the essential DOFs of the segment take x (y = x there); otherwise x is
// not read and their bits are set in *ess instead.
template <class T, int P, class I, bool XESS = true>
__device__ __forceinline__ int gather_segment(const HexGrid &g, const T *__restrict__ ye,
                                              const T *__restrict__ x, int dirichlet, int ex, int gy,
                                              int gz, size_t row0, T (&out)[P + 1], unsigned *ess = nullptr) {
    constexpr int N1 = P + 1;
    const int cnt = ex == g.nx - 1 ? P + 1 : P;
    // y and z candidates: [0] = (e0 - 1, local P) on an element face, [1] = (e0, i0)
    int ce[2][2], ci[2][2];
    bool cv[2][2];
    const int gg[2] = {gy, gz}, ne[2] = {g.ny, g.nz};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const int e0 = gg[d] / P, i0 = gg[d] - e0 * P;
        ce[d][0] = e0 - 1; ci[d][0] = P; cv[d][0] = i0 == 0 && e0 > 0;
        ce[d][1] = e0;     ci[d][1] = i0; cv[d][1] = e0 < ne[d];
    }
    const bool rowbdr = dirichlet && (gy == 0 || gy == g.Ny - 1 || gz == 0 || gz == g.Nz - 1);
#pragma unroll
    for (int i = 0; i <= P; ++i) out[i] = T(0);
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            if (!(cv[1][c] && cv[0][b])) continue;
            if (g.ymerge && b == 0 && cv[0][1]) {
                // lower y candidate already added into the upper one's j = 0 row when
                // both elements lie in one 256-element block
                const size_t e_lo = ((size_t)ce[1][c] * g.ny + (size_t)ce[0][0]) * g.nx + (size_t)ex;
                if ((e_lo >> 8) == ((e_lo + (size_t)g.nx) >> 8)) continue;
            }
            const I erow = (I)ce[1][c] * (I)g.ny + (I)ce[0][b];
            const I row = ((erow * N1 + (I)ci[1][c]) * N1 + (I)ci[0][b]) * (I)g.Nxp;
            const T *rp = ye + row + (I)ex * P;  // this segment's values on the row
            T v[P];
            if constexpr (P % 2 == 0) {  // even row stride and even P: aligned pairs
                using T2 = typename Pair<T>::type;
#pragma unroll
                for (int i = 0; i < P; i += 2) {
                    const T2 w = *reinterpret_cast<const T2 *>(rp + i);
                    v[i] = w.x;
                    v[i + 1] = w.y;
                }
            } else {
#pragma unroll
                for (int i = 0; i < P; ++i) v[i] = rp[i];
            }
            T m0 = v[0];
            if (ex > 0) {
                const I e = erow * (I)g.nx + (I)ex;
                if (e % (I)g.ne == 0)
                    m0 += ye[(I)seam_base(g, N1) + ((e / (I)g.ne) * N1 + (I)ci[1][c]) * N1 + (I)ci[0][b]];
            }
            out[0] += m0;
#pragma unroll
            for (int i = 1; i < P; ++i) out[i] += v[i];
            if (ex == g.nx - 1) out[P] += rp[P];
        }
    const size_t base = row0 + (size_t)ex * P;
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        if (i >= cnt) break;
        const int gx = ex * P + i;
        if (rowbdr || (dirichlet && (gx == 0 || gx == g.Nx - 1))) {
            if constexpr (XESS) out[i] = x[base + i];
            else *ess |= 1u << i;
        }
    }
    return cnt;
}

// Segment coordinates of linear segment id sid = ex + nx (gy + Ny gz).
__device__ __forceinline__ void seg_coords(const HexGrid &g, size_t sid, int &ex, int &gy, int &gz) {
    if ((double)g.nx * g.Ny * g.Nz < 4294967296.0) {
        const unsigned u = (unsigned)sid, nx = (unsigned)g.nx, ny = (unsigned)g.Ny;
        const unsigned r = u / nx;
        ex = (int)(u - r * nx);
        gz = (int)(r / ny);
        gy = (int)(r - (unsigned)gz * ny);
    } else {
        const size_t r = sid / (size_t)g.nx;
        ex = (int)(sid - r * (size_t)g.nx);
        gz = (int)(r / (size_t)g.Ny);
        gy = (int)(r - (size_t)gz * g.Ny);
    }
}

// y (+)= gathered values; one thread per segment (consecutive threads along x).
// mode 0: y = s, 1: y += s.
template <class T, int P, class I>
__global__ __launch_bounds__(256) void k_hex_gather(const HexGrid g, const T *__restrict__ ye,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    int mode, int dirichlet, const int *skip) {
    if (skip && *skip) return;
    const size_t nseg = (size_t)g.nx * g.Ny * g.Nz;
    const size_t sid = blockIdx.x * (size_t)256 + threadIdx.x;
    if (sid >= nseg) return;
    int ex, gy, gz;
    seg_coords(g, sid, ex, gy, gz);
    const size_t row0 = (size_t)g.Nx * ((size_t)gy + (size_t)g.Ny * gz);
    T t[P + 1];
    const int cnt = gather_segment<T, P, I>(g, ye, x, dirichlet, ex, gy, gz, row0, t);
    T *yp = y + row0 + (size_t)ex * P;
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        if (i >= cnt) break;
        yp[i] = mode == 0 ? t[i] : yp[i] + t[i];
    }
}

// Gather fused with the Chebyshev step that consumes it (the solver's k_chebk):
// t = (A x)_i; d = c1 d + c2 inv (r - t); z += d; with last, partials of r.z
// over the fixed red_blocks(n) x 256 grid-stride grid of segments (and d is
// not stored: the polynomial ends there).  x may
// alias z: each thread reads x only at its own segment's DOFs, before it
// writes them.  FIRST: z and d equal (1/theta) inv r on entry (the element
// kernel applied A to d): that value is recomputed from r and inv instead of
// read, z is written, not read, and x is not read either (at the essential
// DOFs t = d).
template <class T, int P, class I, int BS, bool FIRST, bool LAST>
__global__ __launch_bounds__(BS) void k_hex_gather_cheb(
    const HexGrid g, const T *__restrict__ ye, const T *x, const T *__restrict__ r,
    const T *__restrict__ inv, T *z, T *__restrict__ d, double c1, double c2,
    double *__restrict__ part, int dirichlet, const int *skip, double inv_theta) {
    if (skip && *skip) return;
    const size_t nseg = (size_t)g.nx * g.Ny * g.Nz;
    double v[1] = {0.0};
    for (size_t sid = blockIdx.x * (size_t)BS + threadIdx.x; sid < nseg; sid += (size_t)gridDim.x * BS) {
        int ex, gy, gz;
        seg_coords(g, sid, ex, gy, gz);
        const size_t row0 = (size_t)g.Nx * ((size_t)gy + (size_t)g.Ny * gz);
        T t[P + 1];
        unsigned ess = 0;
        const int cnt = gather_segment<T, P, I, !FIRST>(g, ye, x, dirichlet, ex, gy, gz, row0, t, &ess);
        const size_t base = row0 + (size_t)ex * P;
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            if (i >= cnt) break;
            const size_t k = base + i;
            const T dk = FIRST ? cheb_first<T>(inv_theta, inv[k], r[k]) : d[k];
            const T ti = (FIRST && ((ess >> i) & 1u)) ? dk : t[i];
            const T di = (T)c1 * dk + (T)c2 * (inv[k] * (r[k] - ti));
            if constexpr (!LAST) d[k] = di;  // the last step's d is not read again
            const T zi = (FIRST ? dk : z[k]) + di;
            z[k] = zi;
            if constexpr (LAST) v[0] += (double)r[k] * (double)zi;
        }
    }
    if constexpr (LAST) {
        block_sum<1, BS>(v);
        if (threadIdx.x == 0) part[blockIdx.x] = v[0];
    }
}

// Gather fused with the dot product that follows it (the solver's k_dot):
// y = A x and partials of x.y in slot 0, on the solver's reduction grid.
template <class T, int P, class I, int BS>
__global__ __launch_bounds__(BS) void k_hex_gather_dot(const HexGrid g, const T *__restrict__ ye,
                                                        const T *__restrict__ x, T *__restrict__ y,
                                                        double *__restrict__ part, int dirichlet,
                                                        const int *skip) {
    if (skip && *skip) return;
    const size_t nseg = (size_t)g.nx * g.Ny * g.Nz;
    double v[1] = {0.0};
    for (size_t sid = blockIdx.x * (size_t)BS + threadIdx.x; sid < nseg; sid += (size_t)gridDim.x * BS) {
        int ex, gy, gz;
        seg_coords(g, sid, ex, gy, gz);
        const size_t row0 = (size_t)g.Nx * ((size_t)gy + (size_t)g.Ny * gz);
        T t[P + 1];
        const int cnt = gather_segment<T, P, I>(g, ye, x, dirichlet, ex, gy, gz, row0, t);
        const size_t base = row0 + (size_t)ex * P;
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            if (i >= cnt) break;
            y[base + i] = t[i];
            v[0] += (double)x[base + i] * (double)t[i];
        }
    }
    block_sum<1, BS>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// mode 0: y = bdr ? x : 0   (apply_zero)
// mode 1: y += bdr ? x : 0  (accumulate, only with essential DOFs)
// mode 2: y = bdr ? 1 : y   (diagonal fix-up)
// mode 3: y = bdr ? 0 : y   (right-hand side masking)
template <class T>
__global__ void k_hex_bc(const HexGrid g, const T *__restrict__ x, T *__restrict__ y, int mode,
                         int dirichlet, const int *skip) {
    if (skip && *skip) return;
    const size_t N = (size_t)g.Nx * g.Ny * g.Nz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        const long long gx = (long long)(i % g.Nx);
        const long long r = (long long)(i / g.Nx);
        const long long gy = r % g.Ny, gz = r / g.Ny;
        const bool b = dirichlet && on_bdr(g, gx, gy, gz);
        if (mode == 0) y[i] = b ? x[i] : T(0);
        else if (mode == 1) { if (b) y[i] += x[i]; }
        else if (mode == 2) { if (b) y[i] = T(1); }
        else { if (b) y[i] = T(0); }
    }
}

// Quadrature data from the trilinear geometry; one thread per (element, point).
// Computed in double, stored in the operator's precision.
struct CoefDev {
    int kind;
    double value;
    int nf;
    double scale, lo, hi;
    const double *wf, *bf, *qp;
};

__device__ inline double coef_eval(const CoefDev &c, const double xp[3], long long e, int qp, int Q3) {
    switch (c.kind) {
    case FEMGPU_COEF_CONST: return c.value;
    case FEMGPU_COEF_SINPROD:
        return 1.0 + 0.5 * sin(2.0 * M_PI * xp[0]) * sin(2.0 * M_PI * xp[1]) * sin(2.0 * M_PI * xp[2]);
    case FEMGPU_COEF_GRF: {
        double s = 0.0;
        for (int m = 0; m < c.nf; ++m)
            s += cos(c.wf[3 * m] * xp[0] + c.wf[3 * m + 1] * xp[1] + c.wf[3 * m + 2] * xp[2] + c.bf[m]);
        double k = exp(c.scale * s);
        return fmin(fmax(k, c.lo), c.hi);
    }
    default: return c.qp[(size_t)e * Q3 + qp];
    }
}

// The trilinear (HexQ1) map of element (ex, ey, ez) at reference point xi:
// J[i][d] = dX_i/dxi_d, the physical point xp; returns det J.  HexQ1 vertex
// order (lagrange/hex.rs:27-36).
__device__ __forceinline__ double hex_map(const HexGrid &g, const double *__restrict__ coords, int ex, int ey,
                                          int ez, const double (&xi)[3], double (&J)[3][3], double (&xp)[3]) {
    const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
              sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        xp[i] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) J[i][d] = 0.0;
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const long long vi = ex + (sx[v] > 0), vj = ey + (sy[v] > 0), vk = ez + (sz[v] > 0);
        const size_t gv = (size_t)vi + (size_t)(g.nx + 1) * ((size_t)vj + (size_t)(g.ny + 1) * vk);
        const double X0 = coords[3 * gv], X1 = coords[3 * gv + 1], X2 = coords[3 * gv + 2];
        const double fx = 1.0 + sx[v] * xi[0], fy = 1.0 + sy[v] * xi[1], fz = 1.0 + sz[v] * xi[2];
        const double phi = 0.125 * fx * fy * fz;
        const double gr[3] = {0.125 * sx[v] * fy * fz, 0.125 * sy[v] * fx * fz, 0.125 * sz[v] * fx * fy};
        const double X[3] = {X0, X1, X2};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            xp[i] += phi * X[i];
#pragma unroll
            for (int d = 0; d < 3; ++d) J[i][d] += X[i] * gr[d];
        }
    }
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// D = s J^-1 J^-T (6 symmetric components: 00, 01, 02, 11, 12, 22)
__device__ __forceinline__ void hex_metric(const double (&J)[3][3], double det, double s, double (&D)[6]) {
    const double id = 1.0 / det;
    double I[3][3];
    I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    I[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    D[0] = s * (I[0][0] * I[0][0] + I[0][1] * I[0][1] + I[0][2] * I[0][2]);
    D[1] = s * (I[0][0] * I[1][0] + I[0][1] * I[1][1] + I[0][2] * I[1][2]);
    D[2] = s * (I[0][0] * I[2][0] + I[0][1] * I[2][1] + I[0][2] * I[2][2]);
    D[3] = s * (I[1][0] * I[1][0] + I[1][1] * I[1][1] + I[1][2] * I[1][2]);
    D[4] = s * (I[1][0] * I[2][0] + I[1][1] * I[2][1] + I[1][2] * I[2][2]);
    D[5] = s * (I[2][0] * I[2][0] + I[2][1] * I[2][1] + I[2][2] * I[2][2]);
}

template <class T>
__global__ void k_hex_setup(const HexGrid g, int Q, const double *__restrict__ xq,
                            const double *__restrict__ wq, const double *__restrict__ coords,
                            CoefDev coef, double rho, int form, T *__restrict__ qd,
                            T *__restrict__ qm, int mass_only_weight, int cart, int tpe, int geo) {
    const int Q3 = Q * Q * Q, T2 = Q * Q;
    const size_t total = (size_t)g.E * Q3;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total;
         id += (size_t)gridDim.x * blockDim.x) {
        const long long e = (long long)(id / Q3);
        const int qp = (int)(id - (size_t)e * Q3);
        const int qx = qp % Q, qy = (qp / Q) % Q, qz = qp / T2;
        const int ex = (int)(e % g.nx);
        const int ey = (int)((e / g.nx) % g.ny), ez = (int)(e / ((long long)g.nx * g.ny));
        const double xi[3] = {xq[qx], xq[qy], xq[qz]};
        double J[3][3], xp[3];
        const double det = hex_map(g, coords, ex, ey, ez, xi, J, xp);
        const double wd = wq[qx] * wq[qy] * wq[qz] * fabs(det);
        const int t2 = qx + Q * qy;
        if ((form & 1) && qm) qm[qd_index(Q, 1, e, 0, qz, t2)] = (T)(mass_only_weight ? wd : wd * rho);
        if ((form & 2) && qd && geo) {
            // geometry on the fly: the element kernel forms |J| J^-1 J^-T itself
            qd[qd_index(Q, 1, e, 0, qz, t2)] = (T)(wq[qx] * wq[qy] * wq[qz] * coef_eval(coef, xp, e, qp, Q3));
        } else if ((form & 2) && qd && cart) {
            qd[tpe ? qd_index_tpe(Q, 1, e, 0, qp) : qd_index(Q, 1, e, 0, qz, t2)] =
                (T)(wd * coef_eval(coef, xp, e, qp, Q3));
        } else if ((form & 2) && qd) {
            double D[6];
            hex_metric(J, det, wd * coef_eval(coef, xp, e, qp, Q3), D);
#pragma unroll
            for (int comp = 0; comp < 6; ++comp)
                qd[tpe ? qd_index_tpe(Q, 6, e, comp, qp) : qd_index(Q, 6, e, comp, qz, t2)] = (T)D[comp];
        }
    }
}

// Assembled diagonal: one thread per (element, local DOF); setup-time only.
// geo: qd holds w kappa per point (geometry on the fly); the metric and the
// mass weight rho w |J| are formed here from the vertices (coords, xq, wq).
template <class T>
__global__ void k_hex_diag(const HexGrid g, int N1, int Q, const double *__restrict__ B,
                           const double *__restrict__ G, const T *__restrict__ qd,
                           const T *__restrict__ qm, const T *__restrict__ cf, int cart, int tpe, int form,
                           T *__restrict__ ye, int geo, const double *__restrict__ coords,
                           const double *__restrict__ xq, const double *__restrict__ wq, double rho) {
    // Element diagonals into the element vector; k_hex_gather then sums them
    // per DOF in a fixed order.
    const int L = N1 * N1 * N1;
    const size_t total = (size_t)g.E * L;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total;
         id += (size_t)gridDim.x * blockDim.x) {
        const long long e = (long long)(id / L);
        const int l = (int)(id - (size_t)e * L);
        const int i = l % N1, j = (l / N1) % N1, k = l / (N1 * N1);
        const int ex = (int)(e % g.nx);
        const int ey = (int)((e / g.nx) % g.ny), ez = (int)(e / ((long long)g.nx * g.ny));
        double s = 0.0;
        for (int c = 0; c < Q; ++c)
            for (int b = 0; b < Q; ++b)
                for (int a = 0; a < Q; ++a) {
                    const int t2 = a + Q * b;
                    const double Bi = B[a * N1 + i], Bj = B[b * N1 + j], Bk = B[c * N1 + k];
                    double gD[6] = {0, 0, 0, 0, 0, 0}, gwd = 0.0;  // geo: the metric and w |J| here
                    if (geo) {
                        const double xi[3] = {xq[a], xq[b], xq[c]};
                        double J[3][3], xp[3];
                        const double det = hex_map(g, coords, ex, ey, ez, xi, J, xp);
                        gwd = wq[a] * wq[b] * wq[c] * fabs(det);
                        if (form & 2) hex_metric(J, det, (double)qd[qd_index(Q, 1, e, 0, c, t2)] * fabs(det), gD);
                    }
                    if (form & 2) {
                        const double g0 = G[a * N1 + i] * Bj * Bk, g1 = Bi * G[b * N1 + j] * Bk,
                                     g2 = Bi * Bj * G[c * N1 + k];
                        double d[6];
                        if (geo) {
                            for (int m = 0; m < 6; ++m) d[m] = gD[m];
                        } else if (cart) {
                            const int qp = t2 + Q * Q * c;
                            const double sq = (double)qd[tpe ? qd_index_tpe(Q, 1, e, 0, qp) : qd_index(Q, 1, e, 0, c, t2)];
                            d[0] = sq * (double)cf[ex]; d[3] = sq * (double)cf[g.nx + ey];
                            d[5] = sq * (double)cf[g.nx + g.ny + ez];
                            d[1] = d[2] = d[4] = 0.0;
                        } else {
                            for (int m = 0; m < 6; ++m)
                                d[m] = (double)qd[tpe ? qd_index_tpe(Q, 6, e, m, t2 + Q * Q * c) : qd_index(Q, 6, e, m, c, t2)];
                        }
                        s += d[0] * g0 * g0 + d[3] * g1 * g1 + d[5] * g2 * g2 +
                             2.0 * (d[1] * g0 * g1 + d[2] * g0 * g2 + d[4] * g1 * g2);
                    }
                    if (form & 1) {
                        const double ph = Bi * Bj * Bk;
                        const double mw = geo ? gwd * rho : (double)qm[qd_index(Q, 1, e, 0, c, t2)];
                        s += mw * ph * ph;
                    }
                }
        ye[(size_t)e * L + l] = (T)s;
    }
}

// Per-element values old[e][k][j][i] -> the x-merged element vector, with the
// element kernel's sums and seams.  One thread per (row, gx).
template <class T>
__global__ void k_evec_merge(const HexGrid g, int N1, const T *__restrict__ old, T *__restrict__ ye) {
    const int P = N1 - 1, L = N1 * N1 * N1;
    const size_t total = seam_base(g, N1);
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total;
         id += (size_t)gridDim.x * blockDim.x) {
        const long long gx = (long long)(id % (size_t)g.Nxp);
        const size_t row = id / (size_t)g.Nxp;
        if (gx >= g.Nx) continue;  // row padding
        const int j = (int)(row % N1), k = (int)((row / N1) % N1);
        const long long erow = (long long)(row / ((size_t)N1 * N1));
        auto val = [&](long long ex, int i) { return old[(size_t)(erow * g.nx + ex) * L + (k * N1 + j) * N1 + i]; };
        long long ex = gx / P;
        const int i = (int)(gx - ex * P);
        T v;
        if (ex == g.nx) v = val(g.nx - 1, P);
        else if (i != 0 || ex == 0) v = val(ex, i);
        else {
            const long long e = erow * g.nx + ex;
            if (e % g.ne == 0) {
                v = val(ex - 1, P);
                ye[seam_index(g, N1, e / g.ne, j, k)] = val(ex, 0);
            } else {
                v = val(ex - 1, P) + val(ex, 0);
            }
        }
        ye[id] = v;
    }
}

__global__ void k_to_f32(size_t n, const double *__restrict__ a, float *__restrict__ b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

template <class T>
__global__ void k_random_field(size_t n, uint64_t seed, T *__restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = (T)(2.0 * philox_u01(seed, 7, i) - 1.0);
}

static unsigned grid1(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 1u << 20) b = 1u << 20;
    return (unsigned)(b ? b : 1);
}

// qm_mode (k_pa_hex's QM): 0 general qdata, 1 Cartesian (cf: per-axis
// factors), 2 geometry on the fly (cf: mesh vertices in T, rho: the constant
// mass coefficient)
template <int N1, int Q, class T>
static void launch_form(const HexGrid &g, int form, int qm_mode, const T *qd, const T *qm, const T *cf,
                        const T *x, T *y, int dir, const int *skip, T rho, hipStream_t s) {
    constexpr int NE = ne_of(Q);
    // geometry on the fly: a multiple of 8 workgroups (the XCD-contiguous
    // order); the extra ones hold no element
    const long long ng = (g.E + NE - 1) / NE;
    const dim3 grid((unsigned)(qm_mode == 2 ? (ng + 7) / 8 * 8 : ng)), block(NE * Q * Q);
#define L_(F_, M_) hipLaunchKernelGGL((k_pa_hex<N1, Q, T, F_, M_>), grid, block, 0, s, g, qd, qm, cf, x, y, dir, skip, rho)
    if (form == 1) L_(1, 0);
    else if (form == 2 && qm_mode == 1) L_(2, 1);
    else if (form == 2 && qm_mode == 2) L_(2, 2);
    else if (form == 2) L_(2, 0);
    else if (qm_mode == 1) L_(3, 1);
    else if (qm_mode == 2) L_(3, 2);
    else L_(3, 0);
#undef L_
}

// tpe: the operator keeps qdata in the qd_index_tpe layout, for the
// thread-per-element kernel (Q2 diffusion, 3^3 points) or the quad-per-element
// kernel (Q3 diffusion, 4^3 points, Cartesian qdata)
template <class T>
static bool launch_pa(int n1, int q, const HexGrid &g, int form, const T *qd, const T *qm, const T *cf,
                      const T *x, T *y, int dir, const int *skip, hipStream_t s, bool tpe = false,
                      bool geo = false, T rho = T(0)) {
    if (tpe && form == 2 && n1 == 3 && q == 3) {
        const dim3 grid((unsigned)((g.E + 255) / 256)), block(256);
        if (cf) hipLaunchKernelGGL((k_pa_hex_tpe<3, 3, T, true>), grid, block, 0, s, g, qd, cf, x, y, dir, skip);
        else hipLaunchKernelGGL((k_pa_hex_tpe<3, 3, T, false>), grid, block, 0, s, g, qd, cf, x, y, dir, skip);
        return true;
    }
    if (tpe && form == 2 && n1 == 4 && q == 4 && cf) {
        FG_REQUIRE(g.ne == 16 && !g.ymerge, FEMGPU_ERR_INVALID_ARG, "pa_hex: quad kernel needs 16-element groups");
        const dim3 grid((unsigned)((g.E + 63) / 64)), block(256);
        hipLaunchKernelGGL((k_pa_hex_quad<T>), grid, block, 0, s, g, qd, cf, x, y, dir, skip);
        return true;
    }
#define CASE(A, B)                                                                 \
    if (n1 == A && q == B) {                                                       \
        launch_form<A, B, T>(g, form, geo ? 2 : cf ? 1 : 0, qd, qm, cf, x, y, dir, skip, rho, s); \
        return true;                                                               \
    }
    CASE(2, 2) CASE(2, 3) CASE(3, 3) CASE(3, 4) CASE(4, 4) CASE(4, 5) CASE(5, 5) CASE(5, 6)
    CASE(6, 6)
#undef CASE
    return false;
}

// element vectors below 2^31 entries use 32-bit index arithmetic in the gathers
static inline bool evec_small(const HexGrid &g, int n1) { return (double)evec_count(g, n1) < 2147483648.0; }

template <class T, class I>
static void launch_gather_i(int p, const HexGrid &g, const T *ye, const T *x, T *y, int mode, int dir,
                            const int *skip, hipStream_t s) {
    const size_t nseg = (size_t)g.nx * g.Ny * g.Nz;
    const dim3 grid((unsigned)((nseg + 255) / 256));
    switch (p) {
    case 1: hipLaunchKernelGGL((k_hex_gather<T, 1, I>), grid, dim3(256), 0, s, g, ye, x, y, mode, dir, skip); break;
    case 2: hipLaunchKernelGGL((k_hex_gather<T, 2, I>), grid, dim3(256), 0, s, g, ye, x, y, mode, dir, skip); break;
    case 3: hipLaunchKernelGGL((k_hex_gather<T, 3, I>), grid, dim3(256), 0, s, g, ye, x, y, mode, dir, skip); break;
    case 4: hipLaunchKernelGGL((k_hex_gather<T, 4, I>), grid, dim3(256), 0, s, g, ye, x, y, mode, dir, skip); break;
    default: hipLaunchKernelGGL((k_hex_gather<T, 5, I>), grid, dim3(256), 0, s, g, ye, x, y, mode, dir, skip); break;
    }
}

template <class T>
static void launch_gather(int p, const HexGrid &g, const T *ye, const T *x, T *y, int mode, int dir,
                          const int *skip, hipStream_t s) {
    if (evec_small(g, p + 1)) launch_gather_i<T, uint32_t>(p, g, ye, x, y, mode, dir, skip, s);
    else launch_gather_i<T, size_t>(p, g, ye, x, y, mode, dir, skip, s);
}

static bool supported(int n1, int q) {
    return (n1 >= 2 && n1 <= 6) && (q == n1 || (q == n1 + 1 && n1 <= 5));
}

struct HexOp final : femgpu_op {
    femgpu_hex_desc desc{};
    HexGrid g{};
    DevBuf coords, qd, qm, kqp, grf_w, grf_b, xq, wq, Bd, Gd;
    DevBuf evec;  // x-merged element vector between the two apply kernels
    bool cart = false;           // axis-aligned box mesh: Cartesian qdata (one value per point)
    bool tpe = false;            // qd_index_tpe qdata layout and a thread-level element kernel
    bool quad = false;           // that kernel is k_pa_hex_quad (Q3), else k_pa_hex_tpe (Q2)
    bool geo = false;            // geometry on the fly: qd = w kappa per point, J from the vertices
    int geo_req = -1;            // requested at creation (-1: from FEMGPU_HEX_QDATA)
    std::vector<double> cf_host;  // Cartesian per-axis factors [nx | ny | nz]
    DevBuf cf;                    // the same, in the operator's type; with geo, the vertices in that type

    template <class T> const T *cf_ptr() const { return cart || geo ? (const T *)cf.ptr : nullptr; }
    // values of quadrature data per Gauss point (stored, read by every apply)
    int qdata_per_point() const {
        if (geo) return 1;
        return ((desc.form & 2) ? (cart ? 1 : 6) : 0) + ((desc.form & 1) ? 1 : 0);
    }
    // 0 general, 1 Cartesian, 2 geometry on the fly
    int qdata_mode() const { return geo ? 2 : cart ? 1 : 0; }

    const char *name() const override { return "pa_hex"; }

    // y (+)= A x: element kernel into the element vector, then the DOF gather
    // (which also applies the essential rows y = x).
    template <class T> void apply_t(const T *x, T *y, bool acc, const int *skip) {
        hipStream_t s = ctx->stream;
        const int dir = desc.dirichlet ? 1 : 0;
        if (prof_begin) FG_HIP(hipEventRecord(prof_begin, s));
        launch_pa<T>(desc.p + 1, desc.q, g, desc.form, qd.as<T>(), qm.as<T>(), cf_ptr<T>(), x, evec.as<T>(), dir,
                     skip, s, tpe, geo, (T)desc.rho);
        if (prof_mid) FG_HIP(hipEventRecord(prof_mid, s));
        launch_gather<T>(desc.p, g, evec.as<T>(), x, y, acc ? 1 : 0, dir, skip, s);
        if (prof_end) FG_HIP(hipEventRecord(prof_end, s));
    }

    void apply_dev(const void *x, void *y, bool acc, const int *skip) override {
        if (dtype == FEMGPU_F64) apply_t<double>((const double *)x, (double *)y, acc, skip);
        else apply_t<float>((const float *)x, (float *)y, acc, skip);
    }

    template <class T> void fused_t(const T *x, const T *r, const T *inv, T *z, T *d, double c1, double c2,
                                    int last, T *y, double *part, const int *skip, int first = 0,
                                    double inv_theta = 0.0) {
        hipStream_t s = ctx->stream;
        const int dir = desc.dirichlet ? 1 : 0;
        launch_pa<T>(desc.p + 1, desc.q, g, desc.form, qd.as<T>(), qm.as<T>(), cf_ptr<T>(), x, evec.as<T>(), dir,
                     skip, s, tpe, geo, (T)desc.rho);
        const dim3 grid((unsigned)red_blocks(n));
        const T *ye = evec.as<T>();
        // Threads per block on the fixed red_blocks(n) grid (the reduction
        // partials stay one per block): 1024 from 2^24 segments on (Q4 128^3:
        // 3.5% faster per PCG-Chebyshev iteration than 256, more waves in
        // flight for the dependent loads), 256 below (Q2 64^3: 1% faster);
        // FEMGPU_GATHER_BS=256/512/1024 overrides, for comparison
        static const int gbs_env = [] {
            const char *e = std::getenv("FEMGPU_GATHER_BS");
            return e ? std::atoi(e) : 0;
        }();
        const double nseg = (double)g.nx * (double)g.Ny * (double)g.Nz;
        const int gbs = gbs_env ? gbs_env : (nseg >= 16777216.0 ? 1024 : 256);
#define GF(P_)                                                                                              \
    do {                                                                                                    \
        if (small) GF2(P_, uint32_t);                                                                       \
        else GF2(P_, size_t);                                                                               \
    } while (0)
#define GF2(P_, I_)                                                                                         \
    do {                                                                                                    \
        if (gbs == 512) GF3(P_, I_, 512);                                                                  \
        else if (gbs == 1024) GF3(P_, I_, 1024);                                                           \
        else GF3(P_, I_, 256);                                                                             \
    } while (0)
#define GFC(P_, I_, BS_, F_, L_)                                                                            \
    hipLaunchKernelGGL((k_hex_gather_cheb<T, P_, I_, BS_, F_, L_>), grid, block, 0, s, g, ye, x, r, inv, z, d, c1, \
                       c2, part, dir, skip, inv_theta)
#define GF3(P_, I_, BS_)                                                                                    \
    do {                                                                                                    \
        const dim3 block(BS_);                                                                              \
        if (y) hipLaunchKernelGGL((k_hex_gather_dot<T, P_, I_, BS_>), grid, block, 0, s, g, ye, x, y, part, dir, skip); \
        else if (first && last) GFC(P_, I_, BS_, true, true);                                              \
        else if (first) GFC(P_, I_, BS_, true, false);                                                      \
        else if (last) GFC(P_, I_, BS_, false, true);                                                       \
        else GFC(P_, I_, BS_, false, false);                                                     \
    } while (0)
        const bool small = evec_small(g, desc.p + 1);
        switch (desc.p) {
        case 1: GF(1); break;
        case 2: GF(2); break;
        case 3: GF(3); break;
        case 4: GF(4); break;
        default: GF(5); break;
        }
#undef GFC
#undef GF3
#undef GF2
#undef GF
    }

    bool apply_cheb_dev(const void *x, const void *r, const void *inv, void *z, void *d, double c1, double c2,
                        int last, double *part, const int *skip, int first, double inv_theta) override {
        if (dtype == FEMGPU_F64)
            fused_t<double>((const double *)x, (const double *)r, (const double *)inv, (double *)z, (double *)d,
                            c1, c2, last, nullptr, part, skip, first, inv_theta);
        else
            fused_t<float>((const float *)x, (const float *)r, (const float *)inv, (float *)z, (float *)d, c1, c2,
                           last, nullptr, part, skip, first, inv_theta);
        return true;
    }

    bool apply_dot_dev(const void *x, void *y, double *part, const int *skip) override {
        if (dtype == FEMGPU_F64)
            fused_t<double>((const double *)x, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, (double *)y, part, skip);
        else
            fused_t<float>((const float *)x, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, (float *)y, part, skip);
        return true;
    }

    template <class T> void diag_t(T *d) {
        hipStream_t s = ctx->stream;
        const int n1 = desc.p + 1;
        const size_t L = (size_t)g.E * n1 * n1 * n1;
        DevBuf old(L * sizeof(T));
        hipLaunchKernelGGL(k_hex_diag<T>, dim3(grid1(L)), dim3(256), 0, s, g, n1, desc.q, Bd.as<double>(),
                           Gd.as<double>(), qd.as<T>(), qm.as<T>(), (const T *)cf.ptr, cart ? 1 : 0, tpe ? 1 : 0,
                           desc.form, old.as<T>(), geo ? 1 : 0, coords.as<double>(), xq.as<double>(),
                           wq.as<double>(), desc.rho);
        HexGrid gd = g;
        gd.ymerge = 0;  // k_evec_merge builds the element vector without the y-merge
        hipLaunchKernelGGL(k_evec_merge<T>, dim3(grid1(seam_base(gd, n1))), dim3(256), 0, s, gd, n1,
                           (const T *)old.as<T>(), evec.as<T>());
        launch_gather<T>(desc.p, gd, evec.as<T>(), d, d, 0, 0, nullptr, s);
        hipLaunchKernelGGL(k_hex_bc<T>, dim3(grid1(n)), dim3(256), 0, s, g, (const T *)d, d, 2,
                           desc.dirichlet, nullptr);
        FG_HIP(hipStreamSynchronize(s));  // before the temporary is freed
    }

    void diagonal_dev(void *d) override {
        if (dtype == FEMGPU_F64) diag_t<double>((double *)d);
        else diag_t<float>((float *)d);
    }

    // Builds qdata of the given form into (qd_, qm_) in type T.
    template <class T> void setup_t(int form, double rho, bool mass_weight_only, T *qd_, T *qm_) {
        CoefDev c{};
        const femgpu_coef &k = desc.kappa;
        c.kind = k.kind; c.value = k.value; c.nf = k.n_features; c.lo = k.lo; c.hi = k.hi;
        c.scale = k.kind == FEMGPU_COEF_GRF ? k.sigma * std::sqrt(2.0 / k.n_features) : 0.0;
        c.wf = grf_w.as<double>(); c.bf = grf_b.as<double>(); c.qp = kqp.as<double>();
        const size_t total = (size_t)g.E * desc.q * desc.q * desc.q;
        hipLaunchKernelGGL(k_hex_setup<T>, dim3(grid1(total)), dim3(256), 0, ctx->stream, g, desc.q,
                           xq.as<double>(), wq.as<double>(), coords.as<double>(), c, rho, form,
                           qd_, qm_, mass_weight_only ? 1 : 0, cart ? 1 : 0, tpe ? 1 : 0, geo ? 1 : 0);
        FG_HIP(hipGetLastError());
    }

    void build() {
        const int n1 = desc.p + 1, q = desc.q;
        double x[8], w[8], B[64], G[64];
        gauss_legendre(q, x, w);
        lagrange_1d(desc.p, q, x, B, G);
        xq.alloc(8 * q); wq.alloc(8 * q); Bd.alloc(8 * q * n1); Gd.alloc(8 * q * n1);
        hipStream_t s = ctx->stream;
        FG_HIP(hipMemcpyAsync(xq.ptr, x, 8 * q, hipMemcpyHostToDevice, s));
        FG_HIP(hipMemcpyAsync(wq.ptr, w, 8 * q, hipMemcpyHostToDevice, s));
        FG_HIP(hipMemcpyAsync(Bd.ptr, B, 8 * q * n1, hipMemcpyHostToDevice, s));
        FG_HIP(hipMemcpyAsync(Gd.ptr, G, 8 * q * n1, hipMemcpyHostToDevice, s));
        const size_t nq = qd_count_any(g.E, q);
        const size_t es = esize();
        // Geometry on the fly on meshes that are not axis-aligned boxes (diffusion
        // forms): faster than streaming the six-value qdata in both precisions
        // (Q4 128^3 perturbed: 4.21 -> 3.33 ms FP64, 2.42 -> 1.85 ms FP32) and a
        // seventh of its memory.  FEMGPU_HEX_QDATA=general keeps stored qdata.
        const char *qmode = std::getenv("FEMGPU_HEX_QDATA");
        const bool want_geo = geo_req >= 0 ? geo_req == 1 : !(qmode && std::strcmp(qmode, "general") == 0);
        geo = (desc.form & 2) && !cart && want_geo;
        const char *kmode = std::getenv("FEMGPU_HEX_KERNEL");
        // with the general six-value qdata the workgroup kernel is as fast (both HBM-bound) at higher occupancy
        const bool lds_forced = kmode && std::strcmp(kmode, "lds") == 0;
        tpe = n1 == 3 && q == 3 && desc.form == 2 && cart && !lds_forced;
        // Q3 with 4^3 points on Cartesian qdata: the quad-per-element kernel, with
        // the workgroup kernel's 16-element groups (g.ne = ne_of(4) = 16)
        quad = n1 == 4 && q == 4 && desc.form == 2 && cart && !lds_forced;
        if (quad) tpe = true;
        if (tpe && !quad) {
            g.ne = 64;                     // element-vector seams every wavefront
            // rows must tile the 256-element block (nx = 64 or 128): then every
            // wavefront of a row pairs with the row above alike, and no x-face is
            // shared by a merged and an unmerged element
            g.ymerge = (g.nx == 64 || g.nx == 128) ? 1 : 0;
        }
        if (desc.form & 2) qd.alloc((cart || geo ? 1 : 6) * nq * es);
        if ((desc.form & 1) && !geo) qm.alloc(nq * es);
        if (geo) {
            // the vertices in the operator's type, for the element kernel
            const size_t nv3 = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1) * 3;
            cf.alloc(nv3 * es);
            if (dtype == FEMGPU_F64) {
                FG_HIP(hipMemcpyAsync(cf.ptr, coords.ptr, nv3 * 8, hipMemcpyDeviceToDevice, s));
            } else {
                hipLaunchKernelGGL(k_to_f32, dim3(grid1(nv3)), dim3(256), 0, s, nv3, coords.as<double>(),
                                   cf.as<float>());
                FG_HIP(hipGetLastError());
            }
        }
        if (cart) {
            cf.alloc(cf_host.size() * es);
            if (dtype == FEMGPU_F64) {
                FG_HIP(hipMemcpyAsync(cf.ptr, cf_host.data(), cf_host.size() * 8, hipMemcpyHostToDevice, s));
            } else {
                std::vector<float> f(cf_host.begin(), cf_host.end());
                FG_HIP(hipMemcpyAsync(cf.ptr, f.data(), f.size() * 4, hipMemcpyHostToDevice, s));
                FG_HIP(hipStreamSynchronize(s));
            }
        }
        HexGrid gl = g;
        gl.ne = ne_of(q);  // the workgroup kernel's grouping (mass-weight applies)
        gl.ymerge = 0;
        evec.alloc(std::max(evec_count(g, n1), evec_count(gl, n1)) * es);
        if (dtype == FEMGPU_F64) setup_t<double>(desc.form, desc.rho, false, qd.as<double>(), qm.as<double>());
        else setup_t<float>(desc.form, desc.rho, false, qd.as<float>(), qm.as<float>());
        FG_HIP(hipStreamSynchronize(s));
    }
};

static void grf_features_host(const femgpu_coef &c, std::vector<double> &w, std::vector<double> &b) {
    // Random Fourier features of a squared-exponential covariance (Rahimi &
    // Recht 2007): w ~ N(0, I / length^2) by Box-Muller, b ~ U[0, 2 pi).
    w.assign(3 * c.n_features, 0.0);
    b.assign(c.n_features, 0.0);
    for (int m = 0; m < c.n_features; ++m) {
        for (int k = 0; k < 3; ++k) {
            const double u1 = philox_u01(c.seed, 10 + 2 * k, (uint64_t)m);
            const double u2 = philox_u01(c.seed, 11 + 2 * k, (uint64_t)m);
            const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
            w[3 * m + k] = r * std::cos(2.0 * M_PI * u2) / c.length;
        }
        b[m] = 2.0 * M_PI * philox_u01(c.seed, 20, (uint64_t)m);
    }
}

// Axis-aligned box mesh test on the (nx+1)(ny+1)(nz+1) x 3 vertex array:
// x of vertex (i, j, k) equals x of (i, 0, 0) for all j, k (likewise y, z),
// and no spacing is zero.  On success f = [1/J_xx^2 | 1/J_yy^2 | 1/J_zz^2]
// with J_dd = (spacing) / 2 on the reference cube [-1, 1]^3.
static bool cartesian_factors(int nx, int ny, int nz, const double *X, std::vector<double> &f) {
    const size_t sx = 1, sy = (size_t)nx + 1, sz = sy * ((size_t)ny + 1);
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const double *v = X + 3 * (i * sx + j * sy + k * sz);
                if (v[0] != X[3 * (i * sx)] || v[1] != X[3 * (j * sy) + 1] || v[2] != X[3 * (k * sz) + 2]) return false;
            }
    f.assign((size_t)nx + ny + nz, 0.0);
    const int n[3] = {nx, ny, nz};
    const size_t st[3] = {sx, sy, sz};
    size_t o = 0;
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < n[a]; ++i) {
            const double h = 0.5 * (X[3 * ((i + 1) * st[a]) + a] - X[3 * (i * st[a]) + a]);
            if (!(h != 0.0) || !std::isfinite(h)) return false;
            f[o++] = 1.0 / (h * h);
        }
    return true;
}

// geo_req: -1 takes the quadrature-data mode from FEMGPU_HEX_QDATA, 0 / 1
// asks for stored qdata / the geometry on the fly (a converted copy keeps its
// source's mode)
static HexOp *make_hex(femgpu_ctx *c, const femgpu_hex_desc &d, const double *coords_host,
                       const void *coords_dev, const void *kqp_dev, int geo_req = -1) {
    FG_REQUIRE(d.nx > 0 && d.ny > 0 && d.nz > 0, FEMGPU_ERR_INVALID_ARG, "pa_create_hex: empty mesh");
    FG_REQUIRE(d.p >= 1 && supported(d.p + 1, d.q), FEMGPU_ERR_UNSUPPORTED,
               "pa_create_hex: unsupported (p, q) = (" + std::to_string(d.p) + ", " +
                   std::to_string(d.q) + "); supported q = p+1 (p <= 5) or q = p+2 (p <= 4)");
    FG_REQUIRE(d.form >= 1 && d.form <= 3, FEMGPU_ERR_INVALID_ARG, "pa_create_hex: bad form");
    FG_REQUIRE(d.dtype == FEMGPU_F64 || d.dtype == FEMGPU_F32, FEMGPU_ERR_INVALID_ARG,
               "pa_create_hex: bad dtype");
    if (d.form & 2) {
        FG_REQUIRE(d.kappa.kind >= 0 && d.kappa.kind <= 3, FEMGPU_ERR_INVALID_ARG,
                   "pa_create_hex: bad coefficient kind");
        if (d.kappa.kind == FEMGPU_COEF_GRF)
            FG_REQUIRE(d.kappa.n_features > 0 && d.kappa.length > 0, FEMGPU_ERR_INVALID_ARG,
                       "pa_create_hex: GRF needs n_features > 0 and length > 0");
        if (d.kappa.kind == FEMGPU_COEF_QP_ARRAY)
            FG_REQUIRE(d.kappa.qp_values || kqp_dev, FEMGPU_ERR_INVALID_ARG,
                       "pa_create_hex: QP_ARRAY needs qp_values");
    }
    FG_HIP(hipSetDevice(c->device));
    ensure_tables(c->device);
    auto *op = new HexOp();
    try {
        op->ctx = c; op->dtype = d.dtype; op->desc = d; op->desc.kappa.qp_values = nullptr;
        op->geo_req = geo_req;
        HexGrid &g = op->g;
        g.nx = d.nx; g.ny = d.ny; g.nz = d.nz;
        g.E = (long long)d.nx * d.ny * d.nz;
        g.Nx = (long long)d.nx * d.p + 1; g.Ny = (long long)d.ny * d.p + 1; g.Nz = (long long)d.nz * d.p + 1;
        g.ne = ne_of(d.q);
        g.Nxp = g.Nx + (g.Nx & 1);
        div_magic(g.nx, g.mx, g.shx);
        div_magic(g.ny, g.my, g.shy);
        op->n = (uint64_t)g.Nx * g.Ny * g.Nz;
        // the element kernels index DOF rows (gy + Ny gz) and element-vector
        // rows in 32-bit arithmetic
        FG_REQUIRE(op->n < (1ull << 40) && g.Nx * g.Ny < (1ll << 31) && g.Nz < 65536 && g.E < (1ll << 31) &&
                       g.Ny * g.Nz < (1ll << 32) && (long long)d.ny * d.nz * (d.p + 1) * (d.p + 1) < (1ll << 32),
                   FEMGPU_ERR_INVALID_ARG, "pa_create_hex: mesh too large");
        const size_t nv = (size_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
        hipStream_t s = c->stream;
        op->coords.alloc(nv * 3 * 8);
        std::vector<double> h;  // host copy of the vertices unless the caller's array is host memory
        const double *hc = coords_host;
        if (coords_host) {
            FG_HIP(hipMemcpyAsync(op->coords.ptr, coords_host, nv * 24, hipMemcpyHostToDevice, s));
        } else if (coords_dev) {
            FG_HIP(hipMemcpyAsync(op->coords.ptr, coords_dev, nv * 24, hipMemcpyDeviceToDevice, s));
            h.resize(nv * 3);
            FG_HIP(hipMemcpyAsync(h.data(), coords_dev, nv * 24, hipMemcpyDeviceToHost, s));
            hc = h.data();
        } else {
            h.resize(nv * 3);
            hex_vertices(d.nx, d.ny, d.nz, 0.0, 0, h.data());
            FG_HIP(hipMemcpyAsync(op->coords.ptr, h.data(), nv * 24, hipMemcpyHostToDevice, s));
            hc = h.data();
        }
        FG_HIP(hipStreamSynchronize(s));
        const char *qmode = std::getenv("FEMGPU_HEX_QDATA");
        if ((d.form & 2) && !(qmode && std::strcmp(qmode, "general") == 0))
            op->cart = cartesian_factors(d.nx, d.ny, d.nz, hc, op->cf_host);
        if ((d.form & 2) && d.kappa.kind == FEMGPU_COEF_QP_ARRAY) {
            const size_t nq = (size_t)g.E * d.q * d.q * d.q;
            op->kqp.alloc(nq * 8);
            if (d.kappa.qp_values)
                FG_HIP(hipMemcpyAsync(op->kqp.ptr, d.kappa.qp_values, nq * 8, hipMemcpyHostToDevice, s));
            else
                FG_HIP(hipMemcpyAsync(op->kqp.ptr, kqp_dev, nq * 8, hipMemcpyDeviceToDevice, s));
        }
        if ((d.form & 2) && d.kappa.kind == FEMGPU_COEF_GRF) {
            std::vector<double> w, b;
            grf_features_host(d.kappa, w, b);
            op->grf_w.alloc(w.size() * 8);
            op->grf_b.alloc(b.size() * 8);
            FG_HIP(hipMemcpyAsync(op->grf_w.ptr, w.data(), w.size() * 8, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemcpyAsync(op->grf_b.ptr, b.data(), b.size() * 8, hipMemcpyHostToDevice, s));
            FG_HIP(hipStreamSynchronize(s));
        }
        op->build();
    } catch (...) {
        delete op;
        throw;
    }
    return op;
}

// log(kappa)/log(100) at every DOF node (trilinear position of the equispaced
// node), the auxiliary input channel of the learned preconditioner.
__global__ void k_hex_log_kappa_nodes(const HexGrid g, int P, const double *__restrict__ coords,
                                      CoefDev coef, float *__restrict__ out) {
    const size_t N = (size_t)g.Nx * g.Ny * g.Nz;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < N;
         id += (size_t)gridDim.x * blockDim.x) {
        const long long gx = (long long)(id % g.Nx), r = (long long)(id / g.Nx);
        const long long gy = r % g.Ny, gz = r / g.Ny;
        const long long e3[3] = {min(gx / P, (long long)g.nx - 1), min(gy / P, (long long)g.ny - 1),
                                 min(gz / P, (long long)g.nz - 1)};
        const long long gg[3] = {gx, gy, gz};
        double xi[3];
        for (int d = 0; d < 3; ++d) xi[d] = -1.0 + 2.0 * (double)(gg[d] - e3[d] * P) / P;
        const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                  sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
        double xp[3] = {0, 0, 0};
        for (int v = 0; v < 8; ++v) {
            const size_t gv = (size_t)(e3[0] + (sx[v] > 0)) +
                              (size_t)(g.nx + 1) * ((size_t)(e3[1] + (sy[v] > 0)) + (size_t)(g.ny + 1) * (e3[2] + (sz[v] > 0)));
            const double phi = 0.125 * (1 + sx[v] * xi[0]) * (1 + sy[v] * xi[1]) * (1 + sz[v] * xi[2]);
            for (int d = 0; d < 3; ++d) xp[d] += phi * coords[3 * gv + d];
        }
        const double k = coef.kind == FEMGPU_COEF_QP_ARRAY ? 1.0 : coef_eval(coef, xp, 0, 0, 1);
        out[id] = (float)(log(k) / log(100.0));
    }
}

void hex_log_kappa_nodes(femgpu_op *op_, float *out) {
    auto *op = dynamic_cast<HexOp *>(op_);
    FG_REQUIRE(op, FEMGPU_ERR_INVALID_ARG, "log_kappa_nodes: not a hex operator");
    CoefDev c{};
    const femgpu_coef &k = op->desc.kappa;
    c.kind = (op->desc.form & 2) ? k.kind : FEMGPU_COEF_CONST;
    c.value = (op->desc.form & 2) ? k.value : 1.0;
    c.nf = k.n_features; c.lo = k.lo; c.hi = k.hi;
    c.scale = k.kind == FEMGPU_COEF_GRF ? k.sigma * std::sqrt(2.0 / k.n_features) : 0.0;
    c.wf = op->grf_w.as<double>(); c.bf = op->grf_b.as<double>(); c.qp = op->kqp.as<double>();
    hipLaunchKernelGGL(k_hex_log_kappa_nodes, dim3(grid1(op->n)), dim3(256), 0, op->ctx->stream, op->g,
                       op->desc.p, op->coords.as<double>(), c, out);
    FG_HIP(hipGetLastError());
}

bool hex_dirichlet(femgpu_op *op_) {
    auto *op = dynamic_cast<HexOp *>(op_);
    return op && op->desc.dirichlet;
}

bool hex_grid_dims(femgpu_op *op_, int *Nx, int *Ny, int *Nz) {
    auto *op = dynamic_cast<HexOp *>(op_);
    if (!op) return false;
    *Nx = (int)op->g.Nx; *Ny = (int)op->g.Ny; *Nz = (int)op->g.Nz;
    return true;
}

// ---------------------------------------------------------------------------
// Contraction probe: the vector unit against the matrix cores.  One batched
// sweep of the element kernels, Out[line] = [B u | G u] for lines u of N1
// nodal values (Q outputs for each table), on 256 LDS-resident lines per
// workgroup, repeated reps times; each repetition reads the lines from LDS,
// writes the 2Q outputs to LDS and feeds them back (u_k += 2^-10 (Out_k +
// Out_{Q+k})), with the same two barriers for every engine, so the timings
// compare the one-dimensional products and their LDS hand-off alone.
//   ENGINE 0: vector unit, one line per lane, multiply-add chains (k_pa_hex's
//             form without the even-odd tables);
//   ENGINE 1: vector unit, even-odd form (k_pa_hex's geometry kernel);
//   ENGINE 2: matrix cores, 16 lines per v_mfma_*_16x16x4 tile: A = 16 lines x
//             4 nodal values (lane l: line l % 16, value l / 16 + 4 s), B = 4
//             x 16 of [B^T | G^T] (zero-padded to 16 outputs and to whole
//             K-steps of 4), D = 16 x 16 (lane l, result register v: output
//             l % 16 of line MfmaTile<T>::row(l, v); measured on gfx950: rows
//             4 (l / 16) + v in FP32, (l / 16) + 4 v in FP64).
template <class T> struct MfmaTile;
template <> struct MfmaTile<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return lane / 16 + 4 * v; }
};
template <> struct MfmaTile<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane / 16) + v; }
};

template <int N1, int Q, class T, int ENGINE>
__global__ __launch_bounds__(256) void k_sumfac_probe(const T *__restrict__ in, T *__restrict__ out, int reps) {
    constexpr int L = 256, NO = 2 * Q;
    constexpr int SI = N1 | 1, SO = NO + 1;  // odd line strides
    static_assert(NO <= 16, "one MFMA tile holds the outputs");
    __shared__ T sI[L * SI];
    __shared__ T sO[L * SO];
    const int tid = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * L;
    for (int i = tid; i < L * N1; i += L) sI[(i / N1) * SI + i % N1] = in[base * N1 + i];
    __syncthreads();
    constexpr int C = combo(N1, Q);
    const T *Bt = Tab<T>::B(C), *Gt = Tab<T>::G(C);
    constexpr int KS = (N1 + 3) / 4;
    const int lane = tid & 63, w = tid >> 6;
    T breg[KS];
    if constexpr (ENGINE == 2) {
#pragma unroll
        for (int st = 0; st < KS; ++st) {
            const int k = lane / 16 + 4 * st, n = lane % 16;
            breg[st] = (k < N1 && n < Q) ? Bt[n * N1 + k] : (k < N1 && n < NO) ? Gt[(n - Q) * N1 + k] : T(0);
        }
    }
    const T eps = T(1.0 / 1024.0);
    for (int r = 0; r < reps; ++r) {
        if constexpr (ENGINE == 2) {
            using V = typename MfmaTile<T>::v4;
#pragma unroll
            for (int gq = 0; gq < 4; ++gq) {
                const int g = w * 4 + gq;
                V acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int st = 0; st < KS; ++st) {
                    const int k = lane / 16 + 4 * st;
                    const T a = k < N1 ? sI[(16 * g + lane % 16) * SI + k] : T(0);
                    acc = MfmaTile<T>::mma(a, breg[st], acc);
                }
                const int n = lane % 16;
                if (n < NO) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) sO[(16 * g + MfmaTile<T>::row(lane, i)) * SO + n] = acc[i];
                }
            }
        } else {
            T u[N1];
#pragma unroll
            for (int k = 0; k < N1; ++k) u[k] = sI[tid * SI + k];
            T sb[Q], sg[Q];
            if constexpr (ENGINE == 1) {
                const EOs<N1, T> su = eo_split(u);
                eo_contract<Q, N1, 1>(eo_tab<T>(C, 0), su, sb);
                eo_contract<Q, N1, -1>(eo_tab<T>(C, 1), su, sg);
            } else {
#pragma unroll
                for (int a = 0; a < Q; ++a) {
                    T x0 = 0, x1 = 0;
#pragma unroll
                    for (int i = 0; i < N1; ++i) {
                        x0 += Bt[a * N1 + i] * u[i];
                        x1 += Gt[a * N1 + i] * u[i];
                    }
                    sb[a] = x0;
                    sg[a] = x1;
                }
            }
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                sO[tid * SO + a] = sb[a];
                sO[tid * SO + Q + a] = sg[a];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < N1; ++k)
            sI[tid * SI + k] += eps * (sO[tid * SO + k] + sO[tid * SO + Q + k]);
        __syncthreads();
    }
    for (int a = 0; a < NO; ++a) out[(base + tid) * NO + a] = sO[tid * SO + a];
}

}  // namespace femgpu

using namespace femgpu;

extern "C" {

int femgpu_sumfac_probe(femgpu_ctx *c, int p, int q, femgpu_dtype dtype, int engine, int nblocks, int reps,
                        const void *in, void *out, int launches, double *ms) {
    return guard([&] {
        FG_REQUIRE(c && in && out && ms && nblocks > 0 && reps > 0 && launches > 0 && engine >= 0 && engine <= 2,
                   FEMGPU_ERR_INVALID_ARG, "sumfac_probe: bad argument");
        FG_REQUIRE((p == 3 && q == 4) || (p == 4 && q == 5), FEMGPU_ERR_UNSUPPORTED,
                   "sumfac_probe: (p, q) must be (3, 4) or (4, 5)");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        ensure_tables(c->device);
        hipStream_t s = c->stream;
        auto launch = [&]() {
#define SP(N1_, Q_, T_, E_)                                                                                   \
    hipLaunchKernelGGL((k_sumfac_probe<N1_, Q_, T_, E_>), dim3(nblocks), dim3(256), 0, s, (const T_ *)in, \
                       (T_ *)out, reps)
#define SPE(N1_, Q_, T_)                                                                                      \
    do {                                                                                                      \
        if (engine == 0) SP(N1_, Q_, T_, 0);                                                                  \
        else if (engine == 1) SP(N1_, Q_, T_, 1);                                                             \
        else SP(N1_, Q_, T_, 2);                                                                              \
    } while (0)
            if (p == 3) {
                if (dtype == FEMGPU_F64) SPE(4, 4, double);
                else SPE(4, 4, float);
            } else {
                if (dtype == FEMGPU_F64) SPE(5, 5, double);
                else SPE(5, 5, float);
            }
#undef SPE
#undef SP
        };
        launch();  // warm-up
        FG_HIP(hipGetLastError());
        FG_HIP(hipEventRecord(c->t0, s));
        for (int i = 0; i < launches; ++i) launch();
        FG_HIP(hipEventRecord(c->t1, s));
        FG_HIP(hipGetLastError());
        FG_HIP(hipEventSynchronize(c->t1));
        float tot = 0.f;
        FG_HIP(hipEventElapsedTime(&tot, c->t0, c->t1));
        *ms = tot / launches;
        return 0;
    });
}


int femgpu_pa_create_hex(femgpu_ctx *c, const femgpu_hex_desc *d, const double *coords,
                         femgpu_op **out) {
    return guard([&] {
        FG_REQUIRE(c && d && out, FEMGPU_ERR_INVALID_ARG, "pa_create_hex: NULL argument");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        *out = make_hex(c, *d, coords, nullptr, nullptr);
        return 0;
    });
}

int femgpu_pa_hex_qdata_per_point(const femgpu_op *op_, int *values) {
    return guard([&] {
        auto *op = dynamic_cast<const HexOp *>(op_);
        FG_REQUIRE(op && values, FEMGPU_ERR_INVALID_ARG, "pa_hex_qdata_per_point: not a hex operator");
        *values = op->qdata_per_point();
        return 0;
    });
}

int femgpu_pa_hex_qdata_mode(const femgpu_op *op_, int *mode) {
    return guard([&] {
        auto *op = dynamic_cast<const HexOp *>(op_);
        FG_REQUIRE(op && mode, FEMGPU_ERR_INVALID_ARG, "pa_hex_qdata_mode: not a hex operator");
        *mode = op->qdata_mode();
        return 0;
    });
}

int femgpu_pa_hex_element_kernel(const femgpu_op *op_, int *kind) {
    return guard([&] {
        auto *op = dynamic_cast<const HexOp *>(op_);
        FG_REQUIRE(op && kind, FEMGPU_ERR_INVALID_ARG, "pa_hex_element_kernel: not a hex operator");
        *kind = op->quad ? 2 : op->tpe ? 1 : 0;
        return 0;
    });
}

int femgpu_pa_hex_convert(femgpu_op *op_, femgpu_dtype dtype, femgpu_op **out) {
    return guard([&] {
        auto *op = dynamic_cast<HexOp *>(op_);
        FG_REQUIRE(op && out, FEMGPU_ERR_INVALID_ARG, "pa_hex_convert: not a hex operator");
        std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
        femgpu_hex_desc d = op->desc;
        d.dtype = dtype;
        *out = make_hex(op->ctx, d, nullptr, op->coords.ptr, op->kqp.ptr, op->geo ? 1 : 0);
        return 0;
    });
}

int femgpu_pa_hex_rhs_unit(femgpu_op *op_, void *b, femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<HexOp *>(op_);
        FG_REQUIRE(op && b, FEMGPU_ERR_INVALID_ARG, "pa_hex_rhs_unit: not a hex operator");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        const size_t nq = qd_count(op->g.E, op->desc.q);
        const size_t bytes = op->n * op->esize();
        DevBuf tq(nq * op->esize()), ones(bytes), out(mem == FEMGPU_DEVICE ? 0 : bytes);
        void *dst = mem == FEMGPU_DEVICE ? b : out.ptr;
        hipStream_t s = c->stream;
        const int n1 = op->desc.p + 1, q = op->desc.q;
        HexGrid gl = op->g;
        gl.ne = ne_of(q);  // the workgroup kernel's element-vector grouping
        gl.ymerge = 0;
        // b = M_1 1 (mass with weights w|J|), then zero on essential DOFs
        if (op->dtype == FEMGPU_F64) {
            op->setup_t<double>(1, 1.0, true, nullptr, tq.as<double>());
            fill<double>(s, ones.as<double>(), 1.0, op->n);
            launch_pa<double>(n1, q, gl, 1, nullptr, tq.as<double>(), nullptr, ones.as<double>(), op->evec.as<double>(), 0,
                              nullptr, s);
            launch_gather<double>(op->desc.p, gl, op->evec.as<double>(), ones.as<double>(), (double *)dst, 0, 0, nullptr, s);
            hipLaunchKernelGGL(k_hex_bc<double>, dim3(grid1(op->n)), dim3(256), 0, s, op->g,
                               (const double *)dst, (double *)dst, 3, op->desc.dirichlet, nullptr);
        } else {
            op->setup_t<float>(1, 1.0, true, nullptr, tq.as<float>());
            fill<float>(s, ones.as<float>(), 1.f, op->n);
            launch_pa<float>(n1, q, gl, 1, nullptr, tq.as<float>(), nullptr, ones.as<float>(), op->evec.as<float>(), 0,
                             nullptr, s);
            launch_gather<float>(op->desc.p, gl, op->evec.as<float>(), ones.as<float>(), (float *)dst, 0, 0, nullptr, s);
            hipLaunchKernelGGL(k_hex_bc<float>, dim3(grid1(op->n)), dim3(256), 0, s, op->g,
                               (const float *)dst, (float *)dst, 3, op->desc.dirichlet, nullptr);
        }
        FG_HIP(hipGetLastError());
        if (mem != FEMGPU_DEVICE) FG_HIP(hipMemcpyAsync(b, dst, bytes, hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

int femgpu_pa_hex_rhs_random(femgpu_op *op_, uint64_t seed, void *b, femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<HexOp *>(op_);
        FG_REQUIRE(op && b, FEMGPU_ERR_INVALID_ARG, "pa_hex_rhs_random: not a hex operator");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        const size_t bytes = op->n * op->esize();
        DevBuf out(mem == FEMGPU_DEVICE ? 0 : bytes);
        void *dst = mem == FEMGPU_DEVICE ? b : out.ptr;
        hipStream_t s = c->stream;
        if (op->dtype == FEMGPU_F64) {
            hipLaunchKernelGGL(k_random_field<double>, dim3(grid1(op->n)), dim3(256), 0, s, (size_t)op->n, seed, (double *)dst);
            hipLaunchKernelGGL(k_hex_bc<double>, dim3(grid1(op->n)), dim3(256), 0, s, op->g,
                               (const double *)dst, (double *)dst, 3, op->desc.dirichlet, nullptr);
        } else {
            hipLaunchKernelGGL(k_random_field<float>, dim3(grid1(op->n)), dim3(256), 0, s, (size_t)op->n, seed, (float *)dst);
            hipLaunchKernelGGL(k_hex_bc<float>, dim3(grid1(op->n)), dim3(256), 0, s, op->g,
                               (const float *)dst, (float *)dst, 3, op->desc.dirichlet, nullptr);
        }
        FG_HIP(hipGetLastError());
        if (mem != FEMGPU_DEVICE) FG_HIP(hipMemcpyAsync(b, dst, bytes, hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

}  // extern "C"
