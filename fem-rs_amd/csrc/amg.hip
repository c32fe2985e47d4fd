// Smoothed-aggregation algebraic multigrid: host set-up, device V-cycle.
//
// Reference: solve_amg_cg / AmgSolver (amg/src/lib.rs:125-207) build a
// hierarchy with linger's AmgHierarchy (an absent dependency, so its exact
// coarsening is not known) and use one V-cycle as the PCG preconditioner.
// This restates a standard smoothed-aggregation hierarchy (Vanek, Mandel,
// Brezina 1996):
//  - strength: j != i is strong for i when |a_ij| >= theta sqrt(|a_ii a_jj|);
//    rows without strong neighbours (identity rows of row-zeroed Dirichlet
//    DOFs) are not aggregated (zero rows of P) and left to the smoother;
//  - aggregation in three passes over the rows in order (root + its strong
//    neighbours; join an adjacent aggregate of the first pass; new aggregates
//    of what is left);
//  - tentative prolongator: 1 / sqrt(|aggregate|) in each aggregate's column;
//  - P = (I - omega D^-1 A) P_tent, omega = 4 / (3 lambda_max(D^-1 A));
//  - A_c = P^T A P, until at most coarse_max rows remain; the coarsest level
//    is inverted (dense) once.
// A cycle is a V-cycle (one coarse correction per level), a W-cycle (two,
// the second on the residual the first leaves) or an F-cycle (an F-cycle then
// a V-cycle on the coarse level).
// Every product runs row by row with a sparse accumulator and sorted columns,
// so the hierarchy is deterministic.  The cycle runs on the device: per
// level a Chebyshev-Jacobi polynomial smoother (the same polynomial before and
// after the coarse correction, so the cycle is symmetric), residual SpMV,
// restriction by R = P^T, prolongation by P, and a dense solve at the bottom.
#include "csr_row.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace femgpu {

struct HCsr {
    size_t n = 0, m = 0;  // rows, columns
    std::vector<uint64_t> rp;
    std::vector<uint32_t> ci;
    std::vector<double> v;
};

static double diag_of(const HCsr &A, size_t i) {
    for (uint64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
        if (A.ci[k] == i) return A.v[k];
    return 0.0;
}

// C = A B, rows in order, columns sorted (sparse accumulator).
static HCsr spgemm(const HCsr &A, const HCsr &B) {
    HCsr C;
    C.n = A.n; C.m = B.m;
    C.rp.assign(A.n + 1, 0);
    std::vector<double> acc(B.m, 0.0);
    std::vector<char> used(B.m, 0);
    std::vector<uint32_t> cols;
    for (size_t i = 0; i < A.n; ++i) {
        cols.clear();
        for (uint64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const uint32_t a = A.ci[k];
            const double av = A.v[k];
            for (uint64_t t = B.rp[a]; t < B.rp[a + 1]; ++t) {
                const uint32_t j = B.ci[t];
                if (!used[j]) { used[j] = 1; cols.push_back(j); acc[j] = 0.0; }
                acc[j] += av * B.v[t];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (uint32_t j : cols) {
            C.ci.push_back(j);
            C.v.push_back(acc[j]);
            used[j] = 0;
        }
        C.rp[i + 1] = C.ci.size();
    }
    return C;
}

static HCsr transpose(const HCsr &A) {
    HCsr T;
    T.n = A.m; T.m = A.n;
    T.rp.assign(A.m + 1, 0);
    for (uint32_t j : A.ci) ++T.rp[j + 1];
    for (size_t j = 0; j < A.m; ++j) T.rp[j + 1] += T.rp[j];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<uint64_t> pos(T.rp.begin(), T.rp.end() - 1);
    for (size_t i = 0; i < A.n; ++i)  // rows in order: columns of T come out sorted
        for (uint64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const uint64_t p = pos[A.ci[k]]++;
            T.ci[p] = (uint32_t)i;
            T.v[p] = A.v[k];
        }
    return T;
}

// lambda_max of D^-1 A by 20 power steps from the all-ones vector.
static double lmax_dinv_a(const HCsr &A, const std::vector<double> &inv) {
    std::vector<double> v(A.n, 1.0 / std::sqrt((double)std::max<size_t>(A.n, 1))), w(A.n);
    double lam = 0.0;
    for (int it = 0; it < 20; ++it) {
        for (size_t i = 0; i < A.n; ++i) {
            double s = 0.0;
            for (uint64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.v[k] * v[A.ci[k]];
            w[i] = inv[i] * s;
        }
        double nrm = 0.0;
        for (double t : w) nrm += t * t;
        nrm = std::sqrt(nrm);
        if (!(nrm > 0)) break;
        lam = nrm;
        for (size_t i = 0; i < A.n; ++i) v[i] = w[i] / nrm;
    }
    return lam;
}

// Aggregates of the strong-connection graph; -1 = not aggregated.
static std::vector<int64_t> aggregate(const HCsr &A, double theta, size_t &nagg) {
    const size_t n = A.n;
    std::vector<double> d(n);
    for (size_t i = 0; i < n; ++i) d[i] = diag_of(A, i);
    std::vector<uint64_t> srp(n + 1, 0);
    std::vector<uint32_t> sci;
    for (size_t i = 0; i < n; ++i) {
        for (uint64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const uint32_t j = A.ci[k];
            if (j != i && std::fabs(A.v[k]) >= theta * std::sqrt(std::fabs(d[i] * d[j]))) sci.push_back(j);
        }
        srp[i + 1] = sci.size();
    }
    // rows without strong neighbours (identity rows of eliminated DOFs) stay
    // out of every aggregate, also as the neighbour of another row, so that P
    // has zero rows there and P^T A P stays symmetric
    {
        std::vector<char> iso(n);
        for (size_t i = 0; i < n; ++i) iso[i] = srp[i + 1] == srp[i];
        std::vector<uint64_t> rp2(n + 1, 0);
        std::vector<uint32_t> ci2;
        for (size_t i = 0; i < n; ++i) {
            for (uint64_t k = srp[i]; k < srp[i + 1]; ++k)
                if (!iso[sci[k]]) ci2.push_back(sci[k]);
            rp2[i + 1] = ci2.size();
        }
        srp.swap(rp2);
        sci.swap(ci2);
    }
    std::vector<int64_t> agg(n, -1);
    nagg = 0;
    // pass 1: roots whose strong neighbours are all free
    for (size_t i = 0; i < n; ++i) {
        if (agg[i] >= 0 || srp[i + 1] == srp[i]) continue;
        bool free = true;
        for (uint64_t k = srp[i]; k < srp[i + 1] && free; ++k) free = agg[sci[k]] < 0;
        if (!free) continue;
        agg[i] = (int64_t)nagg;
        for (uint64_t k = srp[i]; k < srp[i + 1]; ++k) agg[sci[k]] = (int64_t)nagg;
        ++nagg;
    }
    // pass 2: join the first adjacent aggregate of pass 1
    std::vector<int64_t> a1 = agg;
    for (size_t i = 0; i < n; ++i) {
        if (agg[i] >= 0 || srp[i + 1] == srp[i]) continue;
        for (uint64_t k = srp[i]; k < srp[i + 1]; ++k)
            if (a1[sci[k]] >= 0) { agg[i] = a1[sci[k]]; break; }
    }
    // pass 3: new aggregates from the rest (with their free strong neighbours)
    for (size_t i = 0; i < n; ++i) {
        if (agg[i] >= 0 || srp[i + 1] == srp[i]) continue;
        agg[i] = (int64_t)nagg;
        for (uint64_t k = srp[i]; k < srp[i + 1]; ++k)
            if (agg[sci[k]] < 0) agg[sci[k]] = (int64_t)nagg;
        ++nagg;
    }
    return agg;
}

// ------------------------------------------------------------ kernels --
// y = A x (mode 0), y += A x (mode 1), y = b - A x (mode 2)
// STAGE: the workgroup's entries go through LDS first (levels from 2^19 rows;
// on smaller levels the extra pass and barrier only add latency)
template <bool STAGE>
__global__ __launch_bounds__(256) void k_amg_spmv(size_t n, const uint64_t *__restrict__ rp,
                                                  const uint32_t *__restrict__ ci, const double *__restrict__ v,
                                                  const double *__restrict__ x, const double *__restrict__ b,
                                                  double *__restrict__ y, int mode, const int *skip) {
    if (skip && *skip) return;
    const size_t r0 = blockIdx.x * (size_t)256;
    const size_t i = r0 + threadIdx.x;
    auto xat = [x](uint32_t j) { return x[j]; };
    double s;
    if constexpr (STAGE) {
        __shared__ RowStage<double> st;
        stage_rows<double>(st, rp, ci, v, n, r0);
        if (i >= n) return;
        s = row_staged<double>(st, rp, ci, v, xat, i);
    } else {
        if (i >= n) return;
        s = row_dot_ref<double>(rp, ci, v, xat, i);
    }
    y[i] = mode == 0 ? s : mode == 1 ? y[i] + s : b[i] - s;
}

// Chebyshev-Jacobi from x = 0: first step x = d = inv r / theta; then
// d = c1 d + c2 inv (r - t), x += d with t = A x.
__global__ void k_amg_cheb0(size_t n, const double *__restrict__ r, const double *__restrict__ inv,
                            double *__restrict__ x, double *__restrict__ d, double inv_theta, const int *skip) {
    if (skip && *skip) return;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double z = inv_theta * (inv[i] * r[i]);
    x[i] = z;
    d[i] = z;
}
__global__ void k_amg_chebk(size_t n, const double *__restrict__ r, const double *__restrict__ t,
                            const double *__restrict__ inv, double *__restrict__ x, double *__restrict__ d,
                            double c1, double c2, const int *skip) {
    if (skip && *skip) return;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double di = c1 * d[i] + c2 * (inv[i] * (r[i] - t[i]));
    d[i] = di;
    x[i] += di;
}
__global__ void k_amg_add(size_t n, const double *__restrict__ a, double *__restrict__ x, const int *skip) {
    if (skip && *skip) return;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) x[i] += a[i];
}
// x = M b with M dense row-major [n][n]
__global__ void k_amg_dense(size_t n, const double *__restrict__ M, const double *__restrict__ b,
                            double *__restrict__ x, const int *skip) {
    if (skip && *skip) return;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (size_t j = 0; j < n; ++j) s += M[i * n + j] * b[j];
    x[i] = s;
}

static unsigned g256(size_t n) { return (unsigned)((n + 255) / 256 ? (n + 255) / 256 : 1); }

struct DCsr {
    size_t n = 0, m = 0, nnz = 0;
    bool stage = false;  // k_amg_spmv stages entries in LDS (from 2^19 rows; FEMGPU_CSR_STAGE=1/0 forces)
    DevBuf rp, ci, v;
    void upload(const HCsr &H, hipStream_t s) {
        n = H.n; m = H.m; nnz = H.ci.size();
        stage = n >= (1u << 19);
        if (const char *f = std::getenv("FEMGPU_CSR_STAGE")) stage = std::atoi(f) != 0;
        rp.alloc((n + 1) * 8);
        ci.alloc(std::max<size_t>(nnz, 1) * 4);
        v.alloc(std::max<size_t>(nnz, 1) * 8);
        FG_HIP(hipMemcpyAsync(rp.ptr, H.rp.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
        if (nnz) {
            FG_HIP(hipMemcpyAsync(ci.ptr, H.ci.data(), nnz * 4, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemcpyAsync(v.ptr, H.v.data(), nnz * 8, hipMemcpyHostToDevice, s));
        }
    }
    void spmv(const double *x, const double *b, double *y, int mode, hipStream_t s, const int *skip) const {
        if (n && stage)
            hipLaunchKernelGGL(k_amg_spmv<true>, dim3(g256(n)), dim3(256), 0, s, n, rp.as<uint64_t>(),
                               ci.as<uint32_t>(), v.as<double>(), x, b, y, mode, skip);
        else if (n)
            hipLaunchKernelGGL(k_amg_spmv<false>, dim3(g256(n)), dim3(256), 0, s, n, rp.as<uint64_t>(),
                               ci.as<uint32_t>(), v.as<double>(), x, b, y, mode, skip);
    }
};

struct AmgLevel {
    HCsr A, P;  // host copies (P: this level -> next)
    DCsr dA, dP, dR;
    DevBuf inv, b, x, r, t, d, z;
    DevBuf b2, x2;  // right-hand side and result of a second coarse-level visit (W, F)
    double lmax = 0.0;
};

}  // namespace femgpu

// A cycle captured as a HIP graph for one (input, output, stop flag, cycle kind).
struct AmgGraph {
    const double *r = nullptr;
    double *z = nullptr;
    const int *skip = nullptr;
    int cycle = 0;
    hipGraphExec_t exec = nullptr;
};

struct femgpu_amg {
    femgpu_ctx *ctx = nullptr;
    femgpu_op *op = nullptr;  // borrowed level-0 operator
    femgpu_amg_config cfg{};
    int cycle = FEMGPU_CYCLE_V;
    std::vector<femgpu::AmgLevel> L;
    femgpu::DevBuf coarse_inv;  // dense inverse of the coarsest operator
    std::vector<AmgGraph> graphs;  // most recent last, at most kAmgGraphs
    ~femgpu_amg() {
        for (AmgGraph &g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
    }
};

namespace femgpu {

static void amg_build(femgpu_amg *h, HCsr A0) {
    const femgpu_amg_config &c = h->cfg;
    hipStream_t s = h->ctx->stream;
    h->L.clear();
    h->L.emplace_back();
    h->L[0].A = std::move(A0);
    for (;;) {
        AmgLevel &lv = h->L.back();
        const HCsr &A = lv.A;
        std::vector<double> inv(A.n);
        for (size_t i = 0; i < A.n; ++i) {
            const double dd = diag_of(A, i);
            inv[i] = std::fabs(dd) > 1e-30 ? 1.0 / dd : 1.0;
        }
        lv.lmax = lmax_dinv_a(A, inv);
        const bool last = A.n <= c.coarse_max || (int)h->L.size() >= c.max_levels;
        size_t nagg = 0;
        std::vector<int64_t> agg;
        if (!last) agg = aggregate(A, c.theta, nagg);
        if (last || nagg == 0 || nagg >= A.n) break;
        // tentative prolongator
        std::vector<size_t> cnt(nagg, 0);
        for (int64_t a : agg) if (a >= 0) ++cnt[a];
        HCsr Pt;
        Pt.n = A.n; Pt.m = nagg;
        Pt.rp.assign(A.n + 1, 0);
        for (size_t i = 0; i < A.n; ++i) {
            if (agg[i] >= 0) {
                Pt.ci.push_back((uint32_t)agg[i]);
                Pt.v.push_back(1.0 / std::sqrt((double)cnt[agg[i]]));
            }
            Pt.rp[i + 1] = Pt.ci.size();
        }
        // P = Pt - omega D^-1 A Pt
        const double omega = lv.lmax > 0 ? 4.0 / (3.0 * lv.lmax) : 0.0;
        HCsr AP = spgemm(A, Pt);
        HCsr P;
        P.n = A.n; P.m = nagg;
        P.rp.assign(A.n + 1, 0);
        for (size_t i = 0; i < A.n; ++i) {
            uint64_t a = AP.rp[i], ae = AP.rp[i + 1], b = Pt.rp[i], be = Pt.rp[i + 1];
            while (a < ae || b < be) {  // merge the two sorted rows
                uint32_t j;
                double val = 0.0;
                if (b >= be || (a < ae && AP.ci[a] < Pt.ci[b])) {
                    j = AP.ci[a]; val = -omega * inv[i] * AP.v[a]; ++a;
                } else if (a >= ae || Pt.ci[b] < AP.ci[a]) {
                    j = Pt.ci[b]; val = Pt.v[b]; ++b;
                } else {
                    j = Pt.ci[b]; val = Pt.v[b] - omega * inv[i] * AP.v[a]; ++a; ++b;
                }
                if (val != 0.0) {  // rows of eliminated DOFs come out as explicit zeros
                    P.ci.push_back(j);
                    P.v.push_back(val);
                }
            }
            P.rp[i + 1] = P.ci.size();
        }
        HCsr Ac = spgemm(transpose(P), spgemm(A, P));
        lv.P = std::move(P);
        h->L.emplace_back();
        h->L.back().A = std::move(Ac);
    }
    // device copies
    for (size_t l = 0; l < h->L.size(); ++l) {
        AmgLevel &lv = h->L[l];
        const size_t n = lv.A.n;
        lv.dA.upload(lv.A, s);
        std::vector<double> inv(n);
        for (size_t i = 0; i < n; ++i) {
            const double dd = diag_of(lv.A, i);
            inv[i] = std::fabs(dd) > 1e-30 ? 1.0 / dd : 1.0;
        }
        lv.inv.alloc(std::max<size_t>(n, 1) * 8);
        if (n) FG_HIP(hipMemcpyAsync(lv.inv.ptr, inv.data(), n * 8, hipMemcpyHostToDevice, s));
        for (DevBuf *b : {&lv.b, &lv.x, &lv.r, &lv.t, &lv.d, &lv.z, &lv.b2, &lv.x2})
            b->alloc(std::max<size_t>(n, 1) * 8);
        // R is a named local: the asynchronous copies read it until the sync
        HCsr R;
        if (l + 1 < h->L.size()) {
            R = transpose(lv.P);
            lv.dP.upload(lv.P, s);
            lv.dR.upload(R, s);
        }
        FG_HIP(hipStreamSynchronize(s));  // host vectors above go out of scope
    }
    // dense inverse of the coarsest operator (Gauss-Jordan, partial pivoting)
    const HCsr &Ac = h->L.back().A;
    const size_t nc = Ac.n;
    std::vector<double> M(nc * nc, 0.0), I(nc * nc, 0.0);
    for (size_t i = 0; i < nc; ++i) {
        I[i * nc + i] = 1.0;
        for (uint64_t k = Ac.rp[i]; k < Ac.rp[i + 1]; ++k) M[i * nc + Ac.ci[k]] += Ac.v[k];
    }
    for (size_t col = 0; col < nc; ++col) {
        size_t piv = col;
        for (size_t r = col + 1; r < nc; ++r)
            if (std::fabs(M[r * nc + col]) > std::fabs(M[piv * nc + col])) piv = r;
        FG_REQUIRE(std::fabs(M[piv * nc + col]) > 1e-300, FEMGPU_ERR_BREAKDOWN, "amg: singular coarse operator");
        if (piv != col)
            for (size_t k = 0; k < nc; ++k) {
                std::swap(M[piv * nc + k], M[col * nc + k]);
                std::swap(I[piv * nc + k], I[col * nc + k]);
            }
        const double p = M[col * nc + col];
        for (size_t k = 0; k < nc; ++k) { M[col * nc + k] /= p; I[col * nc + k] /= p; }
        for (size_t r = 0; r < nc; ++r) {
            if (r == col) continue;
            const double f = M[r * nc + col];
            if (f == 0.0) continue;
            for (size_t k = 0; k < nc; ++k) { M[r * nc + k] -= f * M[col * nc + k]; I[r * nc + k] -= f * I[col * nc + k]; }
        }
    }
    h->coarse_inv.alloc(std::max<size_t>(nc * nc, 1) * 8);
    if (nc) FG_HIP(hipMemcpyAsync(h->coarse_inv.ptr, I.data(), nc * nc * 8, hipMemcpyHostToDevice, s));
    FG_HIP(hipStreamSynchronize(s));
}

// x = S(r): Chebyshev-Jacobi of the given degree on [lmax / ratio, lmax] from x = 0.
static void amg_smooth(const femgpu_amg *h, const AmgLevel &lv, const double *r, double *x, hipStream_t s,
                       const int *skip) {
    const size_t n = lv.A.n;
    const double hi = 1.1 * lv.lmax, lo = hi / h->cfg.smoother_ratio;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho_prev = 1.0 / sigma;
    double *d = lv.d.as<double>(), *t = lv.t.as<double>();
    hipLaunchKernelGGL(k_amg_cheb0, dim3(g256(n)), dim3(256), 0, s, n, r, lv.inv.as<double>(), x, d, 1.0 / theta,
                       skip);
    for (int k = 1; k < h->cfg.smoother_degree; ++k) {
        const double rho = 1.0 / (2.0 * sigma - rho_prev);
        lv.dA.spmv(x, nullptr, t, 0, s, skip);
        hipLaunchKernelGGL(k_amg_chebk, dim3(g256(n)), dim3(256), 0, s, n, r, (const double *)t,
                           lv.inv.as<double>(), x, d, rho * rho_prev, 2.0 * rho / delta, skip);
        rho_prev = rho;
    }
}

// x = C(l) b for the cycle kind (V, W, F).  A visit of level l uses that
// level's r, z, t, d as scratch and the next level's b, x (first coarse
// visit) and b2, x2 (second), so b and x must not be this level's scratch.
static void amg_cycle(const femgpu_amg *h, size_t l, int kind, const double *b, double *x, hipStream_t s,
                      const int *skip) {
    const AmgLevel &lv = h->L[l];
    const size_t n = lv.A.n;
    if (l + 1 == h->L.size()) {
        hipLaunchKernelGGL(k_amg_dense, dim3(g256(n)), dim3(256), 0, s, n, h->coarse_inv.as<double>(), b, x, skip);
        return;
    }
    const AmgLevel &nx = h->L[l + 1];
    double *r = lv.r.as<double>(), *z = lv.z.as<double>();
    double *bc = nx.b.as<double>(), *xc = nx.x.as<double>();
    amg_smooth(h, lv, b, x, s, skip);                           // pre-smoothing from zero
    lv.dA.spmv(x, b, r, 2, s, skip);                            // r = b - A x
    lv.dR.spmv(r, nullptr, bc, 0, s, skip);                     // b_c = R r
    amg_cycle(h, l + 1, kind, bc, xc, s, skip);
    if (kind != FEMGPU_CYCLE_V && l + 2 < h->L.size()) {        // second coarse visit (exact on the coarsest)
        const size_t nc = nx.A.n;
        double *b2 = nx.b2.as<double>(), *x2 = nx.x2.as<double>();
        nx.dA.spmv(xc, bc, b2, 2, s, skip);                     // b2 = b_c - A_c x_c
        amg_cycle(h, l + 1, kind == FEMGPU_CYCLE_W ? FEMGPU_CYCLE_W : FEMGPU_CYCLE_V, b2, x2, s, skip);
        hipLaunchKernelGGL(k_amg_add, dim3(g256(nc)), dim3(256), 0, s, nc, (const double *)x2, xc, skip);
    }
    lv.dP.spmv(xc, nullptr, x, 1, s, skip);                     // x += P x_c
    lv.dA.spmv(x, b, r, 2, s, skip);                            // r = b - A x
    amg_smooth(h, lv, r, z, s, skip);                           // post-smoothing
    hipLaunchKernelGGL(k_amg_add, dim3(g256(n)), dim3(256), 0, s, n, (const double *)z, x, skip);
}

static bool amg_graphs_enabled() {
    const char *e = std::getenv("FEMGPU_GRAPHS");
    return !(e && e[0] == '0');
}

constexpr size_t kAmgGraphs = 8;

// One cycle from r into z.  The cycle's kernels (about a dozen per level, all
// small on the coarse levels) are captured once per (r, z, skip, cycle kind)
// into a HIP graph and replayed, so a PCG iteration submits one graph instead
// of every kernel; the graph records kernels only (no memset or copy nodes).
// FEMGPU_GRAPHS=0 launches them directly.
void amg_apply(const femgpu_amg *h, const double *r, double *z, const int *skip) {
    hipStream_t s = h->ctx->stream;
    if (!amg_graphs_enabled()) {
        amg_cycle(h, 0, h->cycle, r, z, s, skip);
        return;
    }
    auto *hm = const_cast<femgpu_amg *>(h);
    for (const AmgGraph &g : hm->graphs)
        if (g.r == r && g.z == z && g.skip == skip && g.cycle == h->cycle) {
            FG_HIP(hipGraphLaunch(g.exec, s));
            return;
        }
    hipGraph_t graph = nullptr;
    FG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    amg_cycle(h, 0, h->cycle, r, z, s, skip);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    FG_HIP(le);
    FG_HIP(ce);
    AmgGraph g;
    g.r = r; g.z = z; g.skip = skip; g.cycle = h->cycle;
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    FG_HIP(ie);
    if (hm->graphs.size() >= kAmgGraphs) {
        // the evicted graph may still run from an earlier launch on s
        FG_HIP(hipStreamSynchronize(s));
        (void)hipGraphExecDestroy(hm->graphs.front().exec);
        hm->graphs.erase(hm->graphs.begin());
    }
    hm->graphs.push_back(g);
    FG_HIP(hipGraphLaunch(g.exec, s));
}

femgpu_op *amg_operator(const femgpu_amg *h) { return h->op; }

}  // namespace femgpu

using namespace femgpu;

extern "C" {

int femgpu_amg_create(femgpu_op *op, const femgpu_amg_config *cfg, femgpu_amg **out) {
    return guard([&] {
        FG_REQUIRE(op && out, FEMGPU_ERR_INVALID_ARG, "amg_create: NULL argument");
        CsrView cv;
        FG_REQUIRE(op->dtype == FEMGPU_F64 && op->csr_arrays(&cv), FEMGPU_ERR_INVALID_ARG,
                   "amg_create: needs an f64 CSR matrix");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        auto *h = new femgpu_amg();
        try {
            h->ctx = c; h->op = op;
            h->cfg.theta = cfg && cfg->theta > 0 ? cfg->theta : 0.08;
            h->cfg.max_levels = cfg && cfg->max_levels > 1 ? cfg->max_levels : 10;
            h->cfg.coarse_max = cfg && cfg->coarse_max > 0 ? cfg->coarse_max : 50;
            h->cfg.smoother_degree = cfg && cfg->smoother_degree > 0 ? cfg->smoother_degree : 2;
            h->cfg.smoother_ratio = cfg && cfg->smoother_ratio > 1 ? cfg->smoother_ratio : 10.0;
            HCsr A;
            A.n = A.m = op->n;
            A.rp.resize(op->n + 1);
            FG_HIP(hipMemcpyAsync(A.rp.data(), cv.rp, (op->n + 1) * 8, hipMemcpyDeviceToHost, c->stream));
            FG_HIP(hipStreamSynchronize(c->stream));
            const size_t nnz = A.rp[op->n];
            A.ci.resize(nnz);
            A.v.resize(nnz);
            if (nnz) {
                FG_HIP(hipMemcpyAsync(A.ci.data(), cv.ci, nnz * 4, hipMemcpyDeviceToHost, c->stream));
                FG_HIP(hipMemcpyAsync(A.v.data(), cv.vals, nnz * 8, hipMemcpyDeviceToHost, c->stream));
            }
            FG_HIP(hipStreamSynchronize(c->stream));
            amg_build(h, std::move(A));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return 0;
    });
}

int femgpu_amg_destroy(femgpu_amg *h) {
    if (!h) return 0;
    (void)hipSetDevice(h->ctx->device);
    (void)hipStreamSynchronize(h->ctx->stream);
    delete h;
    return 0;
}

int femgpu_amg_n_levels(const femgpu_amg *h, int *n) {
    return guard([&] {
        FG_REQUIRE(h && n, FEMGPU_ERR_INVALID_ARG, "amg_n_levels: NULL argument");
        *n = (int)h->L.size();
        return 0;
    });
}

int femgpu_amg_level(const femgpu_amg *h, int level, int which, uint64_t *nrows, uint64_t *ncols,
                     uint64_t *nnz, uint64_t *row_ptr, uint32_t *col_idx, double *values) {
    return guard([&] {
        FG_REQUIRE(h && nrows && ncols && nnz, FEMGPU_ERR_INVALID_ARG, "amg_level: NULL argument");
        FG_REQUIRE(level >= 0 && level < (int)h->L.size(), FEMGPU_ERR_INVALID_ARG, "amg_level: no such level");
        FG_REQUIRE(which == 0 || (which == 1 && level + 1 < (int)h->L.size()), FEMGPU_ERR_INVALID_ARG,
                   "amg_level: which = 0 (A) or 1 (P, not on the coarsest level)");
        const HCsr &M = which == 0 ? h->L[level].A : h->L[level].P;
        *nrows = M.n; *ncols = M.m; *nnz = M.ci.size();
        if (row_ptr) std::memcpy(row_ptr, M.rp.data(), (M.n + 1) * 8);
        if (col_idx && !M.ci.empty()) std::memcpy(col_idx, M.ci.data(), M.ci.size() * 4);
        if (values && !M.v.empty()) std::memcpy(values, M.v.data(), M.v.size() * 8);
        return 0;
    });
}

int femgpu_amg_set_cycle(femgpu_amg *h, int cycle) {
    return guard([&] {
        FG_REQUIRE(h, FEMGPU_ERR_INVALID_ARG, "amg_set_cycle: NULL argument");
        FG_REQUIRE(cycle == FEMGPU_CYCLE_V || cycle == FEMGPU_CYCLE_W || cycle == FEMGPU_CYCLE_F,
                   FEMGPU_ERR_INVALID_ARG, "amg_set_cycle: cycle must be V (0), W (1) or F (2)");
        h->cycle = cycle;
        return 0;
    });
}

int femgpu_amg_vcycle(femgpu_amg *h, const void *b, void *x, femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(h && b && x, FEMGPU_ERR_INVALID_ARG, "amg_vcycle: NULL argument");
        femgpu_ctx *c = h->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        const size_t n = h->L[0].A.n, bytes = n * 8;
        const double *bd = (const double *)b;
        double *xd = (double *)x;
        DevBuf sb, sx;
        if (mem == FEMGPU_HOST) {
            sb.alloc(bytes); sx.alloc(bytes);
            FG_HIP(hipMemcpyAsync(sb.ptr, b, bytes, hipMemcpyHostToDevice, c->stream));
            bd = sb.as<double>(); xd = sx.as<double>();
        }
        amg_apply(h, bd, xd, nullptr);
        FG_HIP(hipGetLastError());
        if (mem == FEMGPU_HOST) FG_HIP(hipMemcpyAsync(x, xd, bytes, hipMemcpyDeviceToHost, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

}  // extern "C"
