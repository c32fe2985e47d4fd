// Partial assembly on simplex meshes (Tri3 / Tet4, P1..P3) and the lumped
// mass operator.
//
// Reference: PAMassOperator / PADiffusionOperator::apply
// (assembly/src/partial.rs:85-130, 154-208) and LumpedMassOperator
// (partial.rs:217-274).  The reference re-evaluates bases, Jacobians and the
// coefficient for every element on every apply; here the affine geometry is
// folded at set-up into per-element quadrature data
//   diffusion: D_e = coef |det J| J^-1 J^-T  (dim x dim, symmetric)
//   mass:      m_e = coef |det J|
// and the apply is y_e[i] = sum_q w_q G_q[i] . (D_e G_q^T x_e)  (diffusion) or
// y_e[i] = sum_q w_q m_e phi_q[i] (phi_q . x_e)  (mass), with the reference
// basis tables (phi_q, G_q) in LDS.  One thread per element; scatter with
// hardware atomics.
#include "basis.hpp"
#include "internal.hpp"
#include "simplex.hpp"

#include <algorithm>
#include <cmath>

namespace femgpu {

constexpr int kMaxNl = 20, kMaxQp = 36;

// NL (local DOFs: 3 / 6 / 10 on triangles, 4 / 10 / 20 on tetrahedra) is a
// template parameter so that the element's x and y values stay in registers.
//
// Summation into y: when the workgroup's DOF indices span fewer than kWin
// entries (neighbouring elements of a well-ordered mesh, e.g. a row of a
// structured triangle mesh), its elements first add into an LDS window and the
// workgroup then adds each nonzero window entry to y once, so shared vertices
// cost one global atomic per workgroup instead of one per element.  Otherwise
// every element adds its values to y directly.  Either way the order of the
// additions into y is not fixed (FP64 results agree to rounding).
constexpr int kWin = 4096;

template <class T, int DIM, int NL>
__global__ __launch_bounds__(256) void k_pa_simplex(size_t ne, int nq, int form,
                                                    const uint32_t *__restrict__ edofs,
                                                    const T *__restrict__ qdata,
                                                    const T *__restrict__ cq,
                                                    const double *__restrict__ tab,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    const int *skip) {
    constexpr int nl = NL;
    // tab: [nq] weights, [nq][nl] phi, [nq][nl][DIM] grad (reference element)
    __shared__ T s_tab[kMaxQp * (1 + NL * (1 + DIM))];
    __shared__ T win[kWin];
    __shared__ uint32_t s_lo, s_hi;
    if (skip && *skip) return;
    const int tab_len = nq * (1 + nl * (1 + DIM));
    for (int i = threadIdx.x; i < tab_len; i += blockDim.x) s_tab[i] = (T)tab[i];
    if (threadIdx.x == 0) {
        s_lo = 0xffffffffu;
        s_hi = 0;
    }
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const bool live = e < ne;
    uint32_t dof[NL];
    uint32_t lo = 0xffffffffu, hi = 0;
#pragma unroll
    for (int l = 0; l < nl; ++l) {
        dof[l] = live ? edofs[e * nl + l] : 0u;
        if (live) {
            lo = min(lo, dof[l]);
            hi = max(hi, dof[l]);
        }
    }
    __syncthreads();
    if (live) {
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
    }
    __syncthreads();
    const uint32_t wlo = s_lo;
    const uint32_t span = s_hi - wlo + 1;  // >= 1: every workgroup holds a live element
    const bool windowed = span <= (uint32_t)kWin;  // uniform over the workgroup
    if (windowed)
        for (uint32_t i = threadIdx.x; i < span; i += 256) win[i] = T(0);

    T ye[NL];
#pragma unroll
    for (int l = 0; l < nl; ++l) ye[l] = 0;
    if (live) {
        const T *W = s_tab, *PHI = s_tab + nq, *GR = s_tab + nq + nq * nl;
        T xe[NL];
#pragma unroll
        for (int l = 0; l < nl; ++l) xe[l] = x[dof[l]];
        constexpr int NS = DIM == 2 ? 3 : 6;
        if (form == FEMGPU_DIFFUSION) {
            T D[NS];
            for (int c = 0; c < NS; ++c) D[c] = qdata[e * NS + c];
            for (int q = 0; q < nq; ++q) {
                const T *G = GR + (size_t)q * nl * DIM;
                T gu[3] = {0, 0, 0};
#pragma unroll
                for (int l = 0; l < nl; ++l)
                    for (int d = 0; d < DIM; ++d) gu[d] += xe[l] * G[l * DIM + d];
                T v[3];
                if (DIM == 2) {
                    v[0] = D[0] * gu[0] + D[1] * gu[1];
                    v[1] = D[1] * gu[0] + D[2] * gu[1];
                } else {
                    v[0] = D[0] * gu[0] + D[1] * gu[1] + D[2] * gu[2];
                    v[1] = D[1] * gu[0] + D[3] * gu[1] + D[4] * gu[2];
                    v[2] = D[2] * gu[0] + D[4] * gu[1] + D[5] * gu[2];
                }
                const T wq = cq ? W[q] * cq[e * nq + q] : W[q];
#pragma unroll
                for (int i = 0; i < nl; ++i) {
                    T s = 0;
                    for (int d = 0; d < DIM; ++d) s += G[i * DIM + d] * v[d];
                    ye[i] += wq * s;
                }
            }
        } else {
            const T m = qdata[e];
            for (int q = 0; q < nq; ++q) {
                const T *ph = PHI + (size_t)q * nl;
                T u = 0;
#pragma unroll
                for (int l = 0; l < nl; ++l) u += ph[l] * xe[l];
                const T wu = (cq ? W[q] * cq[e * nq + q] : W[q]) * m * u;
#pragma unroll
                for (int i = 0; i < nl; ++i) ye[i] += wu * ph[i];
            }
        }
    }
    if (!windowed) {
        if (live)
#pragma unroll
            for (int l = 0; l < nl; ++l) unsafeAtomicAdd(&y[dof[l]], ye[l]);
        return;
    }
    __syncthreads();  // window cleared
    if (live)
#pragma unroll
        for (int l = 0; l < nl; ++l)
            __hip_atomic_fetch_add(&win[dof[l] - wlo], ye[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < span; i += 256) {
        const T v = win[i];
        if (v != T(0) || v != v) unsafeAtomicAdd(&y[wlo + i], v);  // untouched entries stay 0
    }
}

template <class T>
__global__ void k_diag_scale(size_t n, const T *__restrict__ d, const T *__restrict__ x,
                             T *__restrict__ y, int mode, const int *skip) {
    if (skip && *skip) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        if (mode == 0) y[i] = d[i] * x[i];
        else if (mode == 1) y[i] += d[i] * x[i];
        else y[i] = fabs((double)d[i]) > 1e-14 ? x[i] / d[i] : T(0);
    }
}

static unsigned blocks_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > (1u << 20)) b = 1u << 20;
    return (unsigned)(b ? b : 1);
}

struct SimplexOp final : femgpu_op {
    int dim = 2, form = FEMGPU_DIFFUSION, nl = 0, nq = 0;
    size_t ne = 0;
    DevBuf edofs, qdata, tab, diag;
    DevBuf cq;  // coefficient per (element, quadrature point), or empty: folded into qdata

    const char *name() const override { return "pa_simplex"; }

    template <class T> void apply_t(const T *x, T *y, bool acc, const int *skip) {
        hipStream_t s = ctx->stream;
        if (!acc) fill<T>(s, y, T(0), n, skip);
        if (!ne) return;
        const dim3 grid((unsigned)((ne + 255) / 256)), block(256);
#define PAS(D_, NL_)                                                                                        \
    hipLaunchKernelGGL((k_pa_simplex<T, D_, NL_>), grid, block, 0, s, ne, nq, form, edofs.as<uint32_t>(),     \
                       qdata.as<T>(), (const T *)cq.ptr, tab.as<double>(), x, y, skip)
        if (dim == 2 && nl == 3) PAS(2, 3);
        else if (dim == 2 && nl == 6) PAS(2, 6);
        else if (dim == 2 && nl == 10) PAS(2, 10);
        else if (dim == 3 && nl == 4) PAS(3, 4);
        else if (dim == 3 && nl == 10) PAS(3, 10);
        else if (dim == 3 && nl == 20) PAS(3, 20);
        else FG_REQUIRE(false, FEMGPU_ERR_UNSUPPORTED, "pa_simplex: unsupported element (dim, local DOFs)");
#undef PAS
    }
    void apply_dev(const void *x, void *y, bool acc, const int *skip) override {
        if (dtype == FEMGPU_F64) apply_t<double>((const double *)x, (double *)y, acc, skip);
        else apply_t<float>((const float *)x, (float *)y, acc, skip);
    }
    void diagonal_dev(void *d) override {
        FG_HIP(hipMemcpyAsync(d, diag.ptr, n * esize(), hipMemcpyDeviceToDevice, ctx->stream));
    }
};

struct LumpedOp final : femgpu_op {
    DevBuf diag;
    const char *name() const override { return "lumped_mass"; }
    template <class T> void run(const T *x, T *y, int mode, const int *skip) {
        hipLaunchKernelGGL(k_diag_scale<T>, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, (size_t)n,
                           diag.as<T>(), x, y, mode, skip);
    }
    void apply_dev(const void *x, void *y, bool acc, const int *skip) override {
        if (dtype == FEMGPU_F64) run<double>((const double *)x, (double *)y, acc ? 1 : 0, skip);
        else run<float>((const float *)x, (float *)y, acc ? 1 : 0, skip);
    }
    void diagonal_dev(void *d) override {
        FG_HIP(hipMemcpyAsync(d, diag.ptr, n * esize(), hipMemcpyDeviceToDevice, ctx->stream));
    }
};

void check_simplex_mesh(int dim, int order, uint64_t n_nodes, const double *coords, uint64_t n_elem,
                        const uint32_t *nodes, const uint32_t *edofs, uint64_t n_dofs, int nl) {
    FG_REQUIRE(coords && nodes && edofs, FEMGPU_ERR_INVALID_ARG, "simplex: NULL mesh array");
    (void)order; (void)dim;
    for (uint64_t i = 0; i < n_elem * (uint64_t)(dim + 1); ++i)
        FG_REQUIRE(nodes[i] < n_nodes, FEMGPU_ERR_INVALID_ARG,
                   "simplex: elem_nodes[" + std::to_string(i) + "] out of range");
    for (uint64_t i = 0; i < n_elem * (uint64_t)nl; ++i)
        FG_REQUIRE(edofs[i] < n_dofs, FEMGPU_ERR_INVALID_ARG,
                   "simplex: elem_dofs[" + std::to_string(i) + "] out of range");
}

template <class T> static void upload(femgpu_ctx *c, DevBuf &b, const std::vector<double> &h) {
    std::vector<T> t(h.begin(), h.end());
    b.alloc(t.size() * sizeof(T));
    if (!t.empty()) FG_HIP(hipMemcpyAsync(b.ptr, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace femgpu

using namespace femgpu;

static int create_simplex(femgpu_ctx *c, int dim, int order, uint64_t n_nodes, const double *coords,
                          uint64_t n_elem, const uint32_t *nodes, const uint32_t *edofs, uint64_t n_dofs,
                          int form, double coef, const double *coef_qp, int quad_order, femgpu_op **out) {
    return guard([&] {
        FG_REQUIRE(c && out, FEMGPU_ERR_INVALID_ARG, "pa_create_simplex: NULL argument");
        FG_REQUIRE(form == FEMGPU_MASS || form == FEMGPU_DIFFUSION, FEMGPU_ERR_INVALID_ARG,
                   "pa_create_simplex: form must be MASS or DIFFUSION");
        SimplexSetup S(dim, order, quad_order);
        FG_REQUIRE(S.nq <= kMaxQp, FEMGPU_ERR_UNSUPPORTED, "pa_create_simplex: too many points");
        check_simplex_mesh(dim, order, n_nodes, coords, n_elem, nodes, edofs, n_dofs, S.nl);
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        auto *op = new SimplexOp();
        try {
            op->ctx = c; op->dtype = FEMGPU_F64; op->n = n_dofs; op->dim = dim; op->form = form;
            op->nl = S.nl; op->nq = S.nq; op->ne = n_elem;
            const int ns = dim == 2 ? 3 : 6;
            std::vector<double> qd((size_t)n_elem * (form == FEMGPU_DIFFUSION ? ns : 1));
            std::vector<double> dg(n_dofs, 0.0);
            for (uint64_t e = 0; e < n_elem; ++e) {
                double J[3][3], I[3][3];
                const double det = S.jac(coords, nodes + e * (dim + 1), J, I);
                // with per-point values the coefficient multiplies each point's weight instead
                const double s = (coef_qp ? 1.0 : coef) * std::fabs(det);
                double D[6] = {0, 0, 0, 0, 0, 0};
                if (form == FEMGPU_DIFFUSION) {
                    int k = 0;
                    for (int a = 0; a < dim; ++a)
                        for (int b = a; b < dim; ++b) {
                            double v = 0.0;
                            for (int m = 0; m < dim; ++m) v += I[a][m] * I[b][m];
                            D[k] = s * v;
                            qd[e * ns + k] = D[k];
                            ++k;
                        }
                } else {
                    qd[e] = s;
                }
                // assembled diagonal for Jacobi
                for (int q = 0; q < S.nq; ++q)
                    for (int i = 0; i < S.nl; ++i) {
                        double v;
                        if (form == FEMGPU_DIFFUSION) {
                            const double *g = &S.grad[((size_t)q * S.nl + i) * dim];
                            if (dim == 2) v = D[0] * g[0] * g[0] + 2 * D[1] * g[0] * g[1] + D[2] * g[1] * g[1];
                            else v = D[0] * g[0] * g[0] + D[3] * g[1] * g[1] + D[5] * g[2] * g[2] +
                                     2 * (D[1] * g[0] * g[1] + D[2] * g[0] * g[2] + D[4] * g[1] * g[2]);
                        } else {
                            const double ph = S.phi[(size_t)q * S.nl + i];
                            v = s * ph * ph;
                        }
                        const double wq = coef_qp ? S.w[q] * coef_qp[e * S.nq + q] : S.w[q];
                        dg[edofs[e * S.nl + i]] += wq * v;
                    }
            }
            upload<double>(c, op->qdata, qd);
            if (coef_qp)
                upload<double>(c, op->cq, std::vector<double>(coef_qp, coef_qp + n_elem * S.nq));
            upload<double>(c, op->tab, S.tab);
            upload<double>(c, op->diag, dg);
            op->edofs.alloc(n_elem * S.nl * 4);
            if (n_elem)
                FG_HIP(hipMemcpyAsync(op->edofs.ptr, edofs, n_elem * S.nl * 4, hipMemcpyHostToDevice, c->stream));
            FG_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
        return 0;
    });
}

static int create_lumped(femgpu_ctx *c, int dim, int order, uint64_t n_nodes, const double *coords,
                         uint64_t n_elem, const uint32_t *nodes, const uint32_t *edofs, uint64_t n_dofs,
                         double rho, const double *rho_qp, int quad_order, femgpu_op **out) {
    return guard([&] {
        FG_REQUIRE(c && out, FEMGPU_ERR_INVALID_ARG, "lumped_mass_create: NULL argument");
        SimplexSetup S(dim, order, quad_order);
        check_simplex_mesh(dim, order, n_nodes, coords, n_elem, nodes, edofs, n_dofs, S.nl);
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        // LumpedMassOperator::assemble (partial.rs:228-262): M_ii = sum_q w|J| rho phi_i sum_j phi_j
        std::vector<double> dg(n_dofs, 0.0);
        for (uint64_t e = 0; e < n_elem; ++e) {
            double J[3][3], I[3][3];
            const double det = S.jac(coords, nodes + e * (dim + 1), J, I);
            for (int q = 0; q < S.nq; ++q) {
                const double w = S.w[q] * std::fabs(det);
                const double *ph = &S.phi[(size_t)q * S.nl];
                double sum = 0.0;
                for (int i = 0; i < S.nl; ++i) sum += ph[i];
                const double r = rho_qp ? rho_qp[e * S.nq + q] : rho;
                for (int i = 0; i < S.nl; ++i) dg[edofs[e * S.nl + i]] += w * r * ph[i] * sum;
            }
        }
        auto *op = new LumpedOp();
        try {
            op->ctx = c; op->dtype = FEMGPU_F64; op->n = n_dofs;
            upload<double>(c, op->diag, dg);
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
        return 0;
    });
}

extern "C" {

int femgpu_pa_create_simplex(femgpu_ctx *c, int dim, int order, uint64_t n_nodes,
                             const double *coords, uint64_t n_elem, const uint32_t *nodes,
                             const uint32_t *edofs, uint64_t n_dofs, int form, double coef,
                             int quad_order, femgpu_op **out) {
    return create_simplex(c, dim, order, n_nodes, coords, n_elem, nodes, edofs, n_dofs, form, coef, nullptr,
                          quad_order, out);
}

int femgpu_pa_create_simplex_qp(femgpu_ctx *c, int dim, int order, uint64_t n_nodes,
                                const double *coords, uint64_t n_elem, const uint32_t *nodes,
                                const uint32_t *edofs, uint64_t n_dofs, int form, const double *coef_qp,
                                int quad_order, femgpu_op **out) {
    if (!coef_qp) return fail(FEMGPU_ERR_INVALID_ARG, "pa_create_simplex_qp: coef_qp is NULL");
    return create_simplex(c, dim, order, n_nodes, coords, n_elem, nodes, edofs, n_dofs, form, 1.0, coef_qp,
                          quad_order, out);
}

int femgpu_lumped_mass_create(femgpu_ctx *c, int dim, int order, uint64_t n_nodes,
                              const double *coords, uint64_t n_elem, const uint32_t *nodes,
                              const uint32_t *edofs, uint64_t n_dofs, double rho, int quad_order,
                              femgpu_op **out) {
    return create_lumped(c, dim, order, n_nodes, coords, n_elem, nodes, edofs, n_dofs, rho, nullptr,
                         quad_order, out);
}

int femgpu_lumped_mass_create_qp(femgpu_ctx *c, int dim, int order, uint64_t n_nodes,
                                 const double *coords, uint64_t n_elem, const uint32_t *nodes,
                                 const uint32_t *edofs, uint64_t n_dofs, const double *rho_qp, int quad_order,
                                 femgpu_op **out) {
    if (!rho_qp) return fail(FEMGPU_ERR_INVALID_ARG, "lumped_mass_create_qp: rho_qp is NULL");
    return create_lumped(c, dim, order, n_nodes, coords, n_elem, nodes, edofs, n_dofs, 1.0, rho_qp,
                         quad_order, out);
}

int femgpu_simplex_qp_coords(int dim, uint64_t n_nodes, const double *coords, uint64_t n_elem,
                             const uint32_t *nodes, int quad_order, double *xq, uint64_t *nq) {
    return guard([&] {
        FG_REQUIRE(coords && nodes && nq, FEMGPU_ERR_INVALID_ARG, "simplex_qp_coords: NULL argument");
        SimplexSetup S(dim, 1, quad_order);
        *nq = (uint64_t)S.nq;
        if (!xq) return 0;
        for (uint64_t i = 0; i < n_elem * (uint64_t)(dim + 1); ++i)
            FG_REQUIRE(nodes[i] < n_nodes, FEMGPU_ERR_INVALID_ARG, "simplex_qp_coords: node index out of range");
        for (uint64_t e = 0; e < n_elem; ++e) {
            const uint32_t *v = nodes + e * (dim + 1);
            for (int q = 0; q < S.nq; ++q) {
                const double *xi = &S.pts[(size_t)q * dim];
                for (int r = 0; r < dim; ++r) {
                    // x = x_0 + J xi, J columns x_k - x_0 (mesh/src/transformation.rs:34-64)
                    double acc = 0.0;
                    for (int k = 0; k < dim; ++k) acc += (coords[dim * v[k + 1] + r] - coords[dim * v[0] + r]) * xi[k];
                    xq[(e * S.nq + q) * dim + r] = coords[dim * v[0] + r] + acc;
                }
            }
        }
        return 0;
    });
}

int femgpu_lumped_mass_apply_inverse(femgpu_op *op_, const void *x, void *y, femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<LumpedOp *>(op_);
        FG_REQUIRE(op && x && y, FEMGPU_ERR_INVALID_ARG, "lumped_mass_apply_inverse: not a lumped mass");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        const size_t bytes = op->n * 8;
        const void *xd = x;
        void *yd = y;
        if (mem == FEMGPU_HOST) {
            op->stage_x.ensure(bytes); op->stage_y.ensure(bytes);
            FG_HIP(hipMemcpyAsync(op->stage_x.ptr, x, bytes, hipMemcpyHostToDevice, c->stream));
            xd = op->stage_x.ptr; yd = op->stage_y.ptr;
        }
        op->run<double>((const double *)xd, (double *)yd, 2, nullptr);
        FG_HIP(hipGetLastError());
        if (mem == FEMGPU_HOST) FG_HIP(hipMemcpyAsync(y, yd, bytes, hipMemcpyDeviceToHost, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

}  // extern "C"

// ------------------------------------------- essential-DOF operator wrapper --
// The matrix-free essential-BC pattern of examples/src/maxwell.rs:142-197:
// y = A x0 with x0 = x with the constrained entries zeroed, then y_d = x_d on
// each constrained DOF d (identity rows, eliminated columns: the operator
// stays symmetric).  The diagonal is A's with 1 on the constrained DOFs.
namespace femgpu {

template <class T>
__global__ void k_zero_rows(size_t m, const uint32_t *__restrict__ dofs, T *__restrict__ v, const int *skip) {
    if (skip && *skip) return;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < m) v[dofs[i]] = T(0);
}

template <class T>
__global__ void k_set_rows(size_t m, const uint32_t *__restrict__ dofs, const T *__restrict__ x, T *__restrict__ y,
                           int one, const int *skip) {
    if (skip && *skip) return;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < m) y[dofs[i]] = one ? T(1) : x[dofs[i]];
}

template <class T>
__global__ void k_axpy1(size_t n, const T *__restrict__ t, T *__restrict__ y, const int *skip) {
    if (skip && *skip) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] += t[i];
}

struct EssentialOp final : femgpu_op {
    femgpu_op *inner = nullptr;  // borrowed
    size_t m = 0;
    DevBuf dofs, xw, tw;

    const char *name() const override { return "essential"; }

    template <class T> void apply_t(const T *x, T *y, bool acc, const int *skip) {
        hipStream_t s = ctx->stream;
        T *w = xw.as<T>();
        FG_HIP(hipMemcpyAsync(w, x, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (m)
            hipLaunchKernelGGL(k_zero_rows<T>, dim3(blocks_for(m)), dim3(256), 0, s, m, dofs.as<uint32_t>(), w, skip);
        T *out = acc ? tw.as<T>() : y;
        inner->apply_dev(w, out, false, skip);
        if (m)
            hipLaunchKernelGGL(k_set_rows<T>, dim3(blocks_for(m)), dim3(256), 0, s, m, dofs.as<uint32_t>(), x, out,
                               0, skip);
        if (acc) hipLaunchKernelGGL(k_axpy1<T>, dim3(blocks_for(n)), dim3(256), 0, s, (size_t)n, (const T *)out, y, skip);
    }
    void apply_dev(const void *x, void *y, bool acc, const int *skip) override {
        if (dtype == FEMGPU_F64) apply_t<double>((const double *)x, (double *)y, acc, skip);
        else apply_t<float>((const float *)x, (float *)y, acc, skip);
    }
    void diagonal_dev(void *d) override {
        inner->diagonal_dev(d);
        if (!m) return;
        if (dtype == FEMGPU_F64)
            hipLaunchKernelGGL(k_set_rows<double>, dim3(blocks_for(m)), dim3(256), 0, ctx->stream, m,
                               dofs.as<uint32_t>(), (const double *)nullptr, (double *)d, 1, nullptr);
        else
            hipLaunchKernelGGL(k_set_rows<float>, dim3(blocks_for(m)), dim3(256), 0, ctx->stream, m,
                               dofs.as<uint32_t>(), (const float *)nullptr, (float *)d, 1, nullptr);
    }
};

}  // namespace femgpu

extern "C" {

int femgpu_essential_create(femgpu_op *inner, uint64_t n_constrained, const uint32_t *dofs, femgpu_op **out) {
    return guard([&] {
        FG_REQUIRE(inner && out && (dofs || n_constrained == 0), FEMGPU_ERR_INVALID_ARG,
                   "essential_create: NULL argument");
        std::vector<uint32_t> d(dofs, dofs + n_constrained);
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());
        for (uint32_t v : d)
            FG_REQUIRE(v < inner->n, FEMGPU_ERR_INVALID_ARG,
                       "essential_create: constrained DOF " + std::to_string(v) + " out of range");
        femgpu_ctx *c = inner->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        auto *op = new EssentialOp();
        try {
            op->ctx = c; op->dtype = inner->dtype; op->n = inner->n; op->inner = inner; op->m = d.size();
            op->dofs.alloc(d.size() * 4);
            op->xw.alloc(op->n * op->esize());
            op->tw.alloc(op->n * op->esize());
            if (!d.empty())
                FG_HIP(hipMemcpyAsync(op->dofs.ptr, d.data(), d.size() * 4, hipMemcpyHostToDevice, c->stream));
            FG_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
        return 0;
    });
}

int femgpu_essential_mask_rhs(femgpu_op *op_, void *b, femgpu_mem mem) {
    return guard([&] {
        auto *op = dynamic_cast<EssentialOp *>(op_);
        FG_REQUIRE(op && b, FEMGPU_ERR_INVALID_ARG, "essential_mask_rhs: not an essential-DOF operator");
        femgpu_ctx *c = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        const size_t bytes = op->n * op->esize();
        void *bd = b;
        if (mem == FEMGPU_HOST) {
            bd = op->tw.ptr;
            FG_HIP(hipMemcpyAsync(bd, b, bytes, hipMemcpyHostToDevice, c->stream));
        }
        if (op->m) {
            if (op->dtype == FEMGPU_F64)
                hipLaunchKernelGGL(k_zero_rows<double>, dim3(blocks_for(op->m)), dim3(256), 0, c->stream, op->m,
                                   op->dofs.as<uint32_t>(), (double *)bd, nullptr);
            else
                hipLaunchKernelGGL(k_zero_rows<float>, dim3(blocks_for(op->m)), dim3(256), 0, c->stream, op->m,
                                   op->dofs.as<uint32_t>(), (float *)bd, nullptr);
        }
        if (mem == FEMGPU_HOST) FG_HIP(hipMemcpyAsync(b, bd, bytes, hipMemcpyDeviceToHost, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

}  // extern "C"
