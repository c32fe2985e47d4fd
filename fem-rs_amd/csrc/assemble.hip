// Device assembly of simplex bilinear and linear forms into a CsrMatrix and a
// load vector (the step before the solve path in the config-1 workflow).
//
// Reference: Assembler::assemble_bilinear / assemble_linear
// (assembly/src/assembler.rs:147-333) with DiffusionIntegrator / MassIntegrator
// (assembly/src/standard/diffusion.rs:34-55, standard/mass.rs:28-40) and
// DomainSourceIntegrator (standard/source.rs:44-50), then
// CooMatrix::into_csr (linalg/src/coo.rs:76-): element matrices become COO
// triplets in element order, which are sorted by (row, col) with duplicates
// summed.  The reference sorts unstably; here the sort is a stable radix sort,
// so duplicates are summed in element order, one key run per thread, and the
// result is deterministic.
//
//   1. k_simplex_elem_mat: one thread per element; J and J^-T from the affine
//      map (mesh/src/transformation.rs:34-64), physical gradients J^-T grad_ref,
//      K_e[i][j] = sum_q w_q |det J| coef (grad phi_i . grad phi_j) (or
//      phi_i phi_j), summed over q in rule order; writes key (row << 32 | col)
//      and the value of each triplet.
//   2. hipcub radix sort of (key, triplet index), stable.
//   3. run heads, exclusive scan -> position of each unique key; one thread per
//      run sums its values in order, writes col_idx / values and counts its row.
//   4. inclusive scan of the row counts -> row_ptr.
// The load vector goes the same way with 32-bit keys (the DOF) and no columns.
#include "internal.hpp"
#include "simplex.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

namespace femgpu {

struct GeomD {
    double J[3][3], JIT[3][3], x0[3], det;
};

// The affine map of element e: J columns x_k - x_0, det, J^-T.
template <int DIM>
__device__ __forceinline__ void simplex_geom_dev(const double *__restrict__ coords,
                                                 const uint32_t *__restrict__ nodes, GeomD &g) {
#pragma clang fp contract(off)
    for (int r = 0; r < DIM; ++r) g.x0[r] = coords[DIM * nodes[0] + r];
    for (int c = 0; c < DIM; ++c)
        for (int r = 0; r < DIM; ++r) g.J[r][c] = coords[DIM * nodes[c + 1] + r] - g.x0[r];
    const double(&J)[3][3] = g.J;
    if (DIM == 2) {
        g.det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        g.JIT[0][0] = J[1][1] / g.det; g.JIT[0][1] = -J[1][0] / g.det;
        g.JIT[1][0] = -J[0][1] / g.det; g.JIT[1][1] = J[0][0] / g.det;
    } else {
        g.det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        const double id = 1.0 / g.det;
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
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) g.JIT[r][c] = I[c][r];
    }
}

// tab: [nq] weights, [nq][nl] phi, [nq][nl][DIM] reference gradients
template <int DIM>
__global__ __launch_bounds__(256) void k_simplex_elem_mat(size_t ne, int nl, int nq, int form, double coef,
                                                          const double *__restrict__ cq,
                                                          const double *__restrict__ coords,
                                                          const uint32_t *__restrict__ nodes,
                                                          const uint32_t *__restrict__ edofs,
                                                          const double *__restrict__ tab,
                                                          uint64_t *__restrict__ keys,
                                                          uint32_t *__restrict__ idx,
                                                          double *__restrict__ vals) {
#pragma clang fp contract(off)
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= ne) return;
    GeomD g;
    simplex_geom_dev<DIM>(coords, nodes + e * (DIM + 1), g);
    const double *W = tab, *PHI = tab + nq, *GR = tab + nq + (size_t)nq * nl;
    const double adet = fabs(g.det);
    const size_t base = e * (size_t)nl * nl;
    for (int i = 0; i < nl; ++i) {
        const uint64_t ri = (uint64_t)edofs[e * nl + i] << 32;
        for (int j = 0; j < nl; ++j) {
            double s = 0.0;
            for (int q = 0; q < nq; ++q) {
                const double wc = (W[q] * adet) * (cq ? cq[e * nq + q] : coef);
                if (form == FEMGPU_DIFFUSION) {
                    double gi[3], gj[3];
                    for (int r = 0; r < DIM; ++r) {
                        double a = 0.0, b = 0.0;
                        for (int k = 0; k < DIM; ++k) {
                            a += g.JIT[r][k] * GR[((size_t)q * nl + i) * DIM + k];
                            b += g.JIT[r][k] * GR[((size_t)q * nl + j) * DIM + k];
                        }
                        gi[r] = a;
                        gj[r] = b;
                    }
                    double dot = 0.0;
                    for (int k = 0; k < DIM; ++k) dot += gi[k] * gj[k];
                    s += wc * dot;
                } else {
                    s += wc * PHI[(size_t)q * nl + i] * PHI[(size_t)q * nl + j];
                }
            }
            const size_t t = base + (size_t)i * nl + j;
            keys[t] = ri | edofs[e * nl + j];
            idx[t] = (uint32_t)t;
            vals[t] = s;
        }
    }
}

// Element load vectors: f_e[i] = sum_q w_q |det J| f(x_q) phi_i(x_q), with
// f = value (kind 0) or value * prod_d sin(pi x_d) (kind 1).
template <int DIM>
__global__ __launch_bounds__(256) void k_simplex_elem_rhs(size_t ne, int nl, int nq, int kind, double value,
                                                          const double *__restrict__ coords,
                                                          const uint32_t *__restrict__ nodes,
                                                          const uint32_t *__restrict__ edofs,
                                                          const double *__restrict__ tab,
                                                          const double *__restrict__ pts,
                                                          uint32_t *__restrict__ keys,
                                                          uint32_t *__restrict__ idx,
                                                          double *__restrict__ vals) {
#pragma clang fp contract(off)
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= ne) return;
    GeomD g;
    simplex_geom_dev<DIM>(coords, nodes + e * (DIM + 1), g);
    const double *W = tab, *PHI = tab + nq;
    const double adet = fabs(g.det);
    for (int i = 0; i < nl; ++i) {
        double s = 0.0;
        for (int q = 0; q < nq; ++q) {
            const double w = W[q] * adet;
            double f = value;
            if (kind == 1) {
                for (int r = 0; r < DIM; ++r) {
                    double xr = 0.0;
                    for (int c = 0; c < DIM; ++c) xr += g.J[r][c] * pts[q * DIM + c];
                    f = f * sin(M_PI * (g.x0[r] + xr));
                }
            }
            s += w * f * PHI[(size_t)q * nl + i];
        }
        const size_t t = e * (size_t)nl + i;
        keys[t] = edofs[t];
        idx[t] = (uint32_t)t;
        vals[t] = s;
    }
}

template <class K>
__global__ void k_run_heads(size_t m, const K *__restrict__ keys, uint32_t *__restrict__ head) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= m) return;
    head[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1u : 0u;
}

// One thread per key run: the run's values summed in sorted (element) order.
__global__ void k_csr_runs(size_t m, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ perm,
                           const double *__restrict__ vals, const uint32_t *__restrict__ head,
                           const uint32_t *__restrict__ pos, uint32_t *__restrict__ col_idx,
                           double *__restrict__ values, unsigned long long *__restrict__ rowcnt) {
#pragma clang fp contract(off)
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= m || !head[t]) return;
    double s = vals[perm[t]];
    for (size_t u = t + 1; u < m && !head[u]; ++u) s += vals[perm[u]];
    const uint32_t o = pos[t];
    col_idx[o] = (uint32_t)(keys[t] & 0xffffffffu);
    values[o] = s;
    atomicAdd(&rowcnt[(keys[t] >> 32) + 1], 1ull);
}

__global__ void k_rhs_runs(size_t m, const uint32_t *__restrict__ keys, const uint32_t *__restrict__ perm,
                           const double *__restrict__ vals, const uint32_t *__restrict__ head,
                           double *__restrict__ b) {
#pragma clang fp contract(off)
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= m || !head[t]) return;
    double s = vals[perm[t]];
    for (size_t u = t + 1; u < m && !head[u]; ++u) s += vals[perm[u]];
    b[keys[t]] = s;
}

static unsigned nblocks(size_t n) { return (unsigned)((n + 255) / 256 ? (n + 255) / 256 : 1); }

// Device copies of the mesh arrays and the reference tables.
struct SimplexDev {
    DevBuf coords, nodes, edofs, tab, pts;
    SimplexDev(femgpu_ctx *c, const SimplexSetup &S, uint64_t n_nodes, const double *coords_h, uint64_t ne,
               const uint32_t *nodes_h, const uint32_t *edofs_h) {
        hipStream_t s = c->stream;
        coords.alloc(n_nodes * S.dim * 8);
        nodes.alloc(ne * (S.dim + 1) * 4);
        edofs.alloc(ne * S.nl * 4);
        tab.alloc(S.tab.size() * 8);
        pts.alloc(S.pts.size() * 8);
        if (n_nodes) FG_HIP(hipMemcpyAsync(coords.ptr, coords_h, n_nodes * S.dim * 8, hipMemcpyHostToDevice, s));
        if (ne) {
            FG_HIP(hipMemcpyAsync(nodes.ptr, nodes_h, ne * (S.dim + 1) * 4, hipMemcpyHostToDevice, s));
            FG_HIP(hipMemcpyAsync(edofs.ptr, edofs_h, ne * S.nl * 4, hipMemcpyHostToDevice, s));
        }
        FG_HIP(hipMemcpyAsync(tab.ptr, S.tab.data(), S.tab.size() * 8, hipMemcpyHostToDevice, s));
        FG_HIP(hipMemcpyAsync(pts.ptr, S.pts.data(), S.pts.size() * 8, hipMemcpyHostToDevice, s));
    }
};

// Stable radix sort of (key, index) pairs on the context stream.
template <class K>
static void sort_pairs(femgpu_ctx *c, const K *kin, K *kout, const uint32_t *iin, uint32_t *iout, size_t m,
                       int end_bit) {
    size_t tmp = 0;
    FG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, iin, iout, m, 0, end_bit, c->stream));
    DevBuf t(tmp ? tmp : 1);
    FG_HIP(hipcub::DeviceRadixSort::SortPairs(t.ptr, tmp, kin, kout, iin, iout, m, 0, end_bit, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));  // before the temporary is freed
}

static int bits_for(uint64_t n) {
    int b = 1;
    while (b < 64 && (1ull << b) < n) ++b;
    return b;
}

}  // namespace femgpu

using namespace femgpu;

static int assemble_simplex(femgpu_ctx *c, int dim, int order, uint64_t n_nodes, const double *coords,
                            uint64_t n_elem, const uint32_t *elem_nodes, const uint32_t *elem_dofs,
                            uint64_t n_dofs, int form, double coef, const double *coef_qp, int quad_order,
                            femgpu_op **out) {
    return guard([&] {
        FG_REQUIRE(c && out, FEMGPU_ERR_INVALID_ARG, "csr_assemble_simplex: NULL argument");
        FG_REQUIRE(form == FEMGPU_MASS || form == FEMGPU_DIFFUSION, FEMGPU_ERR_INVALID_ARG,
                   "csr_assemble_simplex: form must be MASS or DIFFUSION");
        SimplexSetup S(dim, order, quad_order);
        check_simplex_mesh(dim, order, n_nodes, coords, n_elem, elem_nodes, elem_dofs, n_dofs, S.nl);
        FG_REQUIRE(n_dofs < (1ull << 32), FEMGPU_ERR_INVALID_ARG, "csr_assemble_simplex: too many DOFs");
        const size_t m = (size_t)n_elem * S.nl * S.nl;
        FG_REQUIRE(m < (1ull << 31) && n_dofs < (1ull << 31), FEMGPU_ERR_INVALID_ARG,
                   "csr_assemble_simplex: too many triplets");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        SimplexDev D(c, S, n_nodes, coords, n_elem, elem_nodes, elem_dofs);
        DevBuf dcq;  // coefficient per (element, quadrature point), when given
        if (coef_qp && n_elem) {
            dcq.alloc((size_t)n_elem * S.nq * 8);
            FG_HIP(hipMemcpyAsync(dcq.ptr, coef_qp, (size_t)n_elem * S.nq * 8, hipMemcpyHostToDevice, s));
        }
        DevBuf keys(m * 8 + 8), keys2(m * 8 + 8), idx(m * 4 + 4), perm(m * 4 + 4), vals(m * 8 + 8);
        DevBuf head(m * 4 + 4), pos(m * 4 + 4);
        DevBuf rp((n_dofs + 1) * 8);
        FG_HIP(hipMemsetAsync(rp.ptr, 0, (n_dofs + 1) * 8, s));
        uint64_t nnz = 0;
        DevBuf ci, vv;
        if (m) {
            if (dim == 2)
                hipLaunchKernelGGL(k_simplex_elem_mat<2>, dim3(nblocks(n_elem)), dim3(256), 0, s, (size_t)n_elem,
                                   S.nl, S.nq, form, coef, (const double *)dcq.ptr, D.coords.as<double>(),
                                   D.nodes.as<uint32_t>(),
                                   D.edofs.as<uint32_t>(), D.tab.as<double>(), keys.as<uint64_t>(),
                                   idx.as<uint32_t>(), vals.as<double>());
            else
                hipLaunchKernelGGL(k_simplex_elem_mat<3>, dim3(nblocks(n_elem)), dim3(256), 0, s, (size_t)n_elem,
                                   S.nl, S.nq, form, coef, (const double *)dcq.ptr, D.coords.as<double>(),
                                   D.nodes.as<uint32_t>(),
                                   D.edofs.as<uint32_t>(), D.tab.as<double>(), keys.as<uint64_t>(),
                                   idx.as<uint32_t>(), vals.as<double>());
            FG_HIP(hipGetLastError());
            sort_pairs<uint64_t>(c, keys.as<uint64_t>(), keys2.as<uint64_t>(), idx.as<uint32_t>(),
                                 perm.as<uint32_t>(), m, 32 + bits_for(n_dofs));
            hipLaunchKernelGGL(k_run_heads<uint64_t>, dim3(nblocks(m)), dim3(256), 0, s, m,
                               (const uint64_t *)keys2.as<uint64_t>(), head.as<uint32_t>());
            size_t tmp = 0;
            FG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, head.as<uint32_t>(), pos.as<uint32_t>(), m, s));
            {
                DevBuf t(tmp ? tmp : 1);
                FG_HIP(hipcub::DeviceScan::ExclusiveSum(t.ptr, tmp, head.as<uint32_t>(), pos.as<uint32_t>(), m, s));
                uint32_t last[2];
                FG_HIP(hipMemcpyAsync(&last[0], pos.as<uint32_t>() + (m - 1), 4, hipMemcpyDeviceToHost, s));
                FG_HIP(hipMemcpyAsync(&last[1], head.as<uint32_t>() + (m - 1), 4, hipMemcpyDeviceToHost, s));
                FG_HIP(hipStreamSynchronize(s));
                nnz = (uint64_t)last[0] + last[1];
            }
            ci.alloc(nnz * 4);
            vv.alloc(nnz * 8);
            hipLaunchKernelGGL(k_csr_runs, dim3(nblocks(m)), dim3(256), 0, s, m, (const uint64_t *)keys2.as<uint64_t>(),
                               (const uint32_t *)perm.as<uint32_t>(), (const double *)vals.as<double>(),
                               (const uint32_t *)head.as<uint32_t>(), (const uint32_t *)pos.as<uint32_t>(),
                               ci.as<uint32_t>(), vv.as<double>(), rp.as<unsigned long long>());
            tmp = 0;
            FG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, rp.as<unsigned long long>(),
                                                    rp.as<unsigned long long>(), n_dofs + 1, s));
            DevBuf t(tmp ? tmp : 1);
            FG_HIP(hipcub::DeviceScan::InclusiveSum(t.ptr, tmp, rp.as<unsigned long long>(),
                                                    rp.as<unsigned long long>(), n_dofs + 1, s));
            FG_HIP(hipStreamSynchronize(s));
        }
        *out = csr_adopt(c, n_dofs, nnz, std::move(rp), std::move(ci), std::move(vv));
        return 0;
    });
}

extern "C" {

int femgpu_csr_assemble_simplex(femgpu_ctx *c, int dim, int order, uint64_t n_nodes, const double *coords,
                                uint64_t n_elem, const uint32_t *elem_nodes, const uint32_t *elem_dofs,
                                uint64_t n_dofs, int form, double coef, int quad_order, femgpu_op **out) {
    return assemble_simplex(c, dim, order, n_nodes, coords, n_elem, elem_nodes, elem_dofs, n_dofs, form, coef,
                            nullptr, quad_order, out);
}

int femgpu_csr_assemble_simplex_qp(femgpu_ctx *c, int dim, int order, uint64_t n_nodes, const double *coords,
                                   uint64_t n_elem, const uint32_t *elem_nodes, const uint32_t *elem_dofs,
                                   uint64_t n_dofs, int form, const double *coef_qp, int quad_order,
                                   femgpu_op **out) {
    if (!coef_qp) return fail(FEMGPU_ERR_INVALID_ARG, "csr_assemble_simplex_qp: coef_qp is NULL");
    return assemble_simplex(c, dim, order, n_nodes, coords, n_elem, elem_nodes, elem_dofs, n_dofs, form, 1.0,
                            coef_qp, quad_order, out);
}

int femgpu_rhs_assemble_simplex(femgpu_ctx *c, int dim, int order, uint64_t n_nodes, const double *coords,
                                uint64_t n_elem, const uint32_t *elem_nodes, const uint32_t *elem_dofs,
                                uint64_t n_dofs, int source, double value, int quad_order, void *b,
                                femgpu_mem mem) {
    return guard([&] {
        FG_REQUIRE(c && b, FEMGPU_ERR_INVALID_ARG, "rhs_assemble_simplex: NULL argument");
        FG_REQUIRE(source == FEMGPU_SOURCE_CONST || source == FEMGPU_SOURCE_SINPROD, FEMGPU_ERR_INVALID_ARG,
                   "rhs_assemble_simplex: unknown source kind");
        SimplexSetup S(dim, order, quad_order);
        check_simplex_mesh(dim, order, n_nodes, coords, n_elem, elem_nodes, elem_dofs, n_dofs, S.nl);
        FG_REQUIRE(n_dofs < (1ull << 32), FEMGPU_ERR_INVALID_ARG, "rhs_assemble_simplex: too many DOFs");
        const size_t m = (size_t)n_elem * S.nl;
        FG_REQUIRE(m < (1ull << 31), FEMGPU_ERR_INVALID_ARG, "rhs_assemble_simplex: too many entries");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        SimplexDev D(c, S, n_nodes, coords, n_elem, elem_nodes, elem_dofs);
        DevBuf out(mem == FEMGPU_DEVICE ? 0 : n_dofs * 8);
        double *bd = mem == FEMGPU_DEVICE ? (double *)b : out.as<double>();
        if (n_dofs) FG_HIP(hipMemsetAsync(bd, 0, n_dofs * 8, s));
        if (m) {
            DevBuf keys(m * 4), keys2(m * 4), idx(m * 4), perm(m * 4), vals(m * 8), head(m * 4);
            if (dim == 2)
                hipLaunchKernelGGL(k_simplex_elem_rhs<2>, dim3(nblocks(n_elem)), dim3(256), 0, s, (size_t)n_elem,
                                   S.nl, S.nq, source, value, D.coords.as<double>(), D.nodes.as<uint32_t>(),
                                   D.edofs.as<uint32_t>(), D.tab.as<double>(), D.pts.as<double>(),
                                   keys.as<uint32_t>(), idx.as<uint32_t>(), vals.as<double>());
            else
                hipLaunchKernelGGL(k_simplex_elem_rhs<3>, dim3(nblocks(n_elem)), dim3(256), 0, s, (size_t)n_elem,
                                   S.nl, S.nq, source, value, D.coords.as<double>(), D.nodes.as<uint32_t>(),
                                   D.edofs.as<uint32_t>(), D.tab.as<double>(), D.pts.as<double>(),
                                   keys.as<uint32_t>(), idx.as<uint32_t>(), vals.as<double>());
            FG_HIP(hipGetLastError());
            sort_pairs<uint32_t>(c, keys.as<uint32_t>(), keys2.as<uint32_t>(), idx.as<uint32_t>(),
                                 perm.as<uint32_t>(), m, bits_for(n_dofs));
            hipLaunchKernelGGL(k_run_heads<uint32_t>, dim3(nblocks(m)), dim3(256), 0, s, m,
                               (const uint32_t *)keys2.as<uint32_t>(), head.as<uint32_t>());
            hipLaunchKernelGGL(k_rhs_runs, dim3(nblocks(m)), dim3(256), 0, s, m, (const uint32_t *)keys2.as<uint32_t>(),
                               (const uint32_t *)perm.as<uint32_t>(), (const double *)vals.as<double>(),
                               (const uint32_t *)head.as<uint32_t>(), bd);
            FG_HIP(hipGetLastError());
        }
        if (mem != FEMGPU_DEVICE && n_dofs) FG_HIP(hipMemcpyAsync(b, bd, n_dofs * 8, hipMemcpyDeviceToHost, s));
        FG_HIP(hipStreamSynchronize(s));
        return 0;
    });
}

}  // extern "C"
