// Host-side reference-element tables and affine geometry of Tri3 / Tet4
// simplices, shared by the simplex PA operator and the device assembly.
#pragma once
#include "basis.hpp"
#include "internal.hpp"

#include <cmath>
#include <vector>

namespace femgpu {

struct SimplexSetup {
    int dim, order, nl, nq;
    std::vector<double> tab;  // [nq] weights, [nq][nl] phi, [nq][nl][dim] reference gradients
    std::vector<double> w, phi, grad, pts;

    SimplexSetup(int dim_, int order_, int quad_order) : dim(dim_), order(order_) {
        FG_REQUIRE(dim == 2 || dim == 3, FEMGPU_ERR_INVALID_ARG, "simplex: dim must be 2 or 3");
        FG_REQUIRE(order >= 1 && order <= 3, FEMGPU_ERR_UNSUPPORTED, "simplex: order must be 1..3");
        nl = simplex_nlocal(dim, order);
        double p[120], wts[40];
        nq = dim == 2 ? tri_rule(quad_order, p, wts) : tet_rule(quad_order, p, wts);
        w.assign(wts, wts + nq);
        pts.assign(p, p + (size_t)nq * dim);
        phi.assign((size_t)nq * nl, 0.0);
        grad.assign((size_t)nq * nl * dim, 0.0);
        for (int q = 0; q < nq; ++q)
            simplex_basis(dim, order, p + dim * q, &phi[(size_t)q * nl], &grad[(size_t)q * nl * dim]);
        tab = w;
        tab.insert(tab.end(), phi.begin(), phi.end());
        tab.insert(tab.end(), grad.begin(), grad.end());
    }

    // Affine map of ElementTransformation::from_simplex_nodes
    // (mesh/src/transformation.rs:34-64): J columns x_k - x_0; I = J^-1.
    double jac(const double *coords, const uint32_t *nodes, double J[3][3], double I[3][3]) const {
        for (int c = 0; c < dim; ++c)
            for (int r = 0; r < dim; ++r)
                J[r][c] = coords[dim * nodes[c + 1] + r] - coords[dim * nodes[0] + r];
        double det;
        if (dim == 2) {
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            I[0][0] = J[1][1] / det; I[0][1] = -J[0][1] / det;
            I[1][0] = -J[1][0] / det; I[1][1] = J[0][0] / det;
        } else {
            det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                  J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            const double id = 1.0 / det;
            I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
            I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
            I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
            I[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
            I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
            I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
            I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
            I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
            I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        }
        return det;
    }
};

void check_simplex_mesh(int dim, int order, uint64_t n_nodes, const double *coords, uint64_t n_elem,
                        const uint32_t *nodes, const uint32_t *edofs, uint64_t n_dofs, int nl);

}  // namespace femgpu
