// Host-side quadrature rules and reference-element bases used at operator
// set-up.  Conventions follow fem-rs: Gauss-Legendre on [-1,1]
// (element/src/quadrature.rs:12-31, extended to n = 5 by the closed form),
// equispaced 1-D Lagrange nodes on [-1,1] (QuadQ2 nodes, lagrange/quad.rs),
// triangle / tetrahedron rules and P1..P3 simplex bases (quadrature.rs:169-330,
// lagrange/tri.rs, lagrange/tet.rs).
#pragma once
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace femgpu {

// n = 1..5 Gauss-Legendre points, ascending.
inline void gauss_legendre(int n, double *x, double *w) {
    switch (n) {
    case 1: x[0] = 0.0; w[0] = 2.0; return;
    case 2: { const double a = 1.0 / std::sqrt(3.0); x[0] = -a; x[1] = a; w[0] = w[1] = 1.0; return; }
    case 3: {
        const double a = std::sqrt(3.0 / 5.0);
        x[0] = -a; x[1] = 0.0; x[2] = a; w[0] = w[2] = 5.0 / 9.0; w[1] = 8.0 / 9.0; return;
    }
    case 4: {
        const double a = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
        const double b = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
        const double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
        x[0] = -b; x[1] = -a; x[2] = a; x[3] = b; w[0] = w[3] = wb; w[1] = w[2] = wa; return;
    }
    case 5: {
        // roots of P5: 0, +-(1/3) sqrt(5 -+ 2 sqrt(10/7)); weights 128/225, (322 +- 13 sqrt 70)/900
        const double a = std::sqrt(5.0 - 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        const double b = std::sqrt(5.0 + 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        const double wa = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        const double wb = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0;
        x[0] = -b; x[1] = -a; x[2] = 0.0; x[3] = a; x[4] = b;
        w[0] = w[4] = wb; w[1] = w[3] = wa; w[2] = 128.0 / 225.0; return;
    }
    case 6: {
        // Golub-Welsch would be general; n = 6 by Newton on P6 from Chebyshev guesses.
        for (int i = 0; i < 6; ++i) {
            double z = std::cos(M_PI * (i + 0.75) / 6.5), dp = 1.0;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= 6; ++k) { double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
                dp = 6 * (z * p1 - p0) / (z * z - 1.0);
                const double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16) break;
            }
            x[5 - i] = z; w[5 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
        }
        return;
    }
    default: throw std::invalid_argument("gauss_legendre: n must be 1..6");
    }
}

// B[q*(p+1)+i] = l_i(x_q), G = l_i'(x_q), nodes -1 + 2i/p.
inline void lagrange_1d(int p, int nq, const double *xq, double *B, double *G) {
    const int n1 = p + 1;
    double nodes[16];
    for (int i = 0; i < n1; ++i) nodes[i] = -1.0 + 2.0 * i / p;
    for (int q = 0; q < nq; ++q) {
        const double x = xq[q];
        for (int i = 0; i < n1; ++i) {
            double v = 1.0, ds = 0.0;
            for (int j = 0; j < n1; ++j)
                if (j != i) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
            for (int m = 0; m < n1; ++m) {
                if (m == i) continue;
                double t = 1.0 / (nodes[i] - nodes[m]);
                for (int j = 0; j < n1; ++j)
                    if (j != i && j != m) t *= (x - nodes[j]) / (nodes[i] - nodes[j]);
                ds += t;
            }
            B[q * n1 + i] = v;
            G[q * n1 + i] = ds;
        }
    }
}

// Triangle rule (quadrature.rs:169-246); returns points (xi,eta) and weights.
inline int tri_rule(int order, double *p, double *w) {
    if (order <= 1) { p[0] = p[1] = 1.0 / 3.0; w[0] = 0.5; return 1; }
    if (order <= 3) {
        const double a = 1.0 / 6.0, b = 2.0 / 3.0;
        const double pts[6] = {a, a, b, a, a, b};
        std::memcpy(p, pts, sizeof pts);
        w[0] = w[1] = w[2] = a;
        return 3;
    }
    if (order <= 5) {
        const double s1 = 0.10128650732345633, s2 = 0.7974269853530872;
        const double s3 = 0.47014206410511505, t3 = 0.05971587178976981;
        const double w1 = 0.12593918054482717 / 2.0, w2 = 0.13239415278850616 / 2.0, w3 = 0.225 / 2.0;
        const double pts[14] = {s1, s1, s2, s1, s1, s2, s3, s3, t3, s3, s3, t3, 1.0 / 3.0, 1.0 / 3.0};
        const double wts[7] = {w1, w1, w1, w2, w2, w2, w3};
        std::memcpy(p, pts, sizeof pts); std::memcpy(w, wts, sizeof wts);
        return 7;
    }
    const double a1 = 0.06308901449150223, b1 = 1.0 - 2.0 * a1, w1 = 0.02542245318510341;
    const double a2 = 0.24928674517091043, b2 = 1.0 - 2.0 * a2, w2 = 0.05839313786118756;
    const double r3 = 0.6365024991213996, s3 = 0.3103524510337858, t3 = 1.0 - r3 - s3;
    const double w3 = 0.041425537809186785;
    const double pts[24] = {a1, a1, b1, a1, a1, b1, a2, a2, b2, a2, a2, b2,
                            s3, t3, r3, t3, t3, r3, s3, r3, r3, s3, t3, s3};
    const double wts[12] = {w1, w1, w1, w2, w2, w2, w3, w3, w3, w3, w3, w3};
    std::memcpy(p, pts, sizeof pts); std::memcpy(w, wts, sizeof wts);
    return 12;
}

// Grundmann-Moller rule of index s on the unit tetrahedron (exact for degree
// 2s+1; Grundmann & Moller, SIAM J. Numer. Anal. 15, 1978), weights scaled to
// sum to 1/6.
inline int gm_tet(int s, double *p, double *w) {
    auto fact = [](int n) { double f = 1.0; for (int i = 2; i <= n; ++i) f *= i; return f; };
    const int d = 3;
    int np = 0;
    for (int i = 0; i <= s; ++i) {
        const int si = s - i;
        const double m = 2.0 * si + d + 1;
        const double wi = std::pow(-1.0, i) * std::pow(m, 2 * s + 1) / (fact(i) * fact(d + 2 * s - i + 1));
        for (int j1 = 0; j1 <= si; ++j1)
            for (int j2 = 0; j2 <= si - j1; ++j2)
                for (int j3 = 0; j3 <= si - j1 - j2; ++j3) {
                    p[3 * np] = (2.0 * j1 + 1.0) / m;
                    p[3 * np + 1] = (2.0 * j2 + 1.0) / m;
                    p[3 * np + 2] = (2.0 * j3 + 1.0) / m;
                    w[np++] = wi;
                }
    }
    double sum = 0.0;
    for (int i = 0; i < np; ++i) sum += w[i];
    for (int i = 0; i < np; ++i) w[i] *= (1.0 / 6.0) / sum;
    return np;
}

inline int tet_rule(int order, double *p, double *w) {
    if (order <= 1) { p[0] = p[1] = p[2] = 0.25; w[0] = 1.0 / 6.0; return 1; }
    if (order <= 2) {
        const double a = 0.1381966011250105, b = 0.5854101966249685;
        const double pts[12] = {a, a, a, b, a, a, a, b, a, a, a, b};
        std::memcpy(p, pts, sizeof pts);
        for (int i = 0; i < 4; ++i) w[i] = 1.0 / 24.0;
        return 4;
    }
    return gm_tet(order <= 5 ? 2 : 3, p, w);
}

inline int simplex_nlocal(int dim, int order) {
    return dim == 2 ? (order + 1) * (order + 2) / 2 : (order + 1) * (order + 2) * (order + 3) / 6;
}

// P1..P3 Lagrange basis on the reference simplex in barycentric form
// (lambda_0 = 1 - sum xi, lambda_k = xi_{k-1}); local DOF order as TriP*/TetP*.
inline int simplex_basis(int dim, int order, const double *xi, double *phi, double *grad) {
    const int nv = dim + 1;
    double L[4], dL[4][3];
    L[0] = 1.0;
    for (int k = 0; k < dim; ++k) { L[0] -= xi[k]; L[k + 1] = xi[k]; }
    for (int a = 0; a < nv; ++a)
        for (int k = 0; k < dim; ++k) dL[a][k] = a == 0 ? -1.0 : (a == k + 1 ? 1.0 : 0.0);
    static const int tri_e[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    static const int tet_e[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    const int ne = dim == 2 ? 3 : 6;
    auto E = [&](int e, int s) { return dim == 2 ? tri_e[e][s] : tet_e[e][s]; };
    auto setg = [&](int idx, double ca, int a, double cb, int b) {
        for (int k = 0; k < dim; ++k) grad[idx * dim + k] = ca * dL[a][k] + cb * dL[b][k];
    };
    int n = 0;
    if (order == 1) {
        for (int a = 0; a < nv; ++a) {
            phi[a] = L[a];
            for (int k = 0; k < dim; ++k) grad[a * dim + k] = dL[a][k];
        }
        return nv;
    }
    if (order == 2) {
        for (int a = 0; a < nv; ++a) {
            phi[a] = L[a] * (2.0 * L[a] - 1.0);
            for (int k = 0; k < dim; ++k) grad[a * dim + k] = (4.0 * L[a] - 1.0) * dL[a][k];
        }
        n = nv;
        for (int e = 0; e < ne; ++e) {
            const int a = E(e, 0), b = E(e, 1);
            phi[n] = 4.0 * L[a] * L[b];
            setg(n, 4.0 * L[b], a, 4.0 * L[a], b);
            ++n;
        }
        return n;
    }
    if (order != 3) throw std::invalid_argument("simplex_basis: order must be 1..3");
    for (int a = 0; a < nv; ++a) {
        phi[a] = 0.5 * L[a] * (3.0 * L[a] - 1.0) * (3.0 * L[a] - 2.0);
        const double dp = 0.5 * (27.0 * L[a] * L[a] - 18.0 * L[a] + 2.0);
        for (int k = 0; k < dim; ++k) grad[a * dim + k] = dp * dL[a][k];
    }
    n = nv;
    for (int e = 0; e < ne; ++e) {
        const int a = E(e, 0), b = E(e, 1);
        phi[n] = 4.5 * L[a] * L[b] * (3.0 * L[a] - 1.0);
        setg(n, 4.5 * L[b] * (6.0 * L[a] - 1.0), a, 4.5 * L[a] * (3.0 * L[a] - 1.0), b);
        ++n;
        phi[n] = 4.5 * L[a] * L[b] * (3.0 * L[b] - 1.0);
        setg(n, 4.5 * L[b] * (3.0 * L[b] - 1.0), a, 4.5 * L[a] * (6.0 * L[b] - 1.0), b);
        ++n;
    }
    static const int tri_f[1][3] = {{0, 1, 2}};
    static const int tet_f[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    const int nf = dim == 2 ? 1 : 4;
    for (int f = 0; f < nf; ++f) {
        const int a = dim == 2 ? tri_f[f][0] : tet_f[f][0];
        const int b = dim == 2 ? tri_f[f][1] : tet_f[f][1];
        const int c = dim == 2 ? tri_f[f][2] : tet_f[f][2];
        phi[n] = 27.0 * L[a] * L[b] * L[c];
        for (int k = 0; k < dim; ++k)
            grad[n * dim + k] = 27.0 * (L[b] * L[c] * dL[a][k] + L[a] * L[c] * dL[b][k] +
                                        L[a] * L[b] * dL[c][k]);
        ++n;
    }
    return n;
}

}  // namespace femgpu
