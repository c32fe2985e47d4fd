/*
 * femref.c — CPU oracle (test infrastructure only; see femref.h).
 *
 * Restates the reference algorithms of javagg/fem-rs in plain C.  Each
 * function names the reference file:line it follows; the paths are relative
 * to the fem-rs checkout.  Pieces with no reference counterpart (hex Qp,
 * sum factorisation, Gauss-Legendre n = 5, textbook Chebyshev, GRF, conv3d)
 * are marked so and are pinned by the tests in tests/test_oracle_*.py.
 */
#include "femref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FR_PI 3.14159265358979323846

/* ======================================================================= */
/* RNG: Philox4x32-10                                                       */
/* ======================================================================= */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void fr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    memcpy(out, c, sizeof c);
}

double fr_uniform(uint64_t seed, uint32_t stream, uint64_t index) {
    uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    fr_philox4x32_10(ctr, key, o);
    uint64_t bits = (((uint64_t)o[0] << 32) | o[1]) >> 11;
    return (double)bits * (1.0 / 9007199254740992.0);
}

/* ======================================================================= */
/* Quadrature                                                               */
/* ======================================================================= */
/* element/src/quadrature.rs:12-31 for n <= 4; n = 5 closed form
 * (nodes 0, +-1/3 sqrt(5 -+ 2 sqrt(10/7)), weights 128/225, (322 +- 13 sqrt 70)/900). */
int fr_gauss_legendre(int n, double *x, double *w) {
    switch (n) {
    case 1: x[0] = 0.0; w[0] = 2.0; return 0;
    case 2: {
        double a = 1.0 / sqrt(3.0);
        x[0] = -a; x[1] = a; w[0] = 1.0; w[1] = 1.0; return 0;
    }
    case 3: {
        double a = sqrt(3.0 / 5.0);
        x[0] = -a; x[1] = 0.0; x[2] = a;
        w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0; return 0;
    }
    case 4: {
        double a = sqrt(3.0 / 7.0 - 2.0 / 7.0 * sqrt(6.0 / 5.0));
        double b = sqrt(3.0 / 7.0 + 2.0 / 7.0 * sqrt(6.0 / 5.0));
        double wa = (18.0 + sqrt(30.0)) / 36.0, wb = (18.0 - sqrt(30.0)) / 36.0;
        x[0] = -b; x[1] = -a; x[2] = a; x[3] = b;
        w[0] = wb; w[1] = wa; w[2] = wa; w[3] = wb; return 0;
    }
    case 5: {
        double a = sqrt(5.0 - 2.0 * sqrt(10.0 / 7.0)) / 3.0;
        double b = sqrt(5.0 + 2.0 * sqrt(10.0 / 7.0)) / 3.0;
        double wa = (322.0 + 13.0 * sqrt(70.0)) / 900.0;
        double wb = (322.0 - 13.0 * sqrt(70.0)) / 900.0;
        x[0] = -b; x[1] = -a; x[2] = 0.0; x[3] = a; x[4] = b;
        w[0] = wb; w[1] = wa; w[2] = 128.0 / 225.0; w[3] = wa; w[4] = wb; return 0;
    }
    case 6: {
        /* no closed form: Newton on P6 (three-term recurrence) from the
         * Chebyshev-like guesses cos(pi (i + 3/4) / (n + 1/2)); w = 2 / ((1 - x^2) P6'(x)^2) */
        for (int i = 0; i < 6; ++i) {
            double z = cos(FR_PI * (i + 0.75) / 6.5), dp = 1.0;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= 6; ++k) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                dp = 6.0 * (z * p1 - p0) / (z * z - 1.0);
                double dz = p1 / dp;
                z -= dz;
                if (fabs(dz) < 1e-16) break;
            }
            x[5 - i] = z;
            w[5 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
        }
        return 0;
    }
    default: return -1;
    }
}

/* quadrature.rs:169-246 */
int fr_tri_rule(int order, double *p, double *w) {
    if (order <= 1) { p[0] = 1.0 / 3.0; p[1] = 1.0 / 3.0; w[0] = 0.5; return 1; }
    if (order <= 3) {
        double a = 1.0 / 6.0, b = 2.0 / 3.0;
        double pts[6] = {a, a, b, a, a, b};
        memcpy(p, pts, sizeof pts);
        w[0] = w[1] = w[2] = a; return 3;
    }
    if (order <= 5) {
        double s1 = 0.10128650732345633, s2 = 0.7974269853530872;
        double s3 = 0.47014206410511505, t3 = 0.05971587178976981;
        double w1 = 0.12593918054482717 / 2.0, w2 = 0.13239415278850616 / 2.0, w3 = 0.225 / 2.0;
        double pts[14] = {s1, s1, s2, s1, s1, s2, s3, s3, t3, s3, s3, t3, 1.0 / 3.0, 1.0 / 3.0};
        double wts[7] = {w1, w1, w1, w2, w2, w2, w3};
        memcpy(p, pts, sizeof pts); memcpy(w, wts, sizeof wts); return 7;
    }
    {
        double a1 = 0.06308901449150223, b1 = 1.0 - 2.0 * a1, w1 = 0.02542245318510341;
        double a2 = 0.24928674517091043, b2 = 1.0 - 2.0 * a2, w2 = 0.05839313786118756;
        double r3 = 0.6365024991213996, s3 = 0.3103524510337858, t3 = 1.0 - r3 - s3;
        double w3 = 0.041425537809186785;
        double pts[24] = {a1, a1, b1, a1, a1, b1, a2, a2, b2, a2, a2, b2,
                          s3, t3, r3, t3, t3, r3, s3, r3, r3, s3, t3, s3};
        double wts[12] = {w1, w1, w1, w2, w2, w2, w3, w3, w3, w3, w3, w3};
        memcpy(p, pts, sizeof pts); memcpy(w, wts, sizeof wts); return 12;
    }
}

static double fact(int n) { double f = 1.0; for (int i = 2; i <= n; ++i) f *= i; return f; }

/* Grundmann-Moller rule on the unit tetrahedron, index s (exact for degree
 * 2s+1), closed-form weights (Grundmann & Moller, SIAM J. Numer. Anal. 15
 * (1978)).  The reference (quadrature.rs:280-330) builds the same lattice and
 * solves for per-level weights; for an exact rule these coincide. */
static int gm_tet(int s, double *p, double *w) {
    const int d = 3;
    int np = 0;
    for (int i = 0; i <= s; ++i) {
        int si = s - i;
        double m = 2.0 * si + d + 1;
        /* level weight (-1)^i m^(2s+1) / (i! (d+2s-i+1)!), up to a common
         * factor that the normalisation below fixes */
        double wi = pow(-1.0, i) * pow(m, 2 * s + 1) / (fact(i) * fact(d + 2 * s - i + 1));
        /* enumerate compositions j0+j1+j2+j3 = si */
        for (int j1 = 0; j1 <= si; ++j1)
            for (int j2 = 0; j2 <= si - j1; ++j2)
                for (int j3 = 0; j3 <= si - j1 - j2; ++j3) {
                    p[3 * np + 0] = (2.0 * j1 + 1.0) / m;
                    p[3 * np + 1] = (2.0 * j2 + 1.0) / m;
                    p[3 * np + 2] = (2.0 * j3 + 1.0) / m;
                    w[np] = wi;
                    ++np;
                }
    }
    /* normalise so that the weights sum to 1/6 */
    double sum = 0.0;
    for (int i = 0; i < np; ++i) sum += w[i];
    for (int i = 0; i < np; ++i) w[i] *= (1.0 / 6.0) / sum;
    return np;
}

/* quadrature.rs:259-279 (orders <= 2), Grundmann-Moller above. */
int fr_tet_rule(int order, double *p, double *w) {
    if (order <= 1) { p[0] = p[1] = p[2] = 0.25; w[0] = 1.0 / 6.0; return 1; }
    if (order <= 2) {
        double a = 0.1381966011250105, b = 0.5854101966249685;
        double pts[12] = {a, a, a, b, a, a, a, b, a, a, a, b};
        memcpy(p, pts, sizeof pts);
        for (int i = 0; i < 4; ++i) w[i] = 1.0 / 24.0;
        return 4;
    }
    if (order <= 5) return gm_tet(2, p, w);
    return gm_tet(3, p, w);
}

/* ======================================================================= */
/* Bases                                                                    */
/* ======================================================================= */
void fr_lagrange_1d(int p, int nq, const double *xq, double *B, double *G) {
    const int n1 = p + 1;
    double nodes[16];
    for (int i = 0; i < n1; ++i) nodes[i] = -1.0 + 2.0 * i / p;
    for (int q = 0; q < nq; ++q) {
        const double x = xq[q];
        for (int i = 0; i < n1; ++i) {
            double v = 1.0, dsum = 0.0;
            for (int j = 0; j < n1; ++j) {
                if (j == i) continue;
                v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
            }
            for (int m = 0; m < n1; ++m) {
                if (m == i) continue;
                double t = 1.0 / (nodes[i] - nodes[m]);
                for (int j = 0; j < n1; ++j) {
                    if (j == i || j == m) continue;
                    t *= (x - nodes[j]) / (nodes[i] - nodes[j]);
                }
                dsum += t;
            }
            B[q * n1 + i] = v;
            G[q * n1 + i] = dsum;
        }
    }
}

/* Barycentric restatement of lagrange/tri.rs (TriP1..P3) and lagrange/tet.rs
 * (TetP1..P3): lambda_0 = 1 - sum xi, lambda_k = xi_{k-1}. */
int fr_simplex_basis(int dim, int order, const double *xi, double *phi, double *grad) {
    const int nv = dim + 1;
    double L[4], dL[4][3];
    L[0] = 1.0;
    for (int k = 0; k < dim; ++k) L[0] -= xi[k];
    for (int k = 0; k < dim; ++k) { L[k + 1] = xi[k]; }
    for (int a = 0; a < nv; ++a)
        for (int k = 0; k < dim; ++k) dL[a][k] = (a == 0) ? -1.0 : (a == k + 1 ? 1.0 : 0.0);
    int n = 0;
#define SETG(idx, expr_a, a, expr_b, b)                                              \
    for (int k = 0; k < dim; ++k) grad[(idx) * dim + k] = (expr_a) * dL[a][k] + (expr_b) * dL[b][k];
    if (order == 1) {
        for (int a = 0; a < nv; ++a) {
            phi[a] = L[a];
            for (int k = 0; k < dim; ++k) grad[a * dim + k] = dL[a][k];
        }
        return nv;
    }
    /* edge list in the reference local order */
    static const int tri_e[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    static const int tet_e[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    const int ne = dim == 2 ? 3 : 6;
    const int (*E)[2] = dim == 2 ? tri_e : tet_e;
    if (order == 2) {
        for (int a = 0; a < nv; ++a) {
            phi[a] = L[a] * (2.0 * L[a] - 1.0);
            for (int k = 0; k < dim; ++k) grad[a * dim + k] = (4.0 * L[a] - 1.0) * dL[a][k];
        }
        n = nv;
        for (int e = 0; e < ne; ++e) {
            int a = E[e][0], b = E[e][1];
            phi[n] = 4.0 * L[a] * L[b];
            SETG(n, 4.0 * L[b], a, 4.0 * L[a], b);
            ++n;
        }
        return n;
    }
    if (order == 3) {
        for (int a = 0; a < nv; ++a) {
            phi[a] = 0.5 * L[a] * (3.0 * L[a] - 1.0) * (3.0 * L[a] - 2.0);
            double dphi = 0.5 * (27.0 * L[a] * L[a] - 18.0 * L[a] + 2.0);
            for (int k = 0; k < dim; ++k) grad[a * dim + k] = dphi * dL[a][k];
        }
        n = nv;
        for (int e = 0; e < ne; ++e) {
            int a = E[e][0], b = E[e][1];
            /* DOF near a: 9/2 La Lb (3La - 1); near b: 9/2 La Lb (3Lb - 1) */
            phi[n] = 4.5 * L[a] * L[b] * (3.0 * L[a] - 1.0);
            SETG(n, 4.5 * L[b] * (6.0 * L[a] - 1.0), a, 4.5 * L[a] * (3.0 * L[a] - 1.0), b);
            ++n;
            phi[n] = 4.5 * L[a] * L[b] * (3.0 * L[b] - 1.0);
            SETG(n, 4.5 * L[b] * (3.0 * L[b] - 1.0), a, 4.5 * L[a] * (6.0 * L[b] - 1.0), b);
            ++n;
        }
        if (dim == 2) {
            phi[n] = 27.0 * L[0] * L[1] * L[2];
            for (int k = 0; k < dim; ++k)
                grad[n * dim + k] = 27.0 * (L[1] * L[2] * dL[0][k] + L[0] * L[2] * dL[1][k] +
                                            L[0] * L[1] * dL[2][k]);
            ++n;
        } else {
            static const int tf[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
            for (int f = 0; f < 4; ++f) {
                int a = tf[f][0], b = tf[f][1], c = tf[f][2];
                phi[n] = 27.0 * L[a] * L[b] * L[c];
                for (int k = 0; k < dim; ++k)
                    grad[n * dim + k] = 27.0 * (L[b] * L[c] * dL[a][k] + L[a] * L[c] * dL[b][k] +
                                                L[a] * L[b] * dL[c][k]);
                ++n;
            }
        }
        return n;
    }
#undef SETG
    return 0;
}

/* ======================================================================= */
/* Hex meshes, coefficients, geometry                                      */
/* ======================================================================= */
void fr_hex_vertices(int nx, int ny, int nz, double amp, uint64_t seed, double *c) {
    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                size_t v = (size_t)i + (size_t)(nx + 1) * ((size_t)j + (size_t)(ny + 1) * k);
                double x = i * hx, y = j * hy, z = k * hz;
                int interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                if (interior && amp != 0.0) {
                    x += amp * hx * (2.0 * fr_uniform(seed, 1, v) - 1.0);
                    y += amp * hy * (2.0 * fr_uniform(seed, 2, v) - 1.0);
                    z += amp * hz * (2.0 * fr_uniform(seed, 3, v) - 1.0);
                }
                c[3 * v] = x; c[3 * v + 1] = y; c[3 * v + 2] = z;
            }
}

/* Random Fourier features for a squared-exponential covariance
 * sigma^2 exp(-|r|^2 / (2 l^2)) (Rahimi & Recht, NIPS 2007): w ~ N(0, I/l^2)
 * by Box-Muller, b ~ U[0, 2 pi). */
void fr_grf_features(const fr_coef *c, double *w, double *b) {
    for (int m = 0; m < c->n_features; ++m) {
        for (int k = 0; k < 3; ++k) {
            double u1 = fr_uniform(c->seed, 10 + 2 * k, (uint64_t)m);
            double u2 = fr_uniform(c->seed, 11 + 2 * k, (uint64_t)m);
            double r = sqrt(-2.0 * log(1.0 - u1));
            w[3 * m + k] = r * cos(2.0 * FR_PI * u2) / c->length;
        }
        b[m] = 2.0 * FR_PI * fr_uniform(c->seed, 20, (uint64_t)m);
    }
}

double fr_coef_eval(const fr_coef *c, const double *wf, const double *bf, const double *x) {
    switch (c->kind) {
    case FR_COEF_CONST: return c->value;
    case FR_COEF_SINPROD:
        return 1.0 + 0.5 * sin(2.0 * FR_PI * x[0]) * sin(2.0 * FR_PI * x[1]) *
                         sin(2.0 * FR_PI * x[2]);
    case FR_COEF_GRF: {
        double s = 0.0;
        for (int m = 0; m < c->n_features; ++m)
            s += cos(wf[3 * m] * x[0] + wf[3 * m + 1] * x[1] + wf[3 * m + 2] * x[2] + bf[m]);
        double g = c->sigma * sqrt(2.0 / c->n_features) * s;
        double k = exp(g);
        if (k < c->lo) k = c->lo;
        if (k > c->hi) k = c->hi;
        return k;
    }
    }
    return 0.0;
}

/* HexQ1 vertex signs (lagrange/hex.rs:27-36). */
static const double Q1S[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

static void hex_elem_vertices(const fr_hex_desc *d, const double *coords, int ex, int ey, int ez,
                              double X[8][3]) {
    for (int v = 0; v < 8; ++v) {
        int i = ex + (Q1S[v][0] > 0), j = ey + (Q1S[v][1] > 0), k = ez + (Q1S[v][2] > 0);
        size_t g = (size_t)i + (size_t)(d->nx + 1) * ((size_t)j + (size_t)(d->ny + 1) * k);
        for (int c = 0; c < 3; ++c) X[v][c] = coords[3 * g + c];
    }
}

/* isoparametric_jacobian (assembly/src/assembler.rs:96-123) with the HexQ1
 * basis (lagrange/hex.rs:40-64): J[i][d] = sum_k X_k[i] dphi_k/dxi_d. */
static double hex_jacobian(const double X[8][3], const double xi[3], double J[3][3], double xp[3]) {
    memset(J, 0, 9 * sizeof(double));
    xp[0] = xp[1] = xp[2] = 0.0;
    for (int k = 0; k < 8; ++k) {
        double fx = 1.0 + Q1S[k][0] * xi[0], fy = 1.0 + Q1S[k][1] * xi[1],
               fz = 1.0 + Q1S[k][2] * xi[2];
        double phi = 0.125 * fx * fy * fz;
        double g[3] = {0.125 * Q1S[k][0] * fy * fz, 0.125 * Q1S[k][1] * fx * fz,
                       0.125 * Q1S[k][2] * fx * fy};
        for (int i = 0; i < 3; ++i) {
            xp[i] += phi * X[k][i];
            for (int dd = 0; dd < 3; ++dd) J[i][dd] += X[k][i] * g[dd];
        }
    }
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

static void inv3(const double J[3][3], double det, double I[3][3]) {
    double id = 1.0 / det;
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

size_t fr_hex_ndofs(const fr_hex_desc *d) {
    return (size_t)(d->nx * d->p + 1) * (size_t)(d->ny * d->p + 1) * (size_t)(d->nz * d->p + 1);
}

static inline size_t hex_gdof(const fr_hex_desc *d, int ex, int ey, int ez, int i, int j, int k) {
    const size_t Nx = (size_t)d->nx * d->p + 1, Ny = (size_t)d->ny * d->p + 1;
    return (size_t)(ex * d->p + i) + Nx * ((size_t)(ey * d->p + j) + Ny * (size_t)(ez * d->p + k));
}

static inline int hex_is_bdr(const fr_hex_desc *d, size_t g) {
    const size_t Nx = (size_t)d->nx * d->p + 1, Ny = (size_t)d->ny * d->p + 1,
                 Nz = (size_t)d->nz * d->p + 1;
    size_t gx = g % Nx, gy = (g / Nx) % Ny, gz = g / (Nx * Ny);
    return gx == 0 || gx == Nx - 1 || gy == 0 || gy == Ny - 1 || gz == 0 || gz == Nz - 1;
}

typedef struct { double x[8], w[8], B[64], G[64]; double *wf, *bf; } hex_tabs;

static void hex_tabs_init(const fr_hex_desc *d, hex_tabs *t) {
    fr_gauss_legendre(d->q, t->x, t->w);
    fr_lagrange_1d(d->p, d->q, t->x, t->B, t->G);
    t->wf = t->bf = NULL;
    if (d->kappa.kind == FR_COEF_GRF) {
        t->wf = (double *)malloc(sizeof(double) * 3 * d->kappa.n_features);
        t->bf = (double *)malloc(sizeof(double) * d->kappa.n_features);
        fr_grf_features(&d->kappa, t->wf, t->bf);
    }
}
static void hex_tabs_free(hex_tabs *t) { free(t->wf); free(t->bf); }

void fr_hex_qdata(const fr_hex_desc *d, const double *coords, double *D, double *M) {
    hex_tabs t;
    hex_tabs_init(d, &t);
    const int q = d->q, Q = q * q * q;
    const size_t E = (size_t)d->nx * d->ny * d->nz;
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < E; ++e) {
        int ex = (int)(e % d->nx), ey = (int)((e / d->nx) % d->ny), ez = (int)(e / ((size_t)d->nx * d->ny));
        double X[8][3];
        hex_elem_vertices(d, coords, ex, ey, ez, X);
        for (int qz = 0; qz < q; ++qz)
            for (int qy = 0; qy < q; ++qy)
                for (int qx = 0; qx < q; ++qx) {
                    int qp = qx + q * (qy + q * qz);
                    double xi[3] = {t.x[qx], t.x[qy], t.x[qz]};
                    double w = t.w[qx] * t.w[qy] * t.w[qz];
                    double J[3][3], xp[3], I[3][3];
                    double det = hex_jacobian(X, xi, J, xp);
                    double wd = w * fabs(det);
                    if (M) M[e * Q + qp] = wd * d->rho;
                    if (D) {
                        inv3(J, det, I);
                        double kap = fr_coef_eval(&d->kappa, t.wf, t.bf, xp);
                        double s = wd * kap;
                        double *o = D + (e * Q + qp) * 6;
                        o[0] = s * (I[0][0] * I[0][0] + I[0][1] * I[0][1] + I[0][2] * I[0][2]);
                        o[1] = s * (I[0][0] * I[1][0] + I[0][1] * I[1][1] + I[0][2] * I[1][2]);
                        o[2] = s * (I[0][0] * I[2][0] + I[0][1] * I[2][1] + I[0][2] * I[2][2]);
                        o[3] = s * (I[1][0] * I[1][0] + I[1][1] * I[1][1] + I[1][2] * I[1][2]);
                        o[4] = s * (I[1][0] * I[2][0] + I[1][1] * I[2][1] + I[1][2] * I[2][2]);
                        o[5] = s * (I[2][0] * I[2][0] + I[2][1] * I[2][1] + I[2][2] * I[2][2]);
                    }
                }
    }
    hex_tabs_free(&t);
}

/* The reference PA element loop (partial.rs:151-208) generalised to hex Qp:
 * per element gather, per quadrature point (hex_rule order: first coordinate
 * slowest, quadrature.rs:405-419) isoparametric J, J^-T gradients, kappa at the
 * physical point, y_e[i] += w|J| kappa grad u . grad phi_i (+ mass term),
 * scatter-add. */
typedef struct { double *phi, *gr, *gp, *xe, *ye; size_t *gd; } refloop_scratch;

static void refloop_scratch_init(refloop_scratch *w, int nl) {
    w->phi = malloc(sizeof(double) * nl); w->gr = malloc(sizeof(double) * nl * 3);
    w->gp = malloc(sizeof(double) * nl * 3); w->xe = malloc(sizeof(double) * nl);
    w->ye = malloc(sizeof(double) * nl); w->gd = malloc(sizeof(size_t) * nl);
}
static void refloop_scratch_free(refloop_scratch *w) {
    free(w->phi); free(w->gr); free(w->gp); free(w->xe); free(w->ye); free(w->gd);
}

/* One element of the reference PA loop: gather, per quadrature point J, J^-T
 * gradients, kappa, element contribution, scatter-add into y. */
static void refloop_element(const fr_hex_desc *d, const hex_tabs *t, const double *coords,
                            const double *x, double *y, int ex, int ey, int ez, refloop_scratch *w) {
    const int n1 = d->p + 1, q = d->q, nl = n1 * n1 * n1;
    double *phi = w->phi, *gr = w->gr, *gp = w->gp, *xe = w->xe, *ye = w->ye;
    size_t *gd = w->gd;
    double X[8][3];
    hex_elem_vertices(d, coords, ex, ey, ez, X);
    for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) {
                int l = i + n1 * (j + n1 * k);
                gd[l] = hex_gdof(d, ex, ey, ez, i, j, k);
                xe[l] = (d->dirichlet && hex_is_bdr(d, gd[l])) ? 0.0 : x[gd[l]];
                ye[l] = 0.0;
            }
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b)
            for (int c = 0; c < q; ++c) {
                double xi[3] = {t->x[a], t->x[b], t->x[c]};
                double J[3][3], xp[3], I[3][3];
                double det = hex_jacobian(X, xi, J, xp);
                double wq = t->w[a] * t->w[b] * t->w[c] * fabs(det);
                inv3(J, det, I);
                for (int k = 0; k < n1; ++k)
                    for (int j = 0; j < n1; ++j)
                        for (int i = 0; i < n1; ++i) {
                            int l = i + n1 * (j + n1 * k);
                            double Bi = t->B[a * n1 + i], Bj = t->B[b * n1 + j], Bk = t->B[c * n1 + k];
                            phi[l] = Bi * Bj * Bk;
                            gr[3 * l] = t->G[a * n1 + i] * Bj * Bk;
                            gr[3 * l + 1] = Bi * t->G[b * n1 + j] * Bk;
                            gr[3 * l + 2] = Bi * Bj * t->G[c * n1 + k];
                            /* grad_phys = J^-T grad_ref (assembler.rs:128-142) */
                            for (int r = 0; r < 3; ++r) {
                                double s = 0.0;
                                for (int m = 0; m < 3; ++m) s += I[m][r] * gr[3 * l + m];
                                gp[3 * l + r] = s;
                            }
                        }
                if (d->kind & FR_DIFFUSION) {
                    double kap = fr_coef_eval(&d->kappa, t->wf, t->bf, xp);
                    double gu[3] = {0, 0, 0};
                    for (int l = 0; l < nl; ++l)
                        for (int r = 0; r < 3; ++r) gu[r] += xe[l] * gp[3 * l + r];
                    for (int l = 0; l < nl; ++l) {
                        double dot = 0.0;
                        for (int r = 0; r < 3; ++r) dot += gu[r] * gp[3 * l + r];
                        ye[l] += wq * kap * dot;
                    }
                }
                if (d->kind & FR_MASS) {
                    double u = 0.0;
                    for (int l = 0; l < nl; ++l) u += phi[l] * xe[l];
                    for (int l = 0; l < nl; ++l) ye[l] += wq * d->rho * phi[l] * u;
                }
            }
    for (int l = 0; l < nl; ++l)
        if (!(d->dirichlet && hex_is_bdr(d, gd[l]))) y[gd[l]] += ye[l];
}

/* The reference PA element loop (partial.rs:151-208) generalised to hex Qp:
 * per element gather, per quadrature point (hex_rule order: first coordinate
 * slowest, quadrature.rs:405-419) isoparametric J, J^-T gradients, kappa at the
 * physical point, y_e[i] += w|J| kappa grad u . grad phi_i (+ mass term),
 * scatter-add.  Serial, elements in lexicographic order, as the reference. */
void fr_hex_apply_refloop(const fr_hex_desc *d, const double *coords, const double *x, double *y,
                          int accumulate) {
    hex_tabs t;
    hex_tabs_init(d, &t);
    const int n1 = d->p + 1, nl = n1 * n1 * n1;
    const size_t N = fr_hex_ndofs(d);
    if (!accumulate) memset(y, 0, N * sizeof(double));
    refloop_scratch w;
    refloop_scratch_init(&w, nl);
    for (int ez = 0; ez < d->nz; ++ez)
        for (int ey = 0; ey < d->ny; ++ey)
            for (int ex = 0; ex < d->nx; ++ex) refloop_element(d, &t, coords, x, y, ex, ey, ez, &w);
    if (d->dirichlet)
        for (size_t g = 0; g < N; ++g)
            if (hex_is_bdr(d, g)) y[g] += x[g];
    refloop_scratch_free(&w);
    hex_tabs_free(&t);
}

/* The same element loop on nthreads OpenMP threads (the CPU baseline's
 * threaded form of the reference algorithm): the elements are split into the
 * 8 parity colours (ex, ey, ez mod 2), whose elements share no DOF, and each
 * colour's elements run in parallel.  Sums into y in another order than the
 * serial loop, so results agree with it to rounding. */
void fr_hex_apply_refloop_mt(const fr_hex_desc *d, const double *coords, const double *x, double *y,
                             int accumulate, int nthreads) {
    hex_tabs t;
    hex_tabs_init(d, &t);
    const int n1 = d->p + 1, nl = n1 * n1 * n1;
    const size_t N = fr_hex_ndofs(d);
    if (!accumulate) memset(y, 0, N * sizeof(double));
    for (int col = 0; col < 8; ++col) {
        const int px = col & 1, py = (col >> 1) & 1, pz = col >> 2;
        const int cx = (d->nx - px + 1) / 2, cy = (d->ny - py + 1) / 2, cz = (d->nz - pz + 1) / 2;
        const long ncol = (long)cx * cy * cz;
#pragma omp parallel num_threads(nthreads)
        {
            refloop_scratch w;
            refloop_scratch_init(&w, nl);
#pragma omp for schedule(static)
            for (long e = 0; e < ncol; ++e) {
                const int ex = px + 2 * (int)(e % cx), ey = py + 2 * (int)((e / cx) % cy),
                          ez = pz + 2 * (int)(e / ((long)cx * cy));
                refloop_element(d, &t, coords, x, y, ex, ey, ez, &w);
            }
            refloop_scratch_free(&w);
        }
    }
    if (d->dirichlet)
        for (size_t g = 0; g < N; ++g)
            if (hex_is_bdr(d, g)) y[g] += x[g];
    hex_tabs_free(&t);
}

/* Sum-factorised apply in double and float. */
#define REAL double
#define SUMFAC_NAME fr_hex_apply_sumfac
#include "femref_sumfac.inc"
#undef REAL
#undef SUMFAC_NAME
#define REAL float
#define SUMFAC_NAME fr_hex_apply_sumfac_f32
#include "femref_sumfac.inc"
#undef REAL
#undef SUMFAC_NAME

void fr_hex_diagonal(const fr_hex_desc *d, const double *D, const double *M, double *diag) {
    hex_tabs t;
    hex_tabs_init(d, &t);
    const int n1 = d->p + 1, q = d->q, Q = q * q * q;
    const size_t N = fr_hex_ndofs(d);
    memset(diag, 0, N * sizeof(double));
    for (int ez = 0; ez < d->nz; ++ez)
        for (int ey = 0; ey < d->ny; ++ey)
            for (int ex = 0; ex < d->nx; ++ex) {
                size_t e = (size_t)ex + (size_t)d->nx * (ey + (size_t)d->ny * ez);
                for (int k = 0; k < n1; ++k)
                    for (int j = 0; j < n1; ++j)
                        for (int i = 0; i < n1; ++i) {
                            double s = 0.0;
                            for (int c = 0; c < q; ++c)
                                for (int b = 0; b < q; ++b)
                                    for (int a = 0; a < q; ++a) {
                                        int qp = a + q * (b + q * c);
                                        double Bi = t.B[a * n1 + i], Bj = t.B[b * n1 + j],
                                               Bk = t.B[c * n1 + k];
                                        if ((d->kind & FR_DIFFUSION) && D) {
                                            const double *o = D + (e * Q + qp) * 6;
                                            double g0 = t.G[a * n1 + i] * Bj * Bk;
                                            double g1 = Bi * t.G[b * n1 + j] * Bk;
                                            double g2 = Bi * Bj * t.G[c * n1 + k];
                                            s += o[0] * g0 * g0 + o[3] * g1 * g1 + o[5] * g2 * g2 +
                                                 2.0 * (o[1] * g0 * g1 + o[2] * g0 * g2 + o[4] * g1 * g2);
                                        }
                                        if ((d->kind & FR_MASS) && M) {
                                            double ph = Bi * Bj * Bk;
                                            s += M[e * Q + qp] * ph * ph;
                                        }
                                    }
                            diag[hex_gdof(d, ex, ey, ez, i, j, k)] += s;
                        }
            }
    if (d->dirichlet)
        for (size_t g = 0; g < N; ++g)
            if (hex_is_bdr(d, g)) diag[g] = 1.0;
    hex_tabs_free(&t);
}

/* ---- COO -> CSR (coo.rs:76-): sort by (row, col), sum duplicates.  The
 * reference sorts unstably; here ties keep insertion order. */
typedef struct { uint32_t r, c; size_t i; } coo_key;
static int coo_cmp(const void *a, const void *b) {
    const coo_key *x = a, *y = b;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i);
}
static size_t coo_to_csr(size_t nrows, size_t nnz_raw, const uint32_t *rows, const uint32_t *cols,
                         const double *vals, uint64_t *row_ptr, uint32_t *col_idx, double *values) {
    coo_key *k = malloc(sizeof(coo_key) * (nnz_raw ? nnz_raw : 1));
    for (size_t i = 0; i < nnz_raw; ++i) { k[i].r = rows[i]; k[i].c = cols[i]; k[i].i = i; }
    qsort(k, nnz_raw, sizeof(coo_key), coo_cmp);
    size_t nnz = 0;
    for (size_t r = 0; r <= nrows; ++r) row_ptr[r] = 0;
    for (size_t t = 0; t < nnz_raw; ++t) {
        if (t > 0 && k[t].r == k[t - 1].r && k[t].c == k[t - 1].c) {
            if (values) values[nnz - 1] += vals[k[t].i];
            continue;
        }
        if (col_idx) { col_idx[nnz] = k[t].c; values[nnz] = vals[k[t].i]; }
        row_ptr[k[t].r + 1]++;
        ++nnz;
    }
    for (size_t r = 0; r < nrows; ++r) row_ptr[r + 1] += row_ptr[r];
    free(k);
    return nnz;
}

size_t fr_hex_assemble_csr(const fr_hex_desc *d, const double *coords, uint64_t *row_ptr,
                           uint32_t *col_idx, double *values) {
    hex_tabs t;
    hex_tabs_init(d, &t);
    const int n1 = d->p + 1, q = d->q, nl = n1 * n1 * n1;
    const size_t E = (size_t)d->nx * d->ny * d->nz, N = fr_hex_ndofs(d);
    const size_t raw = E * nl * nl;
    uint32_t *rr = malloc(sizeof(uint32_t) * raw), *cc = malloc(sizeof(uint32_t) * raw);
    double *vv = calloc(raw, sizeof(double));
    double *phi = malloc(sizeof(double) * nl), *gr = malloc(sizeof(double) * 3 * nl),
           *gp = malloc(sizeof(double) * 3 * nl);
    size_t *gd = malloc(sizeof(size_t) * nl);
    size_t cnt = 0;
    for (int ez = 0; ez < d->nz; ++ez)
        for (int ey = 0; ey < d->ny; ++ey)
            for (int ex = 0; ex < d->nx; ++ex) {
                double X[8][3];
                hex_elem_vertices(d, coords, ex, ey, ez, X);
                double *Ke = vv + cnt;
                for (int k = 0; k < n1; ++k)
                    for (int j = 0; j < n1; ++j)
                        for (int i = 0; i < n1; ++i) gd[i + n1 * (j + n1 * k)] = hex_gdof(d, ex, ey, ez, i, j, k);
                for (int a = 0; a < q; ++a)
                    for (int b = 0; b < q; ++b)
                        for (int c = 0; c < q; ++c) {
                            double xi[3] = {t.x[a], t.x[b], t.x[c]};
                            double J[3][3], xp[3], I[3][3];
                            double det = hex_jacobian(X, xi, J, xp);
                            double w = t.w[a] * t.w[b] * t.w[c] * fabs(det);
                            inv3(J, det, I);
                            for (int l = 0; l < nl; ++l) {
                                int i = l % n1, j = (l / n1) % n1, k = l / (n1 * n1);
                                double Bi = t.B[a * n1 + i], Bj = t.B[b * n1 + j], Bk = t.B[c * n1 + k];
                                phi[l] = Bi * Bj * Bk;
                                gr[3 * l] = t.G[a * n1 + i] * Bj * Bk;
                                gr[3 * l + 1] = Bi * t.G[b * n1 + j] * Bk;
                                gr[3 * l + 2] = Bi * Bj * t.G[c * n1 + k];
                                for (int r = 0; r < 3; ++r) {
                                    double s = 0.0;
                                    for (int m = 0; m < 3; ++m) s += I[m][r] * gr[3 * l + m];
                                    gp[3 * l + r] = s;
                                }
                            }
                            double wk = (d->kind & FR_DIFFUSION) ? w * fr_coef_eval(&d->kappa, t.wf, t.bf, xp) : 0.0;
                            double wr = (d->kind & FR_MASS) ? w * d->rho : 0.0;
                            /* diffusion.rs:36-54, mass.rs:28-40 */
                            for (int i = 0; i < nl; ++i)
                                for (int j = 0; j < nl; ++j) {
                                    double v = 0.0;
                                    if (d->kind & FR_DIFFUSION) {
                                        double dot = 0.0;
                                        for (int r = 0; r < 3; ++r) dot += gp[3 * i + r] * gp[3 * j + r];
                                        v += wk * dot;
                                    }
                                    if (d->kind & FR_MASS) v += wr * phi[i] * phi[j];
                                    Ke[i * nl + j] += v;
                                }
                        }
                for (int i = 0; i < nl; ++i)
                    for (int j = 0; j < nl; ++j) {
                        rr[cnt] = (uint32_t)gd[i]; cc[cnt] = (uint32_t)gd[j];
                        if (d->dirichlet && (hex_is_bdr(d, gd[i]) || hex_is_bdr(d, gd[j])))
                            vv[cnt] = 0.0;
                        ++cnt;
                    }
            }
    size_t nnz = coo_to_csr(N, cnt, rr, cc, vv, row_ptr, col_idx, values);
    if (d->dirichlet && col_idx) {
        for (size_t r = 0; r < N; ++r)
            if (hex_is_bdr(d, r))
                for (uint64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
                    if (col_idx[k] == r) values[k] = 1.0;
    }
    free(rr); free(cc); free(vv); free(phi); free(gr); free(gp); free(gd);
    hex_tabs_free(&t);
    return nnz;
}

void fr_hex_rhs_unit(const fr_hex_desc *d, const double *coords, double *b) {
    hex_tabs t;
    hex_tabs_init(d, &t);
    const int n1 = d->p + 1, q = d->q;
    const size_t N = fr_hex_ndofs(d);
    memset(b, 0, N * sizeof(double));
    for (int ez = 0; ez < d->nz; ++ez)
        for (int ey = 0; ey < d->ny; ++ey)
            for (int ex = 0; ex < d->nx; ++ex) {
                double X[8][3];
                hex_elem_vertices(d, coords, ex, ey, ez, X);
                for (int a = 0; a < q; ++a)
                    for (int bb = 0; bb < q; ++bb)
                        for (int c = 0; c < q; ++c) {
                            double xi[3] = {t.x[a], t.x[bb], t.x[c]};
                            double J[3][3], xp[3];
                            double det = hex_jacobian(X, xi, J, xp);
                            double w = t.w[a] * t.w[bb] * t.w[c] * fabs(det);
                            for (int k = 0; k < n1; ++k)
                                for (int j = 0; j < n1; ++j)
                                    for (int i = 0; i < n1; ++i)
                                        b[hex_gdof(d, ex, ey, ez, i, j, k)] +=
                                            w * t.B[a * n1 + i] * t.B[bb * n1 + j] * t.B[c * n1 + k];
                        }
            }
    if (d->dirichlet)
        for (size_t g = 0; g < N; ++g)
            if (hex_is_bdr(d, g)) b[g] = 0.0;
    hex_tabs_free(&t);
}

/* ======================================================================= */
/* CSR                                                                      */
/* ======================================================================= */
/* csr_row_dot_f64 (linalg/src/csr.rs:38-64) */
static inline double csr_row_dot(const fr_csr *a, const double *x, size_t row) {
    uint64_t s = a->row_ptr[row], e = a->row_ptr[row + 1], k = s;
    double sum = 0.0;
    uint64_t e4 = s + (e - s) / 4 * 4;
    while (k < e4) {
        sum += a->values[k] * x[a->col_idx[k]] + a->values[k + 1] * x[a->col_idx[k + 1]] +
               a->values[k + 2] * x[a->col_idx[k + 2]] + a->values[k + 3] * x[a->col_idx[k + 3]];
        k += 4;
    }
    while (k < e) { sum += a->values[k] * x[a->col_idx[k]]; ++k; }
    return sum;
}

void fr_csr_spmv(const fr_csr *a, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < a->nrows; ++r) y[r] = csr_row_dot(a, x, r);
}

void fr_csr_spmv_add(const fr_csr *a, double alpha, const double *x, double beta, double *y) {
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < a->nrows; ++r) y[r] = alpha * csr_row_dot(a, x, r) + beta * y[r];
}

void fr_csr_spmv_f32(size_t nrows, const uint64_t *row_ptr, const uint32_t *col_idx,
                     const float *values, const float *x, float *y) {
    for (size_t r = 0; r < nrows; ++r) {
        float s = 0.0f;
        for (uint64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += values[k] * x[col_idx[k]];
        y[r] = s;
    }
}

void fr_csr_diagonal(const fr_csr *a, double *d) {
    for (size_t r = 0; r < a->nrows; ++r) {
        d[r] = 0.0;
        for (uint64_t k = a->row_ptr[r]; k < a->row_ptr[r + 1]; ++k)
            if (a->col_idx[k] == r) { d[r] = a->values[k]; break; }
    }
}

/* ======================================================================= */
/* Simplex meshes, assembly, PA                                             */
/* ======================================================================= */
void fr_unit_square_tri(int n, double *c, uint32_t *conn) {
    const int np = n + 1;
    for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
            c[2 * (j * np + i)] = (double)i / n;
            c[2 * (j * np + i) + 1] = (double)j / n;
        }
    size_t e = 0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            uint32_t n0 = j * np + i, n1 = j * np + i + 1, n2 = (j + 1) * np + i + 1,
                     n3 = (j + 1) * np + i;
            conn[3 * e] = n0; conn[3 * e + 1] = n1; conn[3 * e + 2] = n3; ++e;
            conn[3 * e] = n1; conn[3 * e + 1] = n2; conn[3 * e + 2] = n3; ++e;
        }
}

/* ElementTransformation::from_simplex_nodes (mesh/src/transformation.rs:34-64):
 * J columns x_k - x_0; returns det, fills J^-T. */
static double simplex_geom(int dim, const double *coords, const uint32_t *nodes, double J[3][3],
                           double JIT[3][3], double x0[3]) {
    for (int r = 0; r < dim; ++r) x0[r] = coords[dim * nodes[0] + r];
    for (int c = 0; c < dim; ++c)
        for (int r = 0; r < dim; ++r) J[r][c] = coords[dim * nodes[c + 1] + r] - x0[r];
    double det;
    if (dim == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        JIT[0][0] = J[1][1] / det; JIT[0][1] = -J[1][0] / det;
        JIT[1][0] = -J[0][1] / det; JIT[1][1] = J[0][0] / det;
    } else {
        double I[3][3];
        det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
              J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        inv3(J, det, I);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) JIT[r][c] = I[c][r];
    }
    return det;
}

static int simplex_rule(int dim, int order, double *pts, double *wts) {
    return dim == 2 ? fr_tri_rule(order, pts, wts) : fr_tet_rule(order, pts, wts);
}

static int simplex_nl(int dim, int order) {
    if (dim == 2) return (order + 1) * (order + 2) / 2;
    return (order + 1) * (order + 2) * (order + 3) / 6;
}

size_t fr_simplex_assemble_csr(int dim, int order, size_t n_nodes, const double *coords,
                               size_t n_elem, const uint32_t *conn, const uint32_t *edofs,
                               size_t n_dofs, int kind, double coef, const double *coef_qp,
                               int quad_order, uint64_t *row_ptr, uint32_t *col_idx, double *values) {
    (void)n_nodes;
    const int nl = simplex_nl(dim, order), nv = dim + 1;
    double pts[120], wts[40];
    int nq = simplex_rule(dim, quad_order, pts, wts);
    size_t raw = n_elem * nl * nl;
    uint32_t *rr = malloc(sizeof(uint32_t) * raw), *cc = malloc(sizeof(uint32_t) * raw);
    double *vv = calloc(raw, sizeof(double));
    double phi[20], gr[60], gp[60];
    for (size_t e = 0; e < n_elem; ++e) {
        double J[3][3], JIT[3][3], x0[3];
        double det = simplex_geom(dim, coords, conn + e * nv, J, JIT, x0);
        double *Ke = vv + e * nl * nl;
        for (int qi = 0; qi < nq; ++qi) {
            double w = wts[qi] * fabs(det);
            fr_simplex_basis(dim, order, pts + dim * qi, phi, gr);
            for (int l = 0; l < nl; ++l)
                for (int r = 0; r < dim; ++r) {
                    double s = 0.0;
                    for (int k = 0; k < dim; ++k) s += JIT[r][k] * gr[l * dim + k];
                    gp[l * dim + r] = s;
                }
            /* kappa.eval at the point (diffusion.rs:34-55): the caller's value there */
            double wc = w * (coef_qp ? coef_qp[e * nq + qi] : coef);
            for (int i = 0; i < nl; ++i)
                for (int j = 0; j < nl; ++j) {
                    if (kind == FR_DIFFUSION) {
                        double dot = 0.0;
                        for (int k = 0; k < dim; ++k) dot += gp[i * dim + k] * gp[j * dim + k];
                        Ke[i * nl + j] += wc * dot;
                    } else {
                        Ke[i * nl + j] += wc * phi[i] * phi[j];
                    }
                }
        }
        for (int i = 0; i < nl; ++i)
            for (int j = 0; j < nl; ++j) {
                rr[e * nl * nl + i * nl + j] = edofs[e * nl + i];
                cc[e * nl * nl + i * nl + j] = edofs[e * nl + j];
            }
    }
    size_t nnz = coo_to_csr(n_dofs, raw, rr, cc, vv, row_ptr, col_idx, values);
    free(rr); free(cc); free(vv);
    return nnz;
}

void fr_p1_source_rhs(int n, const double *coords, const uint32_t *conn, int quad_order,
                      double *rhs) {
    const size_t E = 2 * (size_t)n * n, N = (size_t)(n + 1) * (n + 1);
    double pts[24], wts[12], phi[3], gr[6];
    int nq = fr_tri_rule(quad_order, pts, wts);
    memset(rhs, 0, N * sizeof(double));
    for (size_t e = 0; e < E; ++e) {
        double J[3][3], JIT[3][3], x0[3];
        double det = simplex_geom(2, coords, conn + 3 * e, J, JIT, x0);
        double fe[3] = {0, 0, 0};
        for (int qi = 0; qi < nq; ++qi) {
            const double *xi = pts + 2 * qi;
            double w = wts[qi] * fabs(det);
            fr_simplex_basis(2, 1, xi, phi, gr);
            double xp[2];
            for (int r = 0; r < 2; ++r) xp[r] = x0[r] + (J[r][0] * xi[0] + J[r][1] * xi[1]);
            double f = 2.0 * FR_PI * FR_PI * sin(FR_PI * xp[0]) * sin(FR_PI * xp[1]);
            for (int i = 0; i < 3; ++i) fe[i] += w * f * phi[i];
        }
        for (int i = 0; i < 3; ++i) rhs[conn[3 * e + i]] += fe[i];
    }
}

size_t fr_unit_square_boundary(int n, uint32_t *bnd) {
    size_t c = 0;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i)
            if (i == 0 || i == n || j == 0 || j == n) bnd[c++] = (uint32_t)(j * (n + 1) + i);
    return c;
}

void fr_csr_dirichlet_row_zeroing(size_t nrows, const uint64_t *row_ptr, const uint32_t *col_idx,
                                  double *values, size_t nb, const uint32_t *bnd,
                                  const double *bval, double *rhs) {
    (void)nrows;
    for (size_t t = 0; t < nb; ++t) {
        uint32_t r = bnd[t];
        for (uint64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) values[k] = 0.0;
        for (uint64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
            if (col_idx[k] == r) { values[k] = 1.0; break; }
        rhs[r] = bval ? bval[t] : 0.0;
    }
}

void fr_simplex_pa_apply(int dim, int order, const double *coords, size_t n_elem,
                         const uint32_t *conn, const uint32_t *edofs, int kind, double coef,
                         const double *coef_qp, int quad_order, const double *x, double *y, int accumulate) {
    const int nl = simplex_nl(dim, order), nv = dim + 1;
    double pts[120], wts[40], phi[20], gr[60], gp[60], xe[20], ye[20];
    int nq = simplex_rule(dim, quad_order, pts, wts);
    if (!accumulate) {
        size_t nd = 0;
        for (size_t i = 0; i < n_elem * nl; ++i) if (edofs[i] + 1 > nd) nd = edofs[i] + 1;
        memset(y, 0, nd * sizeof(double));
    }
    for (size_t e = 0; e < n_elem; ++e) {
        double J[3][3], JIT[3][3], x0[3];
        double det = simplex_geom(dim, coords, conn + e * nv, J, JIT, x0);
        for (int l = 0; l < nl; ++l) { xe[l] = x[edofs[e * nl + l]]; ye[l] = 0.0; }
        for (int qi = 0; qi < nq; ++qi) {
            double w = wts[qi] * fabs(det);
            /* rho / kappa evaluated at the point (partial.rs:116-117, 186-190) */
            const double cq = coef_qp ? coef_qp[e * nq + qi] : coef;
            fr_simplex_basis(dim, order, pts + dim * qi, phi, gr);
            if (kind == FR_MASS) {
                double ux = 0.0;
                for (int l = 0; l < nl; ++l) ux += phi[l] * xe[l];
                for (int i = 0; i < nl; ++i) ye[i] += w * cq * phi[i] * ux;
            } else {
                for (int l = 0; l < nl; ++l)
                    for (int r = 0; r < dim; ++r) {
                        double s = 0.0;
                        for (int k = 0; k < dim; ++k) s += JIT[r][k] * gr[l * dim + k];
                        gp[l * dim + r] = s;
                    }
                double gu[3] = {0, 0, 0};
                for (int r = 0; r < dim; ++r)
                    for (int l = 0; l < nl; ++l) gu[r] += xe[l] * gp[l * dim + r];
                for (int i = 0; i < nl; ++i) {
                    double dot = 0.0;
                    for (int r = 0; r < dim; ++r) dot += gu[r] * gp[i * dim + r];
                    ye[i] += w * cq * dot;
                }
            }
        }
        for (int l = 0; l < nl; ++l) y[edofs[e * nl + l]] += ye[l];
    }
}

/* Physical quadrature points x = x_0 + J xi (transformation.rs:34-64), the
 * points where the reference evaluates its coefficients: xq[(e*nq + q)*dim + r].
 * Returns nq. */
int fr_simplex_qp_coords(int dim, const double *coords, size_t n_elem, const uint32_t *conn, int quad_order,
                         double *xq) {
    double pts[120], wts[40];
    const int nq = simplex_rule(dim, quad_order, pts, wts), nv = dim + 1;
    if (!xq) return nq;
    for (size_t e = 0; e < n_elem; ++e) {
        double J[3][3], JIT[3][3], x0[3];
        simplex_geom(dim, coords, conn + e * nv, J, JIT, x0);
        for (int qi = 0; qi < nq; ++qi)
            for (int r = 0; r < dim; ++r) {
                double acc = 0.0;
                for (int k = 0; k < dim; ++k) acc += J[r][k] * pts[dim * qi + k];
                xq[(e * nq + qi) * dim + r] = x0[r] + acc;
            }
    }
    return nq;
}

void fr_simplex_lumped_mass(int dim, int order, const double *coords, size_t n_elem,
                            const uint32_t *conn, const uint32_t *edofs, size_t n_dofs, double rho,
                            const double *rho_qp, int quad_order, double *diag) {
    const int nl = simplex_nl(dim, order), nv = dim + 1;
    double pts[120], wts[40], phi[20], gr[60];
    int nq = simplex_rule(dim, quad_order, pts, wts);
    memset(diag, 0, n_dofs * sizeof(double));
    for (size_t e = 0; e < n_elem; ++e) {
        double J[3][3], JIT[3][3], x0[3];
        double det = simplex_geom(dim, coords, conn + e * nv, J, JIT, x0);
        for (int qi = 0; qi < nq; ++qi) {
            double w = wts[qi] * fabs(det);
            fr_simplex_basis(dim, order, pts + dim * qi, phi, gr);
            double s = 0.0;
            for (int l = 0; l < nl; ++l) s += phi[l];
            const double r = rho_qp ? rho_qp[e * nq + qi] : rho;
            for (int i = 0; i < nl; ++i) diag[edofs[e * nl + i]] += w * r * phi[i] * s;
        }
    }
}

/* Conforming numbering (see header). */
typedef struct { uint32_t a, b; size_t e; int slot; } edge_rec;
static int edge_cmp(const void *x, const void *y) {
    const edge_rec *p = x, *q = y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return 0;
}
typedef struct { uint32_t v[3]; size_t e; int slot; } face_rec;
static int face_cmp(const void *x, const void *y) {
    const face_rec *p = x, *q = y;
    for (int i = 0; i < 3; ++i)
        if (p->v[i] != q->v[i]) return p->v[i] < q->v[i] ? -1 : 1;
    return 0;
}

size_t fr_simplex_dofs(int dim, int order, size_t n_nodes, size_t n_elem, const uint32_t *conn,
                       uint32_t *ed) {
    const int nv = dim + 1, nl = simplex_nl(dim, order);
    static const int tri_e[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    static const int tet_e[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    static const int tet_f[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    const int ne = dim == 2 ? 3 : 6;
    const int (*EL)[2] = dim == 2 ? tri_e : tet_e;
    for (size_t e = 0; e < n_elem; ++e)
        for (int a = 0; a < nv; ++a) ed[e * nl + a] = conn[e * nv + a];
    size_t next = n_nodes;
    if (order == 1) return next;
    const int per_edge = order - 1;
    edge_rec *er = malloc(sizeof(edge_rec) * n_elem * ne);
    size_t c = 0;
    for (size_t e = 0; e < n_elem; ++e)
        for (int k = 0; k < ne; ++k) {
            uint32_t a = conn[e * nv + EL[k][0]], b = conn[e * nv + EL[k][1]];
            er[c].a = a < b ? a : b; er[c].b = a < b ? b : a; er[c].e = e; er[c].slot = k; ++c;
        }
    qsort(er, c, sizeof(edge_rec), edge_cmp);
    for (size_t t = 0; t < c; ++t) {
        if (t == 0 || edge_cmp(&er[t], &er[t - 1]) != 0) next += per_edge;
        size_t base = next - per_edge;
        size_t e = er[t].e;
        int k = er[t].slot;
        uint32_t first = conn[e * nv + EL[k][0]];
        int forward = first == er[t].a; /* local edge starts at the lower global vertex */
        for (int m = 0; m < per_edge; ++m) {
            int local = nv + k * per_edge + m;
            ed[e * nl + local] = (uint32_t)(base + (forward ? m : per_edge - 1 - m));
        }
    }
    free(er);
    if (order == 3) {
        if (dim == 2) {
            for (size_t e = 0; e < n_elem; ++e) ed[e * nl + 9] = (uint32_t)(next + e);
            next += n_elem;
        } else {
            face_rec *fr = malloc(sizeof(face_rec) * n_elem * 4);
            c = 0;
            for (size_t e = 0; e < n_elem; ++e)
                for (int f = 0; f < 4; ++f) {
                    uint32_t v[3] = {conn[e * 4 + tet_f[f][0]], conn[e * 4 + tet_f[f][1]],
                                     conn[e * 4 + tet_f[f][2]]};
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2 - i; ++j)
                            if (v[j] > v[j + 1]) { uint32_t tt = v[j]; v[j] = v[j + 1]; v[j + 1] = tt; }
                    memcpy(fr[c].v, v, sizeof v); fr[c].e = e; fr[c].slot = f; ++c;
                }
            qsort(fr, c, sizeof(face_rec), face_cmp);
            for (size_t t = 0; t < c; ++t) {
                if (t == 0 || face_cmp(&fr[t], &fr[t - 1]) != 0) ++next;
                ed[fr[t].e * nl + 16 + fr[t].slot] = (uint32_t)(next - 1);
            }
            free(fr);
        }
    }
    return next;
}

/* ======================================================================= */
/* Solvers                                                                  */
/* ======================================================================= */
static double dot(size_t n, const double *a, const double *b) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

void fr_hex_apply_cb(void *ctx, const double *x, double *y) {
    fr_hex_ctx *c = ctx;
    fr_hex_apply_sumfac(c->d, c->D, c->M, x, y, 0, c->nthreads);
}

void fr_csr_apply_cb(void *ctx, const double *x, double *y) { fr_csr_spmv((const fr_csr *)ctx, x, y); }

/* solver/src/lib.rs:198-278 */
int fr_solve_cg_operator(size_t n, fr_apply_fn apply, void *ctx, const double *b, double *x,
                         const fr_solver_config *cfg, fr_solve_result *res) {
    double *r = malloc(n * sizeof(double)), *p = malloc(n * sizeof(double)),
           *ap = malloc(n * sizeof(double));
    apply(ctx, x, ap);
    for (size_t i = 0; i < n; ++i) { r[i] = b[i] - ap[i]; p[i] = r[i]; }
    double norm_b = sqrt(dot(n, b, b));
    double tol = fmax(cfg->atol, cfg->rtol * fmax(norm_b, 1e-32));
    double rs_old = dot(n, r, r), res_norm = sqrt(rs_old);
    int st = FR_OK;
    res->converged = 0; res->iterations = 0; res->final_residual = res_norm;
    if (res_norm <= tol) { res->converged = 1; goto done; }
    for (uint64_t it = 0; it < cfg->max_iter; ++it) {
        apply(ctx, p, ap);
        double pap = dot(n, p, ap);
        if (fabs(pap) < 1e-32) { st = FR_BREAKDOWN; res->iterations = it; goto done; }
        double alpha = rs_old / pap;
        for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * ap[i]; }
        double rs_new = dot(n, r, r);
        res_norm = sqrt(rs_new);
        res->final_residual = res_norm;
        if (res_norm <= tol) { res->converged = 1; res->iterations = it + 1; goto done; }
        double beta = rs_new / rs_old;
        for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        rs_old = rs_new;
    }
    res->iterations = cfg->max_iter;
    st = FR_NOT_CONVERGED;
done:
    res->status = st;
    free(r); free(p); free(ap);
    return st;
}

/* PCG core shared by Jacobi and Chebyshev preconditioning
 * (parallel/src/par_solver.rs:262-346). */
typedef void (*prec_fn)(void *pctx, const double *r, double *z);
typedef struct { size_t n; const double *inv; } jac_ctx;
static void jac_apply(void *c, const double *r, double *z) {
    jac_ctx *j = c;
    for (size_t i = 0; i < j->n; ++i) z[i] = j->inv[i] * r[i];
}
static int pcg_core(size_t n, fr_apply_fn apply, void *ctx, prec_fn prec, void *pctx,
                    const double *b, double *x, const fr_solver_config *cfg, fr_solve_result *res) {
    double *r = malloc(n * sizeof(double)), *z = malloc(n * sizeof(double)),
           *p = malloc(n * sizeof(double)), *ap = malloc(n * sizeof(double));
    apply(ctx, x, ap);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - ap[i];
    prec(pctx, r, z);
    memcpy(p, z, n * sizeof(double));
    double rz = dot(n, r, z);
    double b_norm = sqrt(dot(n, b, b));
    res->status = FR_OK;
    if (b_norm < 1e-30) {
        res->converged = 1; res->iterations = 0; res->final_residual = 0.0;
        goto done;
    }
    for (uint64_t it = 0; it < cfg->max_iter; ++it) {
        apply(ctx, p, ap);
        double pap = dot(n, p, ap);
        if (fabs(pap) < 1e-30) break;
        double alpha = rz / pap;
        for (size_t i = 0; i < n; ++i) x[i] += alpha * p[i];
        for (size_t i = 0; i < n; ++i) r[i] += -alpha * ap[i];
        double rr = dot(n, r, r);
        double res_norm = sqrt(rr) / b_norm;
        if (res_norm < cfg->rtol || sqrt(rr) < cfg->atol) {
            res->converged = 1; res->iterations = it + 1; res->final_residual = res_norm;
            goto done;
        }
        prec(pctx, r, z);
        double rz_new = dot(n, r, z);
        double beta = rz_new / rz;
        for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz = rz_new;
    }
    res->converged = 0; res->iterations = cfg->max_iter;
    res->final_residual = sqrt(dot(n, r, r)) / b_norm;
    res->status = FR_NOT_CONVERGED;
done:
    free(r); free(z); free(p); free(ap);
    return res->status;
}

int fr_pcg_jacobi(size_t n, fr_apply_fn apply, void *ctx, const double *diag, const double *b,
                  double *x, const fr_solver_config *cfg, fr_solve_result *res) {
    double *inv = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) inv[i] = fabs(diag[i]) > 1e-30 ? 1.0 / diag[i] : 1.0;
    jac_ctx j = {n, inv};
    int st = pcg_core(n, apply, ctx, jac_apply, &j, b, x, cfg, res);
    free(inv);
    return st;
}

/* amg/src/lib.rs:53-103 */
void fr_chebyshev_smooth_ref(const fr_csr *a, double *x, const double *b, double lmin, double lmax,
                             int n_iter) {
    const size_t n = a->nrows;
    double alpha = (lmax + lmin) / 2.0, delta = (lmax - lmin) / 2.0;
    double *dinv = malloc(n * sizeof(double)), *r = calloc(n, sizeof(double)),
           *d = calloc(n, sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        double di = 0.0;
        for (uint64_t k = a->row_ptr[i]; k < a->row_ptr[i + 1]; ++k)
            if (a->col_idx[k] == i) { di = a->values[k]; break; }
        dinv[i] = fabs(di) > 1e-14 ? 1.0 / di : 1.0;
    }
    double two_over_alpha = 2.0 / alpha;
    double q = pow(delta / (2.0 * alpha), 2);
    double rho_prev = 0.0;
    for (int k = 0; k < n_iter; ++k) {
        fr_csr_spmv(a, x, r);
        for (size_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        if (k == 0) {
            for (size_t i = 0; i < n; ++i) d[i] = two_over_alpha * dinv[i] * r[i];
            rho_prev = 1.0;
        } else {
            double rho = 1.0 / (1.0 - q * rho_prev);
            for (size_t i = 0; i < n; ++i)
                d[i] = rho * (two_over_alpha * dinv[i] * r[i] + (rho - 1.0) * d[i]);
            rho_prev = rho;
        }
        for (size_t i = 0; i < n; ++i) x[i] += d[i];
    }
    free(dinv); free(r); free(d);
}

/* amg/src/lib.rs:108-123 */
double fr_estimate_spectral_radius(const fr_csr *a, int n_iter) {
    const size_t n = a->nrows;
    double *v = malloc(n * sizeof(double)), *w = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) v[i] = 1.0 / sqrt((double)n);
    double rho = 1.0;
    for (int it = 0; it < n_iter; ++it) {
        fr_csr_spmv(a, v, w);
        rho = sqrt(dot(n, w, w));
        if (rho > 1e-14) {
            double inv = 1.0 / rho;
            for (size_t i = 0; i < n; ++i) v[i] = w[i] * inv;
        }
    }
    free(v); free(w);
    return rho;
}

double fr_estimate_lmax_jacobi(size_t n, fr_apply_fn apply, void *ctx, const double *diag,
                               int n_iter) {
    double *v = malloc(n * sizeof(double)), *w = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) v[i] = 1.0 / sqrt((double)n);
    double rho = 1.0;
    for (int it = 0; it < n_iter; ++it) {
        apply(ctx, v, w);
        for (size_t i = 0; i < n; ++i) w[i] = w[i] / diag[i];
        rho = sqrt(dot(n, w, w));
        if (rho > 1e-14) {
            double inv = 1.0 / rho;
            for (size_t i = 0; i < n; ++i) v[i] = w[i] * inv;
        }
    }
    free(v); free(w);
    return rho;
}

void fr_chebyshev_jacobi_apply(size_t n, fr_apply_fn apply, void *ctx, const double *diag,
                               int degree, double lmin, double lmax, const double *r, double *z) {
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho_prev = 1.0 / sigma;
    double *d = malloc(n * sizeof(double)), *res = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) { d[i] = r[i] / diag[i] / theta; z[i] = d[i]; }
    for (int k = 1; k < degree; ++k) {
        apply(ctx, z, res);
        double rho = 1.0 / (2.0 * sigma - rho_prev);
        double c1 = rho * rho_prev, c2 = 2.0 * rho / delta;
        for (size_t i = 0; i < n; ++i) {
            d[i] = c1 * d[i] + c2 * ((r[i] - res[i]) / diag[i]);
            z[i] += d[i];
        }
        rho_prev = rho;
    }
    free(d); free(res);
}

typedef struct { size_t n; fr_apply_fn apply; void *ctx; const double *diag; int degree; double lmin, lmax; } cheb_ctx;
static void cheb_prec(void *c, const double *r, double *z) {
    cheb_ctx *k = c;
    fr_chebyshev_jacobi_apply(k->n, k->apply, k->ctx, k->diag, k->degree, k->lmin, k->lmax, r, z);
}

int fr_pcg_chebyshev(size_t n, fr_apply_fn apply, void *ctx, const double *diag, int degree,
                     double lmin, double lmax, const double *b, double *x,
                     const fr_solver_config *cfg, fr_solve_result *res) {
    cheb_ctx c = {n, apply, ctx, diag, degree, lmin, lmax};
    return pcg_core(n, apply, ctx, cheb_prec, &c, b, x, cfg, res);
}

/* ======================================================================= */
/* conv3d (learned preconditioner oracle)                                   */
/* ======================================================================= */
void fr_conv3d_fwd(int ci, int co, int nz, int ny, int nx, const float *x, const float *w,
                   const float *b, float *y) {
    const size_t V = (size_t)nz * ny * nx;
#pragma omp parallel for collapse(2) schedule(static)
    for (int o = 0; o < co; ++o)
        for (int z = 0; z < nz; ++z)
            for (int yy = 0; yy < ny; ++yy)
                for (int xx = 0; xx < nx; ++xx) {
                    float s = b ? b[o] : 0.0f;
                    for (int c = 0; c < ci; ++c)
                        for (int kz = 0; kz < 3; ++kz) {
                            int zz = z + kz - 1;
                            if (zz < 0 || zz >= nz) continue;
                            for (int ky = 0; ky < 3; ++ky) {
                                int y2 = yy + ky - 1;
                                if (y2 < 0 || y2 >= ny) continue;
                                for (int kx = 0; kx < 3; ++kx) {
                                    int x2 = xx + kx - 1;
                                    if (x2 < 0 || x2 >= nx) continue;
                                    s += w[(((size_t)o * ci + c) * 3 + kz) * 9 + ky * 3 + kx] *
                                         x[(size_t)c * V + ((size_t)zz * ny + y2) * nx + x2];
                                }
                            }
                        }
                    y[(size_t)o * V + ((size_t)z * ny + yy) * nx + xx] = s;
                }
}

void fr_conv3d_bwd(int ci, int co, int nz, int ny, int nx, const float *x, const float *w,
                   const float *dy, float *dx, float *dw, float *db) {
    const size_t V = (size_t)nz * ny * nx;
    for (int o = 0; o < co; ++o) {
        double s = 0.0;
        for (size_t v = 0; v < V; ++v) s += dy[(size_t)o * V + v];
        db[o] = (float)s;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int o = 0; o < co; ++o)
        for (int c = 0; c < ci; ++c)
            for (int t = 0; t < 27; ++t) {
                int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
                double s = 0.0;
                for (int z = 0; z < nz; ++z) {
                    int zz = z + kz - 1;
                    if (zz < 0 || zz >= nz) continue;
                    for (int yy = 0; yy < ny; ++yy) {
                        int y2 = yy + ky - 1;
                        if (y2 < 0 || y2 >= ny) continue;
                        for (int xx = 0; xx < nx; ++xx) {
                            int x2 = xx + kx - 1;
                            if (x2 < 0 || x2 >= nx) continue;
                            s += (double)dy[(size_t)o * V + ((size_t)z * ny + yy) * nx + xx] *
                                 x[(size_t)c * V + ((size_t)zz * ny + y2) * nx + x2];
                        }
                    }
                }
                dw[((size_t)o * ci + c) * 27 + t] = (float)s;
            }
    if (dx) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int c = 0; c < ci; ++c)
            for (int z = 0; z < nz; ++z)
                for (int yy = 0; yy < ny; ++yy)
                    for (int xx = 0; xx < nx; ++xx) {
                        float s = 0.0f;
                        for (int o = 0; o < co; ++o)
                            for (int t = 0; t < 27; ++t) {
                                int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
                                int zz = z - kz + 1, y2 = yy - ky + 1, x2 = xx - kx + 1;
                                if (zz < 0 || zz >= nz || y2 < 0 || y2 >= ny || x2 < 0 || x2 >= nx) continue;
                                s += w[((size_t)o * ci + c) * 27 + t] *
                                     dy[(size_t)o * V + ((size_t)zz * ny + y2) * nx + x2];
                            }
                        dx[(size_t)c * V + ((size_t)z * ny + yy) * nx + xx] = s;
                    }
    }
}
