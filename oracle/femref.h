/*
 * femref — CPU oracle for the fem-rs solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the reference
 * CPU algorithms of javagg/fem-rs (Rust) for the partial-assembly (PA)
 * operator apply, CSR SpMV, CG / PCG-Jacobi and the Chebyshev smoother, plus
 * the new hex / mixed-precision pieces that the reference does not have.  It
 * is linked only by tests/, by __graft_entry__.smoke() as a checker and by
 * bench.py's cpu_baseline leg.  The product library (libfemgpu.so) never
 * links or calls it.
 *
 * Every function cites the reference file:line it follows.  Functions with
 * no reference counterpart say so; each of them is pinned by closed forms,
 * invariants or brute force in tests/test_oracle_*.py.
 *
 * Floating point: compiled with -ffp-contract=off so that products and sums
 * round exactly as the Rust reference (which never contracts to FMA) does.
 */
#ifndef FEMREF_H
#define FEMREF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- RNG ---- */
/* Philox4x32-10 (Salmon et al., SC'11).  No reference counterpart. */
void fr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Uniform double in [0,1) from (seed, stream, index); 53 random bits. */
double fr_uniform(uint64_t seed, uint32_t stream, uint64_t index);

/* --------------------------------------------------------- quadrature ---- */
/* Gauss-Legendre on [-1,1], n = 1..6.  quadrature.rs:12-31 (n <= 4); n = 5 from
 * the closed form.  Returns 0 on success, -1 for unsupported n. */
int fr_gauss_legendre(int n, double *x, double *w);
/* Triangle rule, reference triangle (0,0),(1,0),(0,1): quadrature.rs:169-246.
 * Returns the number of points (max 12); pts[2*i], wts[i]. */
int fr_tri_rule(int order, double *pts, double *wts);
/* Tetrahedron rule, quadrature.rs:250-330 (Grundmann-Moller for order > 2).
 * Returns the number of points (max 20); pts[3*i], wts[i]. */
int fr_tet_rule(int order, double *pts, double *wts);

/* ------------------------------------------------------------- bases ---- */
/* 1-D Lagrange basis on equispaced nodes -1 + 2i/p, i = 0..p, evaluated at
 * xq[0..nq): B[q*(p+1)+i] = l_i(xq[q]), G = l_i'.  Q2 nodes agree with QuadQ2
 * (lagrange/quad.rs:58-140). */
void fr_lagrange_1d(int p, int nq, const double *xq, double *B, double *G);
/* Simplex P1..P3 bases (lagrange/tri.rs, lagrange/tet.rs).  phi[n], grad[n*dim]
 * row-major as ReferenceElement::eval_grad_basis (reference.rs:31-59).
 * Returns the number of local DOFs. */
int fr_simplex_basis(int dim, int order, const double *xi, double *phi, double *grad);

/* --------------------------------------------------------- hex meshes ---- */
/* Structured (nx+1)(ny+1)(nz+1) vertex grid on [0,1]^3; interior vertices are
 * moved by amp*h*(2u-1) per component, u = fr_uniform(seed, 1+c, vertex).
 * coords[3*v+c].  No reference counterpart (SimplexMesh has no hex generator). */
void fr_hex_vertices(int nx, int ny, int nz, double amp, uint64_t seed, double *coords);

/* Coefficient descriptor shared by the hex routines. */
enum { FR_COEF_CONST = 0, FR_COEF_SINPROD = 1, FR_COEF_GRF = 2 };
typedef struct {
    int kind;          /* FR_COEF_* */
    double value;      /* CONST: the constant */
    /* GRF: kappa = clamp(exp(g), lo, hi), g = sigma*sqrt(2/M) sum cos(w.x+b) */
    int n_features;
    double sigma, length, lo, hi;
    uint64_t seed;
} fr_coef;
/* Random-Fourier-feature parameters: w[3*m], b[m] (M = n_features). */
void fr_grf_features(const fr_coef *c, double *w, double *b);
double fr_coef_eval(const fr_coef *c, const double *wfeat, const double *bfeat, const double *x);

/* Hex operator description (structured lexicographic DOF numbering:
 * dof(gx,gy,gz) = gx + Nx*(gy + Ny*gz), Nd = n*p+1). */
enum { FR_MASS = 1, FR_DIFFUSION = 2, FR_DIFF_MASS = 3 };
typedef struct {
    int p, q;          /* order, Gauss points per direction */
    int nx, ny, nz;
    int kind;          /* FR_MASS | FR_DIFFUSION */
    double rho;        /* mass coefficient (constant) */
    fr_coef kappa;     /* diffusion coefficient */
    int dirichlet;     /* 1: all boundary DOFs essential (y[b] = x[b], x[b] masked) */
} fr_hex_desc;

size_t fr_hex_ndofs(const fr_hex_desc *d);
/* Quadrature data: D[e][qp][6] (xx,xy,xz,yy,yz,zz) = w*kappa*|J|*J^-1 J^-T and
 * M[e][qp] = w*rho*|J|.  qp = qx + q*(qy + q*qz).  Either may be NULL. */
void fr_hex_qdata(const fr_hex_desc *d, const double *coords, double *D, double *M);
/* y (+)= A x by the reference element loop generalised to hex
 * (partial.rs:151-208; geometry assembler.rs:96-123): full tensor basis at
 * every point, physical gradients J^-T grad_ref, no sum factorisation. */
void fr_hex_apply_refloop(const fr_hex_desc *d, const double *coords, const double *x,
                          double *y, int accumulate);
/* fr_hex_apply_refloop on nthreads threads over the 8 element parity colours
 * (the threaded CPU baseline of the reference algorithm). */
void fr_hex_apply_refloop_mt(const fr_hex_desc *d, const double *coords, const double *x, double *y,
                             int accumulate, int nthreads);
/* y (+)= A x by sum factorisation on stored qdata (no reference counterpart;
 * the optimised CPU path, OpenMP over element colours). */
void fr_hex_apply_sumfac(const fr_hex_desc *d, const double *D, const double *M,
                         const double *x, double *y, int accumulate, int nthreads);
/* Same in float, with float qdata (FP32 inner-solve oracle). */
void fr_hex_apply_sumfac_f32(const fr_hex_desc *d, const float *D, const float *M,
                             const float *x, float *y, int accumulate, int nthreads);
/* Diagonal of the hex operator (brute force over element DOFs). */
void fr_hex_diagonal(const fr_hex_desc *d, const double *D, const double *M, double *diag);
/* Assembled CSR of the hex operator (brute force element matrices, COO summed
 * per (row,col) in element order).  row_ptr must hold N+1; returns nnz; call
 * with col_idx == NULL to get nnz only.  Dirichlet rows: row zeroing + diag 1
 * and the matching columns zeroed (the matrix-free BC pattern). */
size_t fr_hex_assemble_csr(const fr_hex_desc *d, const double *coords, uint64_t *row_ptr,
                           uint32_t *col_idx, double *values);
/* b_i = int phi_i (the load of f = 1), 0 on essential DOFs. */
void fr_hex_rhs_unit(const fr_hex_desc *d, const double *coords, double *b);

/* --------------------------------------------------------------- CSR ---- */
typedef struct {
    size_t nrows, ncols;
    const uint64_t *row_ptr;
    const uint32_t *col_idx;
    const double *values;
} fr_csr;
/* y = A x with the 4-way blocked row sum of csr.rs:38-64. */
void fr_csr_spmv(const fr_csr *a, const double *x, double *y);
/* y = alpha*(A x) + beta*y (csr.rs:67-79, 547-569). */
void fr_csr_spmv_add(const fr_csr *a, double alpha, const double *x, double beta, double *y);
/* f32 values: sequential row sum (generic path csr.rs:195-214). */
void fr_csr_spmv_f32(size_t nrows, const uint64_t *row_ptr, const uint32_t *col_idx,
                     const float *values, const float *x, float *y);
/* d_i = A_ii by linear row search, 0 if absent (csr.rs:110-120, 246-252). */
void fr_csr_diagonal(const fr_csr *a, double *d);

/* ------------------------------------------------- P1 Poisson 2-D ---- */
/* SimplexMesh::unit_square_tri(n) (simplex.rs:383-439): coords[2*(j*(n+1)+i)],
 * conn[3*e], squares split [n0,n1,n3],[n1,n2,n3]. */
void fr_unit_square_tri(int n, double *coords, uint32_t *conn);
/* Assembler::assemble_bilinear(DiffusionIntegrator{kappa}) / MassIntegrator on a
 * P1..P3 triangle or P1..P3 tet mesh (assembler.rs:147-249, 409-429,
 * diffusion.rs:34-55, mass.rs:28-40) into CSR via COO (coo.rs:76-).
 * kind: FR_MASS / FR_DIFFUSION; coef constant (or per point: coef_qp[e*nq +
 * q] when not NULL).  Returns nnz (col_idx NULL: count only). */
size_t fr_simplex_assemble_csr(int dim, int order, size_t n_nodes, const double *coords,
                               size_t n_elem, const uint32_t *conn, const uint32_t *elem_dofs,
                               size_t n_dofs, int kind, double coef, const double *coef_qp,
                               int quad_order, uint64_t *row_ptr, uint32_t *col_idx, double *values);
/* Config-1 RHS: DomainSourceIntegrator(2 pi^2 sin(pi x) sin(pi y)) on P1
 * (source.rs:44-50, assembler.rs:434-454). */
void fr_p1_source_rhs(int n, const double *coords, const uint32_t *conn, int quad_order,
                      double *rhs);
/* boundary_dofs for unit_square_tri tags 1..4 (constraints.rs:53-214), P1:
 * returns count, fills bnd (sorted ascending). */
size_t fr_unit_square_boundary(int n, uint32_t *bnd);
/* apply_dirichlet row zeroing (constraints.rs:31-42, csr.rs:297-312). */
void fr_csr_dirichlet_row_zeroing(size_t nrows, const uint64_t *row_ptr,
                                  const uint32_t *col_idx, double *values,
                                  size_t nb, const uint32_t *bnd, const double *bval,
                                  double *rhs);

/* --------------------------------------------------- simplex PA apply ---- */
/* PAMassOperator / PADiffusionOperator::apply (partial.rs:85-130, 154-208):
 * y (+)= A x, element loop with affine Jacobian.  coef constant, or with
 * coef_qp (not NULL) the coefficient's value at each quadrature point,
 * coef_qp[e*nq + q] (the reference evaluates its ScalarCoeff there). */
void fr_simplex_pa_apply(int dim, int order, const double *coords, size_t n_elem,
                         const uint32_t *conn, const uint32_t *elem_dofs, int kind,
                         double coef, const double *coef_qp, int quad_order, const double *x, double *y,
                         int accumulate);
/* LumpedMassOperator::assemble (partial.rs:228-262): row-sum lumped diagonal. */
void fr_simplex_lumped_mass(int dim, int order, const double *coords, size_t n_elem,
                            const uint32_t *conn, const uint32_t *elem_dofs, size_t n_dofs,
                            double rho, const double *rho_qp, int quad_order, double *diag);
/* Physical quadrature points of every element (x = x_0 + J xi); returns nq
 * (xq NULL: count only). */
int fr_simplex_qp_coords(int dim, const double *coords, size_t n_elem, const uint32_t *conn, int quad_order,
                         double *xq);
/* Conforming global DOF numbering for P1..P3 simplex meshes: vertices first
 * (= node ids), then one block per edge (sorted vertex pair, DOFs ordered from
 * the lower vertex id), then faces, then interiors.  Returns n_dofs and fills
 * elem_dofs[e*nl + l] in the reference element's local order.  (The reference
 * DofManager numbers differently; the operator is the same up to a
 * permutation of DOFs.) */
size_t fr_simplex_dofs(int dim, int order, size_t n_nodes, size_t n_elem,
                       const uint32_t *conn, uint32_t *elem_dofs);

/* ------------------------------------------------------------ solvers ---- */
typedef void (*fr_apply_fn)(void *ctx, const double *x, double *y);  /* y = A x */
typedef struct { double rtol, atol; uint64_t max_iter; } fr_solver_config;
typedef struct { int converged; uint64_t iterations; double final_residual; int status; } fr_solve_result;
enum { FR_OK = 0, FR_NOT_CONVERGED = 3, FR_BREAKDOWN = 4, FR_DIM = 2 };

/* solve_cg_operator (solver/src/lib.rs:198-278): unpreconditioned CG, absolute
 * final residual, breakdown if |pAp| < 1e-32, error status on max_iter. */
int fr_solve_cg_operator(size_t n, fr_apply_fn apply, void *ctx, const double *b, double *x,
                         const fr_solver_config *cfg, fr_solve_result *res);
/* par_solve_pcg_jacobi on one rank (parallel/src/par_solver.rs:262-346):
 * relative final residual, converged=false on max_iter. */
int fr_pcg_jacobi(size_t n, fr_apply_fn apply, void *ctx, const double *diag, const double *b,
                  double *x, const fr_solver_config *cfg, fr_solve_result *res);
/* chebyshev_smooth exact recurrence (amg/src/lib.rs:53-103). */
void fr_chebyshev_smooth_ref(const fr_csr *a, double *x, const double *b, double lmin,
                             double lmax, int n_iter);
/* estimate_spectral_radius (amg/src/lib.rs:108-123). */
double fr_estimate_spectral_radius(const fr_csr *a, int n_iter);
/* Power iteration on D^-1 A from v = 1/sqrt(n), returns |D^-1 A v| after n_iter
 * (the Jacobi-scaled analogue of the above; no reference counterpart). */
double fr_estimate_lmax_jacobi(size_t n, fr_apply_fn apply, void *ctx, const double *diag,
                               int n_iter);
/* Textbook Chebyshev-Jacobi preconditioner z = p(D^-1 A) D^-1 r of the given
 * degree on [lmin, lmax] (Saad, Iterative Methods 2nd ed., Alg. 12.1 with
 * x0 = 0).  No reference counterpart. */
void fr_chebyshev_jacobi_apply(size_t n, fr_apply_fn apply, void *ctx, const double *diag,
                               int degree, double lmin, double lmax, const double *r,
                               double *z);
/* PCG (par_solver.rs:262-346 semantics) with the Chebyshev-Jacobi
 * preconditioner above in place of D^-1. */
int fr_pcg_chebyshev(size_t n, fr_apply_fn apply, void *ctx, const double *diag, int degree,
                     double lmin, double lmax, const double *b, double *x,
                     const fr_solver_config *cfg, fr_solve_result *res);

/* Apply callbacks for the two operator families. */
typedef struct { const fr_hex_desc *d; const double *D, *M; int nthreads; } fr_hex_ctx;
void fr_hex_apply_cb(void *ctx, const double *x, double *y);
void fr_csr_apply_cb(void *ctx, const double *x, double *y);

/* ------------------------------------------------------------ conv3d ---- */
/* 3x3x3 convolution, zero padding 1, NCDHW with N = 1, float.  y[co] =
 * b[co] + sum_ci sum_taps w[co][ci][kz][ky][kx] * x[ci][z+kz-1][y+ky-1][x+kx-1].
 * No reference counterpart (learned preconditioner). */
void fr_conv3d_fwd(int ci, int co, int nz, int ny, int nx, const float *x, const float *w,
                   const float *b, float *y);
/* Gradients of the above: dx (may be NULL), dw, db (accumulated = overwritten). */
void fr_conv3d_bwd(int ci, int co, int nz, int ny, int nx, const float *x, const float *w,
                   const float *dy, float *dx, float *dw, float *db);

#ifdef __cplusplus
}
#endif
#endif
