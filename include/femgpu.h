/*
 * femgpu.h — C ABI of the MI355X (gfx950) solve path for fem-rs.
 *
 * This is the drop-in boundary: every entry point takes plain pointers and
 * sizes, returns an femgpu_status, and mirrors one reference (fem-rs, Rust)
 * interface, cited as crate-relative file:line.  A Rust `fem-gpu-sys` crate
 * (or any other FFI) binds these symbols directly; the bindings are shown in
 * INTEGRATION.md.
 *
 * Conventions
 *  - Vectors passed to apply/solve calls live in host memory (FEMGPU_HOST;
 *    the library stages them) or in device memory of the context's GPU
 *    (FEMGPU_DEVICE; e.g. from femgpu_malloc).  Their element type is the
 *    handle's dtype.
 *  - Handles copy every array given at creation time.
 *  - A handle serialises its own work on its context's stream; separate
 *    contexts run concurrently.  Calls on one context from several threads
 *    are serialised by the context's mutex (the reference's Send + Sync).
 *  - On error the message for the calling thread is available from
 *    femgpu_last_error().
 */
#ifndef FEMGPU_H
#define FEMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- status ---- */
/* Error mapping to the reference's SolverError (solver/src/lib.rs:47-55):
 * DIM_MISMATCH -> DimensionMismatch{rows,cols,rhs}; NOT_CONVERGED ->
 * ConvergenceFailed{max_iter,residual}; BREAKDOWN -> Linger(String).
 * CsrMatrix::spmv panics on a length mismatch (csr.rs:140-142); the C ABI
 * returns DIM_MISMATCH instead. */
typedef enum {
    FEMGPU_OK = 0,
    FEMGPU_ERR_INVALID_ARG = 1,
    FEMGPU_ERR_DIM_MISMATCH = 2,
    FEMGPU_ERR_NOT_CONVERGED = 3,
    FEMGPU_ERR_BREAKDOWN = 4,
    FEMGPU_ERR_HIP = 5,
    FEMGPU_ERR_UNSUPPORTED = 6,
    FEMGPU_ERR_OUT_OF_MEMORY = 7
} femgpu_status;

typedef enum { FEMGPU_F64 = 0, FEMGPU_F32 = 1 } femgpu_dtype;
typedef enum { FEMGPU_HOST = 0, FEMGPU_DEVICE = 1 } femgpu_mem;

typedef struct femgpu_ctx femgpu_ctx;
typedef struct femgpu_op femgpu_op;

/* Message of the last failing call on this thread ("" if none). */
const char *femgpu_last_error(void);
/* Library version string. */
const char *femgpu_version(void);

/* ------------------------------------------------------------ context ---- */
/* One context per GPU and caller; owns a HIP stream and solver scratch. */
int femgpu_ctx_create(int device, femgpu_ctx **out);
int femgpu_ctx_destroy(femgpu_ctx *ctx);
int femgpu_ctx_synchronize(femgpu_ctx *ctx);
/* Device memory helpers (for FEMGPU_DEVICE vectors). femgpu_malloc returns
 * zero-filled memory.  kind: 0 = H2D, 1 = D2H, 2 = D2D.  Copies are ordered on
 * the context stream and complete on return. */
int femgpu_malloc(femgpu_ctx *ctx, size_t bytes, void **dptr);
int femgpu_free(femgpu_ctx *ctx, void *dptr);
int femgpu_memcpy(femgpu_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);
/* Vector helpers on device vectors of the context's GPU, ordered on its
 * stream (no reference counterpart; Vector<T>::axpy-style BLAS-1 for callers
 * that keep vectors on the device, linalg/src/vector.rs:39-70):
 * y = alpha x + beta y, and dst = src converted between f64 and f32. */
int femgpu_vec_axpby(femgpu_ctx *ctx, femgpu_dtype dtype, uint64_t n, double alpha, const void *x,
                     double beta, void *y);
int femgpu_vec_convert(femgpu_ctx *ctx, uint64_t n, femgpu_dtype src_dtype, const void *src,
                       femgpu_dtype dst_dtype, void *dst);
/* HIP event timing of everything enqueued between start and stop on the
 * context stream (milliseconds). */
int femgpu_timer_start(femgpu_ctx *ctx);
int femgpu_timer_stop(femgpu_ctx *ctx, double *ms);

/* ---------------------------------------------------------- operators ---- */
/* Every operator below is a square n x n linear operator y = A x.
 *
 * femgpu_op_apply: accumulate = 1 is MatFreeOperator::apply (y += A x,
 * assembly/src/partial.rs:45-47); accumulate = 0 is apply_zero / LinearOperator
 * ::apply / CsrMatrix::spmv (y = A x; partial.rs:50-53, backend.rs:44-45,
 * linalg/src/csr.rs:140-164). */
int femgpu_op_apply(femgpu_op *op, const void *x, void *y, int accumulate, femgpu_mem mem);
/* MatFreeOperator::n_dofs (partial.rs:56) / LinearOperator::nrows. */
int femgpu_op_n_dofs(const femgpu_op *op, uint64_t *n);
int femgpu_op_dtype(const femgpu_op *op, femgpu_dtype *dtype);
/* d_i = A_ii (CsrMatrix::diagonal, csr.rs:246-252; for PA operators the
 * assembled diagonal, computed on the device). */
int femgpu_op_diagonal(femgpu_op *op, void *d, femgpu_mem mem);
int femgpu_op_destroy(femgpu_op *op);
/* Profiling (no reference counterpart; the reference only times with
 * std::time::Instant in benchmarks): runs reps overwrite applies on device
 * vectors and
 * returns the mean time per apply and, for operators that record it (hex
 * PA), the mean duration of the main compute kernel from HIP events recorded
 * around it on the context stream (hex PA: the element kernel, without the
 * DOF gather; -1 if not available). */
int femgpu_op_time_apply(femgpu_op *op, const void *x, void *y, int reps, double *ms_apply,
                         double *ms_kernel);
/* Contraction probe (measurement only; no reference counterpart): times one
 * batched sum-factorisation sweep, Out = [B u | G u] over lines u of p + 1
 * nodal values at q Gauss points, on 256 LDS-resident lines per workgroup,
 * repeated reps times per launch, with engine 0 (vector unit, multiply-add
 * chains), 1 (vector unit, even-odd form) or 2 (matrix cores, MFMA 16x16x4).
 * in: nblocks * 256 * (p + 1) device values of dtype; out: nblocks * 256 * 2q.
 * (p, q) = (3, 4) or (4, 5).  ms = mean time per launch over launches. */
int femgpu_sumfac_probe(femgpu_ctx *ctx, int p, int q, femgpu_dtype dtype, int engine, int nblocks, int reps,
                        const void *in, void *out, int launches, double *ms);

/* An operator whose apply is the caller's function: the `apply: F` closure of
 * solve_cg_operator<F: Fn(&[f64], &mut [f64])> (solver/src/lib.rs:198-208),
 * for matrix-free operators outside this library (its caller in the
 * reference: examples/src/maxwell.rs:173-200).  apply(user, x, y, n) must set
 * y = A x (overwrite) and return 0; a nonzero return aborts the solve or
 * apply with INVALID_ARG.  mem = FEMGPU_HOST: x and y are host buffers of n
 * values of dtype owned by the operator (the library copies the device vector
 * out and the result back; it synchronises its stream first).  mem =
 * FEMGPU_DEVICE: x and y are device pointers on the context's GPU; work the
 * callback enqueues on the context's stream (other femgpu calls on the same
 * context) is ordered, work on other streams must be complete when it
 * returns.  diag (optional, host, n values, copied) is the operator's
 * diagonal, which Jacobi and Chebyshev-Jacobi need; without it those
 * solvers return UNSUPPORTED.  femgpu_op_apply with accumulate = 1 adds the
 * result to y. */
typedef int (*femgpu_apply_fn)(void *user, const void *x, void *y, uint64_t n);
int femgpu_op_create_callback(femgpu_ctx *ctx, uint64_t n, femgpu_dtype dtype, femgpu_mem mem,
                              femgpu_apply_fn apply, void *user, const void *diag, femgpu_op **out);

/* CsrMatrix<T> (linalg/src/csr.rs:81-93): row_ptr u64[nrows+1], col_idx
 * u32[nnz], values T[nnz] (T = f64 or f32 per dtype), exactly the reference
 * layout.  Arrays are host pointers and are copied. */
int femgpu_csr_create(femgpu_ctx *ctx, uint64_t nrows, uint64_t ncols, const uint64_t *row_ptr,
                      const uint32_t *col_idx, const void *values, femgpu_dtype dtype,
                      femgpu_op **out);
/* CsrMatrix::spmv_add: y = alpha*(A x) + beta*y (csr.rs:167-193). */
int femgpu_csr_spmv_add(femgpu_op *csr, double alpha, const void *x, double beta, void *y,
                        femgpu_mem mem);
/* apply_dirichlet (space/src/constraints.rs:31-42) = apply_dirichlet_row_zeroing
 * (csr.rs:297-312) for each (dofs[i], values[i]) in order: zero the row, set
 * its diagonal entry to 1 (if stored), rhs[dof] = value.  Repeated DOFs: the
 * last value wins, as in the reference's loop.  rhs (n entries of the matrix
 * dtype) may be NULL.  A DOF >= nrows returns DIM_MISMATCH (the reference
 * panics).  Invalidates the cached diagonal. */
int femgpu_csr_apply_dirichlet(femgpu_op *csr, uint64_t n_constrained, const uint32_t *dofs,
                               const double *values, void *rhs, femgpu_mem mem);
/* Copy of the nnz matrix values (CsrMatrix::values). */
int femgpu_csr_get_values(femgpu_op *csr, void *values, femgpu_mem mem);
/* The pattern of a CSR operator, CsrMatrix::row_ptr / col_idx (linalg/src/csr.rs:81-93):
 * nnz always; row_ptr u64[nrows + 1] and col_idx u32[nnz] when not NULL. */
int femgpu_csr_get_pattern(femgpu_op *csr, uint64_t *nnz, uint64_t *row_ptr, uint32_t *col_idx,
                           femgpu_mem mem);

/* Coefficient for the hex PA operators (CoeffCtx/ScalarCoeff,
 * assembly/src/coefficient.rs:67-132, evaluated once at set-up into the
 * quadrature data instead of per apply). */
typedef enum {
    FEMGPU_COEF_CONST = 0,    /* value */
    FEMGPU_COEF_SINPROD = 1,  /* 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z) */
    FEMGPU_COEF_GRF = 2,      /* clamp(exp(g), lo, hi); g a Gaussian random field
                                 (random Fourier features, squared-exponential) */
    FEMGPU_COEF_QP_ARRAY = 3  /* values given per element and quadrature point */
} femgpu_coef_kind;

typedef struct {
    int kind;             /* femgpu_coef_kind */
    double value;         /* CONST */
    int n_features;       /* GRF */
    double sigma, length, lo, hi;
    uint64_t seed;
    const double *qp_values; /* QP_ARRAY: host array [E][q^3], qp = qx + q(qy + q qz),
                                e = ex + nx(ey + ny ez) */
} femgpu_coef;

typedef enum { FEMGPU_MASS = 1, FEMGPU_DIFFUSION = 2, FEMGPU_DIFF_MASS = 3 } femgpu_form;

/* Structured hexahedral Q_p mesh of [0,1]^3 with nx*ny*nz elements and
 * trilinear (HexQ1, element/src/lagrange/hex.rs:8-36) geometry; DOFs are the
 * equispaced tensor nodes in lexicographic order, dof(gx,gy,gz) = gx + Nx*(gy
 * + Ny*gz) with Nd = n*p + 1.  q Gauss-Legendre points per direction. */
typedef struct {
    int p, q;
    int nx, ny, nz;
    int form;           /* femgpu_form */
    femgpu_dtype dtype; /* arithmetic and storage precision of the operator */
    int dirichlet;      /* 1: every boundary DOF is essential: x is zeroed there
                           and y = x there (examples/src/maxwell.rs:142-197) */
    double rho;         /* mass (reaction) coefficient, constant */
    femgpu_coef kappa;  /* diffusion coefficient */
} femgpu_hex_desc;

/* PADiffusionOperator::new / PAMassOperator::new (partial.rs:65-79, 138-148)
 * for hex Q_p: the reference supports simplices only (partial.rs:278-288);
 * this is the tensor-product, sum-factorised form.  vertex_coords: host
 * f64[(nx+1)(ny+1)(nz+1)][3], vertex v = i + (nx+1)(j + (ny+1)k). */
int femgpu_pa_create_hex(femgpu_ctx *ctx, const femgpu_hex_desc *desc,
                         const double *vertex_coords, femgpu_op **out);
/* Structured vertex grid with interior vertices moved by amp*h*(2u-1), u
 * uniform from Philox4x32-10 keyed by seed (the perturbed-mesh generator). */
int femgpu_hex_vertices(int nx, int ny, int nz, double amp, uint64_t seed, double *coords);
/* b_i = integral of phi_i (load of f = 1), 0 on essential DOFs. */
int femgpu_pa_hex_rhs_unit(femgpu_op *op, void *b, femgpu_mem mem);
/* Seeded uniform [-1,1] nodal field (0 on essential DOFs when dirichlet). */
int femgpu_pa_hex_rhs_random(femgpu_op *op, uint64_t seed, void *b, femgpu_mem mem);
/* Values of quadrature data stored per Gauss point: 1 (w kappa) with the
 * geometry formed on the fly (the default for the diffusion forms on meshes
 * that are not axis-aligned boxes); 1 on an axis-aligned box mesh (vertex x
 * depends on i only, y on j, z on k; then w kappa |J| is stored with per-axis
 * factors 1/J_dd^2), plus 1 for the mass term; 6 (+ 1) for the stored general
 * form, which FEMGPU_HEX_QDATA=general at creation selects on every mesh.  No
 * counterpart in the reference (its PA is simplex-only,
 * assembly/src/partial.rs). */
int femgpu_pa_hex_qdata_per_point(const femgpu_op *op, int *values);
/* Quadrature-data mode of a hex operator: 0 = general (D = w kappa |J|
 * J^-1 J^-T stored per point), 1 = Cartesian (axis-aligned box mesh), 2 =
 * geometry on the fly (w kappa stored per point; the element kernel forms J
 * from the element's 8 vertices at every point, and the mass weight rho w |J|
 * with the constant rho).  Diffusion forms take 1 on axis-aligned box meshes
 * and 2 elsewhere; FEMGPU_HEX_QDATA=general at creation selects 0; the
 * mass-only form stores w rho |J| (0).  A converted copy keeps its source's
 * mode.  No counterpart in the reference. */
int femgpu_pa_hex_qdata_mode(const femgpu_op *op, int *mode);
/* Element kernel of a hex operator: 0 = one workgroup per group of elements
 * (LDS hand-offs between the contraction phases), 1 = one thread per element
 * (Q2 diffusion, 3^3 points, Cartesian qdata), 2 = one quad of lanes per
 * element (Q3 diffusion, 4^3 points, Cartesian qdata); FEMGPU_HEX_KERNEL=lds
 * at creation selects 0.  No counterpart in the reference. */
int femgpu_pa_hex_element_kernel(const femgpu_op *op, int *kind);
/* Same-mesh copy of a hex operator in another precision, reusing nothing but
 * the description (for the FP32 inner / FP64 outer solver). */
int femgpu_pa_hex_convert(femgpu_op *op, femgpu_dtype dtype, femgpu_op **out);

/* Simplex PA: PAMassOperator / PADiffusionOperator::apply
 * (partial.rs:85-130, 154-208) on Tri3/Tet4 meshes, P1..P3 (TriP1..P3,
 * TetP1..P3).  coords f64[n_nodes*dim]; elem_nodes u32[n_elem*(dim+1)];
 * elem_dofs u32[n_elem*n_local] in the reference element's local order
 * (FESpace::element_dofs, space/src/fe_space.rs:28-65); coef constant;
 * quad_order as in tri_rule/tet_rule (element/src/quadrature.rs:169-330). */
int femgpu_pa_create_simplex(femgpu_ctx *ctx, int dim, int order, uint64_t n_nodes,
                             const double *coords, uint64_t n_elem, const uint32_t *elem_nodes,
                             const uint32_t *elem_dofs, uint64_t n_dofs, int form, double coef,
                             int quad_order, femgpu_op **out);
/* The coefficient-generic forms: PADiffusionOperator<S, K: ScalarCoeff> and
 * PAMassOperator<S, C> evaluate kappa.eval(&CoeffCtx) / rho at every
 * quadrature point of every element (assembly/src/partial.rs:116-117,
 * 137-147, 186-190; coefficient.rs:67-132).  The caller evaluates its
 * coefficient at the physical points femgpu_simplex_qp_coords returns and
 * passes the values: coef_qp[e * nq + q] (f64, copied), q in tri_rule /
 * tet_rule order for quad_order.  The operator multiplies each point's weight
 * by its value. */
int femgpu_pa_create_simplex_qp(femgpu_ctx *ctx, int dim, int order, uint64_t n_nodes,
                                const double *coords, uint64_t n_elem, const uint32_t *elem_nodes,
                                const uint32_t *elem_dofs, uint64_t n_dofs, int form,
                                const double *coef_qp, int quad_order, femgpu_op **out);
/* Physical coordinates of the quadrature points, x = x_0 + J xi with the affine
 * map of mesh/src/transformation.rs:34-64: xq[(e * nq + q) * dim + r]; *nq is
 * set to the rule's point count (pass xq = NULL to query it).  Host only. */
int femgpu_simplex_qp_coords(int dim, uint64_t n_nodes, const double *coords, uint64_t n_elem,
                             const uint32_t *elem_nodes, int quad_order, double *xq, uint64_t *nq);
/* Assembler::assemble_bilinear (assembly/src/assembler.rs:147-249) with
 * DiffusionIntegrator (form FEMGPU_DIFFUSION, standard/diffusion.rs:34-55) or
 * MassIntegrator (FEMGPU_MASS, standard/mass.rs:28-40), constant coefficient,
 * followed by CooMatrix::into_csr (linalg/src/coo.rs:76-), on the device:
 * returns a CsrMatrix<f64> operator (columns sorted within rows, duplicates
 * summed in element order).  Mesh arguments as femgpu_pa_create_simplex. */
int femgpu_csr_assemble_simplex(femgpu_ctx *ctx, int dim, int order, uint64_t n_nodes,
                                const double *coords, uint64_t n_elem, const uint32_t *elem_nodes,
                                const uint32_t *elem_dofs, uint64_t n_dofs, int form, double coef,
                                int quad_order, femgpu_op **out);
/* femgpu_csr_assemble_simplex with the coefficient given per quadrature point
 * (DiffusionIntegrator / MassIntegrator with a ScalarCoeff evaluated at each
 * point, standard/diffusion.rs:34-55, mass.rs:28-40), coef_qp as
 * femgpu_pa_create_simplex_qp. */
int femgpu_csr_assemble_simplex_qp(femgpu_ctx *ctx, int dim, int order, uint64_t n_nodes,
                                   const double *coords, uint64_t n_elem, const uint32_t *elem_nodes,
                                   const uint32_t *elem_dofs, uint64_t n_dofs, int form,
                                   const double *coef_qp, int quad_order, femgpu_op **out);
/* Assembler::assemble_linear with DomainSourceIntegrator
 * (assembly/src/standard/source.rs:44-50): b_i = sum_e sum_q w |det J| f(x_q)
 * phi_i(x_q), f = value (FEMGPU_SOURCE_CONST) or value * prod_d sin(pi x_d)
 * (FEMGPU_SOURCE_SINPROD; -lap u = f with u = sin sin takes value = 2 pi^2).  b: f64[n_dofs]. */
typedef enum { FEMGPU_SOURCE_CONST = 0, FEMGPU_SOURCE_SINPROD = 1 } femgpu_source_kind;
int femgpu_rhs_assemble_simplex(femgpu_ctx *ctx, int dim, int order, uint64_t n_nodes,
                                const double *coords, uint64_t n_elem, const uint32_t *elem_nodes,
                                const uint32_t *elem_dofs, uint64_t n_dofs, int source, double value,
                                int quad_order, void *b, femgpu_mem mem);
/* The matrix-free essential-BC pattern of examples/src/maxwell.rs:142-197 as a
 * reusable operator: y = A x0 (x0 = x with the constrained entries zeroed),
 * then y_d = x_d on every constrained DOF; diagonal 1 there.  The wrapper
 * borrows `inner` (which must outlive it; destroy with femgpu_op_destroy).
 * femgpu_essential_mask_rhs zeroes b on the constrained DOFs. */
int femgpu_essential_create(femgpu_op *inner, uint64_t n_constrained, const uint32_t *dofs,
                            femgpu_op **out);
int femgpu_essential_mask_rhs(femgpu_op *op, void *b, femgpu_mem mem);
/* LumpedMassOperator::assemble (partial.rs:217-262): y = diag(m) x. */
int femgpu_lumped_mass_create(femgpu_ctx *ctx, int dim, int order, uint64_t n_nodes,
                              const double *coords, uint64_t n_elem, const uint32_t *elem_nodes,
                              const uint32_t *elem_dofs, uint64_t n_dofs, double rho,
                              int quad_order, femgpu_op **out);
/* LumpedMassOperator with rho given per quadrature point (rho_qp as coef_qp
 * of femgpu_pa_create_simplex_qp). */
int femgpu_lumped_mass_create_qp(femgpu_ctx *ctx, int dim, int order, uint64_t n_nodes,
                                 const double *coords, uint64_t n_elem, const uint32_t *elem_nodes,
                                 const uint32_t *elem_dofs, uint64_t n_dofs, const double *rho_qp,
                                 int quad_order, femgpu_op **out);
/* LumpedMassOperator::apply_inverse (partial.rs:269-273). */
int femgpu_lumped_mass_apply_inverse(femgpu_op *op, const void *x, void *y, femgpu_mem mem);

/* ---------------------------------------------------------------- AMG ---- */
/* AmgSolver::setup / solve and solve_amg_cg (amg/src/lib.rs:125-207).  The
 * reference builds linger's AmgHierarchy (absent); this is smoothed
 * aggregation: strength threshold theta (default 0.08), at most max_levels
 * levels (10), coarsening until at most coarse_max rows (50, inverted dense),
 * Chebyshev-Jacobi smoothing of degree smoother_degree (2) on
 * [lambda_max / smoother_ratio, lambda_max] (ratio 10).  Zero fields take the
 * defaults.  The hierarchy is built on the host from an f64 CSR operator
 * (which must outlive it); V-cycles and PCG run on the device. */
typedef struct femgpu_amg femgpu_amg;
typedef struct {
    double theta;
    int max_levels;
    uint64_t coarse_max;
    int smoother_degree;
    double smoother_ratio;
} femgpu_amg_config;
int femgpu_amg_create(femgpu_op *csr, const femgpu_amg_config *cfg, femgpu_amg **out);
int femgpu_amg_destroy(femgpu_amg *amg);
/* AmgSolver::n_levels */
int femgpu_amg_n_levels(const femgpu_amg *amg, int *n);
/* Level `level`'s operator (which = 0) or prolongator to the next level
 * (which = 1) as host CSR; arrays may be NULL to query the sizes. */
int femgpu_amg_level(const femgpu_amg *amg, int level, int which, uint64_t *nrows, uint64_t *ncols,
                     uint64_t *nnz, uint64_t *row_ptr, uint32_t *col_idx, double *values);
/* AmgSolver::with_cycle (amg/src/lib.rs:176-179): the cycle the
 * preconditioner runs, V (default), W or F. */
enum { FEMGPU_CYCLE_V = 0, FEMGPU_CYCLE_W = 1, FEMGPU_CYCLE_F = 2 };
int femgpu_amg_set_cycle(femgpu_amg *amg, int cycle);
/* x = one cycle (of the set kind) applied to b (from x = 0): the
 * preconditioner. */
int femgpu_amg_vcycle(femgpu_amg *amg, const void *b, void *x, femgpu_mem mem);

/* ------------------------------------------------------------ solvers ---- */
/* SolverConfig (solver/src/lib.rs:97-113): defaults rtol 1e-8, atol 0,
 * max_iter 1000.  print_level is PrintLevel (lib.rs:79-94): 0 Silent,
 * 1 Summary, 2 Iterations, 3 Debug; the effective level is print_level when
 * not Silent, else Iterations when verbose (effective_print_level,
 * lib.rs:133-142).  Summary prints one line to stderr after a solve;
 * Iterations and Debug also record the residual of every iteration on the
 * device (no host round trip per iteration) and print them after the solve,
 * one line each, as par_solve_pcg_jacobi logs them (parallel/src/par_solver.rs
 * :312-313); femgpu_residual_history returns them.  The CG, PCG-Jacobi,
 * PCG-Chebyshev and AMG-PCG solvers report; the refinement, flexible-CG and
 * allocation-free solvers ignore the level. */
typedef struct {
    double rtol;
    double atol;
    uint64_t max_iter;
    int verbose;
    int print_level;
} femgpu_solver_config;

/* Residual history of the last solve on ctx that recorded one (print level
 * >= 2): len = iterations + 1 entries (entry 0: the initial residual; then the
 * residual each iteration's stopping test used, relative for PCG and absolute
 * for CG, as final_residual), at most 2^20; the first min(cap, len) are
 * copied to out when out is not NULL. */
int femgpu_residual_history(femgpu_ctx *ctx, double *out, uint64_t cap, uint64_t *len);

/* SolveResult (solver/src/lib.rs:71-75).  final_residual is absolute for
 * femgpu_solve_cg_operator (as solve_cg_operator) and relative otherwise
 * (as par_solve_pcg_jacobi); final_residual_is_relative says which. */
typedef struct {
    int converged;
    uint64_t iterations;
    double final_residual;
    int final_residual_is_relative;
    double error_estimate; /* reserved */
} femgpu_solve_result;

/* solve_cg_operator (solver/src/lib.rs:198-278): unpreconditioned CG with the
 * operator's overwrite apply.  Errors: BREAKDOWN if |p^T A p| < 1e-32,
 * NOT_CONVERGED on max_iter (result still filled). */
int femgpu_solve_cg_operator(femgpu_op *op, const void *b, void *x, femgpu_mem mem,
                             const femgpu_solver_config *cfg, femgpu_solve_result *res);
/* solve_pcg_jacobi (solver/src/lib.rs:281-297), with the semantics of the
 * in-tree PCG-Jacobi par_solve_pcg_jacobi (parallel/src/par_solver.rs:262-346):
 * inv_diag = 1/d (1 if |d| <= 1e-30), stop when |r|/|b| < rtol or |r| < atol.
 * Works for any operator (CSR or PA).  NOT_CONVERGED on max_iter. */
int femgpu_solve_pcg_jacobi(femgpu_op *op, const void *b, void *x, femgpu_mem mem,
                            const femgpu_solver_config *cfg, femgpu_solve_result *res);
/* No counterpart in the reference: femgpu_solve_pcg_jacobi for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels, iterates and result
 * as femgpu_solve_pcg_jacobi with FEMGPU_DEVICE vectors, but the call allocates
 * nothing: its work vectors (10 of n entries) live in a fixed device array of
 * FEMGPU_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends.
 * b and x are device pointers; op must be an assembled CSR operator
 * (femgpu_csr_create / femgpu_csr_assemble_simplex), whose apply needs no
 * workspace.  n above 838,848 (f64) or 1,677,696 (f32) returns INVALID_ARG.
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents (each call writes every vector it reads first): calls that overlap
 * on one GPU, from several contexts or threads, race. */
#define FEMGPU_NOALLOC_SCRATCH_BYTES (64u << 20)
int femgpu_solve_pcg_jacobi_noalloc(femgpu_op *op, const void *b, void *x,
                                    const femgpu_solver_config *cfg, femgpu_solve_result *res);
/* AmgSolver::solve (amg/src/lib.rs:188-206): PCG with one V-cycle as the
 * preconditioner, the stopping rule of par_solve_pcg_jacobi (relative
 * residual; converged = 0 at max_iter), on the AMG's operator. */
int femgpu_solve_amg_cg(femgpu_amg *amg, const void *b, void *x, femgpu_mem mem,
                        const femgpu_solver_config *cfg, femgpu_solve_result *res);
/* PCG as above with a textbook Chebyshev-Jacobi polynomial preconditioner
 * z = p(D^-1 A) D^-1 r of the given degree on [lmax/ratio, lmax], lmax =
 * safety * (power-iteration estimate of the largest eigenvalue of D^-1 A,
 * power_iters steps).  No reference counterpart beyond chebyshev_smooth
 * (amg/src/lib.rs:53-103). */
typedef struct {
    int degree;        /* >= 1 */
    double ratio;      /* lmax / lmin */
    double safety;     /* multiplies the estimate */
    int power_iters;
    double lmax;       /* > 0: use this instead of estimating */
} femgpu_cheb_config;
int femgpu_solve_pcg_chebyshev(femgpu_op *op, const femgpu_cheb_config *cheb, const void *b,
                               void *x, femgpu_mem mem, const femgpu_solver_config *cfg,
                               femgpu_solve_result *res, double *lmax_used);
/* chebyshev_smooth (amg/src/lib.rs:53-103), the exact reference recurrence
 * (d0 = (2/alpha) D^-1 r0, rho_k = 1/(1 - (delta/2alpha)^2 rho_{k-1}),
 * d_k = rho_k((2/alpha) D^-1 r_k + (rho_k - 1) d_{k-1})) on any operator. */
int femgpu_chebyshev_smooth(femgpu_op *op, void *x, const void *b, double lambda_min,
                            double lambda_max, int n_iter, femgpu_mem mem);
/* estimate_spectral_radius (amg/src/lib.rs:108-123). */
int femgpu_estimate_spectral_radius(femgpu_op *op, int n_iter, double *rho);

/* Mixed-precision iterative refinement (no reference counterpart): outer
 * FP64 residual r = b - A64 u, update u += e and the norms in FP64 on the
 * GPU; inner Chebyshev-Jacobi PCG in FP32 (op32: same operator in FP32) with
 * FP64 dot products on r/|r|, to max(inner_rtol, 0.5 rtol |b|/|r|), i.e. no
 * further than the outer target needs.  Converges when |r|/|b| < cfg->rtol;
 * cfg->max_iter bounds the total inner iterations.  Stops early when an outer
 * step does not reduce |r|. */
typedef struct {
    double inner_rtol;         /* mode 0: inner relative tolerance, default 1e-4;
                                  mode 1: residual reduction per segment, default 1e-2 */
    uint64_t inner_max_iter;   /* mode 0: per inner solve; 0: cfg->max_iter */
    uint64_t max_outer;        /* mode 0: default 50 */
    femgpu_cheb_config cheb;
    int mode;                  /* 0: restarted inner solves; 1: continued: one FP32 PCG whose
                                  recursive residual is replaced by the FP64 residual
                                  b - A64 u (u updated in FP64 first) whenever it has
                                  dropped by inner_rtol, search direction kept */
    int update_interval;       /* mode 1: at most this many iterations between
                                  replacements; 0: no limit */
} femgpu_ir_config;
int femgpu_solve_ir_mixed(femgpu_op *op64, femgpu_op *op32, const femgpu_ir_config *ir,
                          const void *b, void *x, femgpu_mem mem, const femgpu_solver_config *cfg,
                          femgpu_solve_result *res, uint64_t *inner_iterations);

/* Textbook Chebyshev-Jacobi preconditioner application z = p(D^-1 A) D^-1 r
 * (the polynomial used inside femgpu_solve_pcg_chebyshev). */
int femgpu_chebyshev_jacobi_apply(femgpu_op *op, const femgpu_cheb_config *cheb, const void *r,
                                  void *z, femgpu_mem mem);

/* ------------------------------------------- learned preconditioner ---- */
/* No reference counterpart.  A 3-level 3-D convolutional U-Net (16/32/64
 * channels, BF16 weights and activations on MFMA with FP32 accumulation) on
 * a structured grid of nx*ny*nz points (the hex DOF grid).  Input channels:
 * the residual (scaled to unit RMS) and an auxiliary field (log kappa).
 * Parameters are FP32 masters laid out per layer as W[cout][kpad] (k = tap *
 * cin + c, tap = (dz+1)*9 + (dy+1)*3 + (dx+1)) followed by bias[cout]. */
typedef struct femgpu_unet femgpu_unet;
int femgpu_unet_create(femgpu_ctx *ctx, int nx, int ny, int nz, uint64_t seed, femgpu_unet **out);
/* The same network placed on a coarser grid: for a DOF grid of nx*ny*nz
 * points (n - 1 divisible by coarsen) it runs on every coarsen-th node
 * ((n - 1)/coarsen + 1 per direction).  Residuals reach it by R = P^T and its
 * output returns to the DOF grid by trilinear interpolation P, so its 3
 * levels see a 2^3 * coarsen times wider part of the domain.  The training
 * step, the flexible CG and the refinement take DOF-grid vectors as with
 * femgpu_unet_create; forward / backward / set_aux work on the network's
 * grid (femgpu_unet_grid).  With femgpu_unet_aux_from_hex on an operator
 * whose boundary DOFs are essential, the correction is zero there. */
int femgpu_unet_create_coarse(femgpu_ctx *ctx, int nx, int ny, int nz, int coarsen, uint64_t seed,
                              femgpu_unet **out);
int femgpu_unet_grid(const femgpu_unet *net, int *nx, int *ny, int *nz);
int femgpu_unet_destroy(femgpu_unet *net);
int femgpu_unet_n_params(const femgpu_unet *net, uint64_t *n);
int femgpu_unet_layer_info(const femgpu_unet *net, int layer, int *cin, int *cout, int *kpad,
                           uint64_t *woff, uint64_t *boff);
int femgpu_unet_get_params(femgpu_unet *net, float *host);
int femgpu_unet_set_params(femgpu_unet *net, const float *host);
/* Auxiliary input channel (f32 per grid point). */
int femgpu_unet_set_aux(femgpu_unet *net, const float *aux, femgpu_mem mem);
/* aux = log(kappa)/log(100) at the DOF nodes of a hex operator of the same grid. */
int femgpu_unet_aux_from_hex(femgpu_unet *net, femgpu_op *hex);
/* out = N(x0, aux) (raw, no scaling); keeps activations for the backward. */
int femgpu_unet_forward(femgpu_unet *net, const float *x0, float *out, femgpu_mem mem);
/* grads = d(sum dout . out)/d params at the last forward input. */
int femgpu_unet_backward(femgpu_unet *net, const float *dout, float *grads, femgpu_mem mem);
/* One Adam step on a (residual, error) pair, both normalised to unit RMS;
 * loss = 0.5 |N - t|^2 / |t|^2. */
int femgpu_unet_train_step(femgpu_unet *net, const float *r, const float *e, femgpu_mem mem,
                           float lr, double *loss);
/* One Adam step on the energy-norm loss of the Galerkin-scaled correction:
 * with n = N(r sqrt(V)/|r|) and g = n.r / n.A n,
 *   loss = |e - g n|_A^2 / |e|_A^2 = 1 - (n.r)^2 / ((n.A n)(e.r)),
 * the relative A-norm error the preconditioner leaves (r = A e; e only sets
 * the normalisation).  r, e are f32; A is applied in op's precision. */
int femgpu_unet_train_step_energy(femgpu_unet *net, femgpu_op *op, const float *r, const float *e,
                                  femgpu_mem mem, float lr, double *loss);
/* z = |r|/sqrt(V) N(r sqrt(V)/|r|), vectors of type dtype. */
int femgpu_unet_apply(femgpu_unet *net, const void *r, void *z, femgpu_dtype dtype, femgpu_mem mem);
/* Weights + Adam state, with a 64-bit FNV-1a checksum. */
int femgpu_unet_save(femgpu_unet *net, const char *path);
int femgpu_unet_load(femgpu_unet *net, const char *path);

/* Flexible CG (Notay's FCG, truncated to the last `truncation` directions)
 * with a two-level preconditioner: Chebyshev-Jacobi smoothing, then the
 * network's correction of the smooth residual the smoother leaves, scaled by
 * its optimal (Galerkin) step:
 *   z1 = C r, r1 = r - A z1, n = N(r1), g = n.r1 / n.A n, z = z1 + g n.
 * A p comes from the same recurrence as p (A z = A z1 + g A n).  Falls back to
 * plain PCG + Chebyshev-Jacobi from the current iterate on stagnation (no new
 * best residual for `patience` iterations), a non-finite value or loss of
 * p^T A p > 0.  Convergence is judged on |b - A x|/|b| < rtol with the
 * residual recomputed in the operator's precision. */
typedef struct {
    int patience;              /* iterations without a new best residual; default 20 */
    int truncation;            /* directions kept for A-orthogonalisation; default 8, max 64 */
    uint64_t fcg_iterations;   /* out: iterations done in FCG */
    uint64_t fallback_iterations; /* out: PCG iterations after a fallback */
    int fell_back;             /* out: 1 if the fallback ran */
    int fallback_reason;       /* out: 1 stagnation, 2 non-finite, 3 p^T A p <= 0 */
    uint64_t net_corrections;  /* out: preconditioner applications that used the network */
} femgpu_fcg_info;
int femgpu_solve_fcg_learned(femgpu_op *op, femgpu_unet *net, const femgpu_cheb_config *cheb,
                             const void *b, void *x, femgpu_mem mem,
                             const femgpu_solver_config *cfg, femgpu_solve_result *res,
                             femgpu_fcg_info *info);
/* Mixed precision with the learned preconditioner: the refinement of
 * femgpu_solve_ir_mixed (FP64 residual, update and norms) around an FP32
 * inner solve by the flexible CG above (network + Chebyshev-Jacobi on op32).
 * info (patience, truncation in; totals over the inner solves out). */
int femgpu_solve_ir_learned(femgpu_op *op64, femgpu_op *op32, femgpu_unet *net,
                            const femgpu_ir_config *ir, const void *b, void *x, femgpu_mem mem,
                            const femgpu_solver_config *cfg, femgpu_solve_result *res,
                            uint64_t *inner_iterations, femgpu_fcg_info *info);

#ifdef __cplusplus
}
#endif
#endif
