//! fem-rs's operator and solver interfaces on one MI355X, over libfemgpu.
//!
//! - [`GpuOperator`] implements `fem_assembly::partial::MatFreeOperator`
//!   (assembly/src/partial.rs:45-57: `apply` accumulates, `apply_zero`
//!   overwrites) and `fem_assembly::backend::LinearOperator`
//!   (assembly/src/backend.rs:39-46).
//! - Operators are built from a `CsrMatrix<f64>` (linalg/src/csr.rs:81-93,
//!   copied as is), from an `FESpace` + `ScalarCoeff` (the simplex
//!   `PAMassOperator` / `PADiffusionOperator` of partial.rs:65-209, with the
//!   coefficient evaluated at every quadrature point as the reference does),
//!   as structured hex Q_p operators, or from a closure.
//! - [`solve_pcg_jacobi`], [`solve_cg_operator`], [`chebyshev_smooth`] and
//!   [`estimate_spectral_radius`] keep the reference signatures and
//!   `SolverError` variants (solver/src/lib.rs:47-55, 198-297;
//!   amg/src/lib.rs:53-123).
//! - [`Backend`] adds the `"hip"` arm to `OperatorBackend::parse`
//!   (backend.rs:21-28) and [`HipCeed`] serves the `/gpu/hip` resource of
//!   `FemCeed::from_backend_resource` (reed/context.rs:83-100).
use std::ffi::CStr;
use std::os::raw::{c_int, c_void};
use std::sync::Arc;

use fem_assembly::backend::{LinearOperator, OperatorBackend};
use fem_assembly::coefficient::{CoeffCtx, ScalarCoeff};
use fem_assembly::partial::MatFreeOperator;
use fem_gpu_sys as sys;
use fem_linalg::CsrMatrix;
use fem_mesh::topology::MeshTopology;
use fem_solver::{SolveResult, SolverConfig, SolverError};
use fem_space::fe_space::FESpace;

/// Message of the last failing libfemgpu call on this thread.
pub fn last_error() -> String {
    unsafe { CStr::from_ptr(sys::femgpu_last_error()).to_string_lossy().into_owned() }
}

fn check(rc: c_int) -> Result<(), SolverError> {
    match rc {
        sys::FEMGPU_OK => Ok(()),
        _ => Err(SolverError::Linger(last_error())),
    }
}

// ------------------------------------------------------------------ context
/// One GPU: a libfemgpu context (its stream and solver scratch).
pub struct GpuContext {
    raw: *mut sys::femgpu_ctx,
}
// libfemgpu serialises every call on a context with the context's mutex.
unsafe impl Send for GpuContext {}
unsafe impl Sync for GpuContext {}

impl GpuContext {
    pub fn new(device: i32) -> Result<Arc<Self>, SolverError> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { sys::femgpu_ctx_create(device, &mut raw) })?;
        Ok(Arc::new(GpuContext { raw }))
    }
    pub fn raw(&self) -> *mut sys::femgpu_ctx {
        self.raw
    }
    /// Residuals of the last solve with `print_level >= Iterations` (entry 0: initial).
    pub fn residual_history(&self) -> Vec<f64> {
        let mut len = 0u64;
        unsafe { sys::femgpu_residual_history(self.raw, std::ptr::null_mut(), 0, &mut len) };
        let mut out = vec![0.0; len as usize];
        unsafe { sys::femgpu_residual_history(self.raw, out.as_mut_ptr(), len, &mut len) };
        out
    }
}

impl Drop for GpuContext {
    fn drop(&mut self) {
        unsafe { sys::femgpu_ctx_destroy(self.raw) };
    }
}

// ----------------------------------------------------------------- operator
/// A square operator resident on the GPU (CSR, PA, hex PA or closure).
pub struct GpuOperator {
    ctx: Arc<GpuContext>,
    raw: *mut sys::femgpu_op,
    n: usize,
    // keeps a closure operator's state alive as long as the operator
    _user: Option<Box<dyn std::any::Any + Send + Sync>>,
}
unsafe impl Send for GpuOperator {}
unsafe impl Sync for GpuOperator {}

impl Drop for GpuOperator {
    fn drop(&mut self) {
        unsafe { sys::femgpu_op_destroy(self.raw) };
    }
}

impl GpuOperator {
    fn adopt(ctx: &Arc<GpuContext>, raw: *mut sys::femgpu_op) -> Self {
        let mut n = 0u64;
        unsafe { sys::femgpu_op_n_dofs(raw, &mut n) };
        GpuOperator { ctx: ctx.clone(), raw, n: n as usize, _user: None }
    }
    pub fn raw(&self) -> *mut sys::femgpu_op {
        self.raw
    }

    /// `CsrMatrix<f64>` copied to the device in the reference layout
    /// (row_ptr as u64, col_idx u32, values f64).
    pub fn from_csr(ctx: &Arc<GpuContext>, a: &CsrMatrix<f64>) -> Result<Self, SolverError> {
        let rp: Vec<u64> = a.row_ptr.iter().map(|&v| v as u64).collect();
        let mut raw = std::ptr::null_mut();
        check(unsafe {
            sys::femgpu_csr_create(ctx.raw, a.nrows as u64, a.ncols as u64, rp.as_ptr(), a.col_idx.as_ptr(),
                                   a.values.as_ptr().cast(), sys::FEMGPU_F64, &mut raw)
        })?;
        Ok(Self::adopt(ctx, raw))
    }

    /// `CsrMatrix::spmv_add`: y = alpha (A x) + beta y (csr.rs:167-193).  Panics on
    /// a length mismatch, as the reference does.
    pub fn spmv_add(&self, alpha: f64, x: &[f64], beta: f64, y: &mut [f64]) {
        assert_eq!((x.len(), y.len()), (self.n, self.n), "spmv_add: dimension mismatch");
        let rc = unsafe {
            sys::femgpu_csr_spmv_add(self.raw, alpha, x.as_ptr().cast(), beta, y.as_mut_ptr().cast(),
                                     sys::FEMGPU_HOST)
        };
        assert_eq!(rc, 0, "{}", last_error());
    }

    /// `apply_dirichlet` row zeroing (constraints.rs:31-42) on the device matrix.
    pub fn apply_dirichlet(&self, dofs: &[u32], values: &[f64], rhs: &mut [f64]) -> Result<(), SolverError> {
        assert_eq!(dofs.len(), values.len());
        check(unsafe {
            sys::femgpu_csr_apply_dirichlet(self.raw, dofs.len() as u64, dofs.as_ptr(), values.as_ptr(),
                                            rhs.as_mut_ptr().cast(), sys::FEMGPU_HOST)
        })
    }

    /// `CsrMatrix::diagonal` / the PA operators' assembled diagonal.
    pub fn diagonal(&self) -> Vec<f64> {
        let mut d = vec![0.0; self.n];
        let rc = unsafe { sys::femgpu_op_diagonal(self.raw, d.as_mut_ptr().cast(), sys::FEMGPU_HOST) };
        assert_eq!(rc, 0, "{}", last_error());
        d
    }

    /// `PADiffusionOperator::new(space, kappa, quad_order)` / `PAMassOperator::new`
    /// (partial.rs:65-79, 138-148) on a simplex space (Tri3 / Tet4, P1..P3):
    /// the coefficient is evaluated once at every physical quadrature point,
    /// with the `CoeffCtx` the reference builds there (partial.rs:137-147).
    pub fn simplex_pa<S: FESpace, C: ScalarCoeff>(ctx: &Arc<GpuContext>, space: &S, coef: &C, diffusion: bool,
                                                  quad_order: u8) -> Result<Self, SolverError> {
        let mesh = space.mesh();
        let dim = mesh.dim() as usize;
        let ne = mesh.n_elements();
        let nn = mesh.n_nodes();
        let mut coords = Vec::with_capacity(nn * dim);
        for v in 0..nn as u32 {
            coords.extend_from_slice(mesh.node_coords(v));
        }
        let mut nodes = Vec::with_capacity(ne * (dim + 1));
        let mut edofs = Vec::new();
        for e in 0..ne as u32 {
            nodes.extend_from_slice(&mesh.element_nodes(e)[..dim + 1]);
            edofs.extend_from_slice(space.element_dofs(e));
        }
        let mut nq = 0u64;
        check(unsafe {
            sys::femgpu_simplex_qp_coords(dim as c_int, nn as u64, coords.as_ptr(), ne as u64, nodes.as_ptr(),
                                          quad_order as c_int, std::ptr::null_mut(), &mut nq)
        })?;
        let nq = nq as usize;
        let mut xq = vec![0.0; ne * nq * dim];
        check(unsafe {
            sys::femgpu_simplex_qp_coords(dim as c_int, nn as u64, coords.as_ptr(), ne as u64, nodes.as_ptr(),
                                          quad_order as c_int, xq.as_mut_ptr(), &mut (nq as u64))
        })?;
        let form = if diffusion { sys::FEMGPU_DIFFUSION } else { sys::FEMGPU_MASS };
        let mut raw = std::ptr::null_mut();
        if coef.is_constant() {
            let c = coef.eval(&CoeffCtx::from_qp(&xq[..dim], dim, 0, mesh.element_tag(0), None, None));
            check(unsafe {
                sys::femgpu_pa_create_simplex(ctx.raw, dim as c_int, space.order() as c_int, nn as u64,
                                              coords.as_ptr(), ne as u64, nodes.as_ptr(), edofs.as_ptr(),
                                              space.n_dofs() as u64, form, c, quad_order as c_int, &mut raw)
            })?;
        } else {
            let mut cq = vec![0.0; ne * nq];
            for e in 0..ne {
                let tag = mesh.element_tag(e as u32);
                for q in 0..nq {
                    let x = &xq[(e * nq + q) * dim..(e * nq + q + 1) * dim];
                    cq[e * nq + q] = coef.eval(&CoeffCtx::from_qp(x, dim, e as u32, tag, None, None));
                }
            }
            check(unsafe {
                sys::femgpu_pa_create_simplex_qp(ctx.raw, dim as c_int, space.order() as c_int, nn as u64,
                                                 coords.as_ptr(), ne as u64, nodes.as_ptr(), edofs.as_ptr(),
                                                 space.n_dofs() as u64, form, cq.as_ptr(), quad_order as c_int,
                                                 &mut raw)
            })?;
        }
        Ok(Self::adopt(ctx, raw))
    }

    /// Structured hex Q_p PA operator (no reference counterpart: the reference's
    /// PA is simplex-only, partial.rs:278-288).  `vertices`: (nx+1)(ny+1)(nz+1)
    /// points, i fastest; `None` = the unit cube.
    pub fn hex_pa(ctx: &Arc<GpuContext>, desc: &sys::femgpu_hex_desc, vertices: Option<&[f64]>)
                  -> Result<Self, SolverError> {
        let mut raw = std::ptr::null_mut();
        let v = vertices.map(|v| v.as_ptr()).unwrap_or(std::ptr::null());
        check(unsafe { sys::femgpu_pa_create_hex(ctx.raw, desc, v, &mut raw) })?;
        Ok(Self::adopt(ctx, raw))
    }

    /// The same hex operator in another precision (FP32 inner solves).
    pub fn convert(&self, dtype: sys::femgpu_dtype) -> Result<Self, SolverError> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { sys::femgpu_pa_hex_convert(self.raw, dtype, &mut raw) })?;
        Ok(Self::adopt(&self.ctx, raw))
    }

    /// An operator whose apply is the caller's closure `y <- A x` (the `apply: F` of
    /// solve_cg_operator, solver/src/lib.rs:198-208); `diag` enables Jacobi.
    pub fn from_fn<F>(ctx: &Arc<GpuContext>, n: usize, apply: F, diag: Option<&[f64]>) -> Result<Self, SolverError>
    where
        F: Fn(&[f64], &mut [f64]) + Send + Sync + 'static,
    {
        unsafe extern "C" fn tramp<F: Fn(&[f64], &mut [f64])>(user: *mut c_void, x: *const c_void,
                                                                y: *mut c_void, n: u64) -> c_int {
            let f = &*(user as *const F);
            let xs = std::slice::from_raw_parts(x as *const f64, n as usize);
            let ys = std::slice::from_raw_parts_mut(y as *mut f64, n as usize);
            match std::panic::catch_unwind(std::panic::AssertUnwindSafe(|| f(xs, ys))) {
                Ok(()) => 0,
                Err(_) => 1,
            }
        }
        let boxed: Box<F> = Box::new(apply);
        let user = &*boxed as *const F as *mut c_void;
        let mut raw = std::ptr::null_mut();
        let d = diag.map(|d| d.as_ptr().cast()).unwrap_or(std::ptr::null());
        check(unsafe {
            sys::femgpu_op_create_callback(ctx.raw, n as u64, sys::FEMGPU_F64, sys::FEMGPU_HOST, Some(tramp::<F>),
                                           user, d, &mut raw)
        })?;
        let mut op = Self::adopt(ctx, raw);
        op._user = Some(boxed);
        Ok(op)
    }

    fn apply_host(&self, x: &[f64], y: &mut [f64], accumulate: bool) {
        assert_eq!((x.len(), y.len()), (self.n, self.n), "apply: dimension mismatch");
        let rc = unsafe {
            sys::femgpu_op_apply(self.raw, x.as_ptr().cast(), y.as_mut_ptr().cast(), accumulate as c_int,
                                 sys::FEMGPU_HOST)
        };
        assert_eq!(rc, 0, "{}", last_error());
    }
}

impl MatFreeOperator for GpuOperator {
    /// y += A x (partial.rs:45-47)
    fn apply(&self, x: &[f64], y: &mut [f64]) {
        self.apply_host(x, y, true)
    }
    /// y = A x (partial.rs:50-53)
    fn apply_zero(&self, x: &[f64], y: &mut [f64]) {
        self.apply_host(x, y, false)
    }
    fn n_dofs(&self) -> usize {
        self.n
    }
}

impl LinearOperator for GpuOperator {
    fn nrows(&self) -> usize {
        self.n
    }
    fn ncols(&self) -> usize {
        self.n
    }
    /// y <- A x (backend.rs:44-45)
    fn apply(&self, x: &[f64], y: &mut [f64]) {
        self.apply_host(x, y, false)
    }
}

// ------------------------------------------------------------------ solvers
fn cfg_c(cfg: &SolverConfig) -> sys::femgpu_solver_config {
    use fem_solver::PrintLevel;
    sys::femgpu_solver_config {
        rtol: cfg.rtol,
        atol: cfg.atol,
        max_iter: cfg.max_iter as u64,
        verbose: cfg.verbose as c_int,
        print_level: match cfg.print_level {
            PrintLevel::Silent => 0,
            PrintLevel::Summary => 1,
            PrintLevel::Iterations => 2,
            PrintLevel::Debug => 3,
        },
    }
}

fn result(r: &sys::femgpu_solve_result) -> SolveResult {
    SolveResult { converged: r.converged != 0, iterations: r.iterations as usize, final_residual: r.final_residual }
}

/// solve_pcg_jacobi (solver/src/lib.rs:281-297) with par_solve_pcg_jacobi's
/// semantics (parallel/src/par_solver.rs:262-346): Ok(converged = false) at
/// max_iter.
pub fn solve_pcg_jacobi(a: &GpuOperator, b: &[f64], x: &mut [f64], cfg: &SolverConfig)
                        -> Result<SolveResult, SolverError> {
    if a.n != b.len() || a.n != x.len() {
        return Err(SolverError::DimensionMismatch { rows: a.n, cols: a.n, rhs: b.len() });
    }
    let c = cfg_c(cfg);
    let mut r = sys::femgpu_solve_result::default();
    let rc = unsafe {
        sys::femgpu_solve_pcg_jacobi(a.raw, b.as_ptr().cast(), x.as_mut_ptr().cast(), sys::FEMGPU_HOST, &c, &mut r)
    };
    match rc {
        sys::FEMGPU_OK | sys::FEMGPU_ERR_NOT_CONVERGED => Ok(result(&r)),
        _ => Err(SolverError::Linger(last_error())),
    }
}

/// solve_cg_operator (solver/src/lib.rs:198-278): unpreconditioned CG on the
/// closure `apply(x, y)` (y <- A x); absolute final residual, breakdown ->
/// Linger, max_iter -> ConvergenceFailed.
pub fn solve_cg_operator<F>(ctx: &Arc<GpuContext>, nrows: usize, ncols: usize, apply: F, b: &[f64],
                            x: &mut [f64], cfg: &SolverConfig) -> Result<SolveResult, SolverError>
where
    F: Fn(&[f64], &mut [f64]) + Send + Sync + 'static,
{
    if nrows != ncols || b.len() != nrows || x.len() != ncols {
        return Err(SolverError::DimensionMismatch { rows: nrows, cols: ncols, rhs: b.len() });
    }
    let op = GpuOperator::from_fn(ctx, nrows, apply, None)?;
    solve_cg_on(&op, b, x, cfg)
}

/// solve_cg_operator on an operator already on the GPU.
pub fn solve_cg_on(op: &GpuOperator, b: &[f64], x: &mut [f64], cfg: &SolverConfig)
                   -> Result<SolveResult, SolverError> {
    let c = cfg_c(cfg);
    let mut r = sys::femgpu_solve_result::default();
    let rc = unsafe {
        sys::femgpu_solve_cg_operator(op.raw, b.as_ptr().cast(), x.as_mut_ptr().cast(), sys::FEMGPU_HOST, &c, &mut r)
    };
    match rc {
        sys::FEMGPU_OK => Ok(result(&r)),
        sys::FEMGPU_ERR_NOT_CONVERGED => {
            Err(SolverError::ConvergenceFailed { max_iter: cfg.max_iter, residual: r.final_residual })
        }
        sys::FEMGPU_ERR_BREAKDOWN => Err(SolverError::Linger("CG operator breakdown: p^T A p is near zero".into())),
        _ => Err(SolverError::Linger(last_error())),
    }
}

/// chebyshev_smooth (amg/src/lib.rs:53-103), the reference recurrence, on any
/// GPU operator.
pub fn chebyshev_smooth(a: &GpuOperator, x: &mut [f64], b: &[f64], lambda_min: f64, lambda_max: f64, n_iter: usize) {
    assert_eq!((x.len(), b.len()), (a.n, a.n));
    let rc = unsafe {
        sys::femgpu_chebyshev_smooth(a.raw, x.as_mut_ptr().cast(), b.as_ptr().cast(), lambda_min, lambda_max,
                                     n_iter as c_int, sys::FEMGPU_HOST)
    };
    assert_eq!(rc, 0, "{}", last_error());
}

/// estimate_spectral_radius (amg/src/lib.rs:108-123).
pub fn estimate_spectral_radius(a: &GpuOperator, n_iter: usize) -> f64 {
    let mut rho = 0.0;
    let rc = unsafe { sys::femgpu_estimate_spectral_radius(a.raw, n_iter as c_int, &mut rho) };
    assert_eq!(rc, 0, "{}", last_error());
    rho
}

/// PCG with the textbook Chebyshev-Jacobi preconditioner of `degree`; returns
/// the result and the lambda_max used.
pub fn solve_pcg_chebyshev(a: &GpuOperator, degree: i32, b: &[f64], x: &mut [f64], cfg: &SolverConfig)
                           -> Result<(SolveResult, f64), SolverError> {
    let cheb = sys::femgpu_cheb_config { degree, ratio: 30.0, safety: 1.1, power_iters: 20, lmax: 0.0 };
    let c = cfg_c(cfg);
    let mut r = sys::femgpu_solve_result::default();
    let mut lmax = 0.0;
    let rc = unsafe {
        sys::femgpu_solve_pcg_chebyshev(a.raw, &cheb, b.as_ptr().cast(), x.as_mut_ptr().cast(), sys::FEMGPU_HOST,
                                        &c, &mut r, &mut lmax)
    };
    match rc {
        sys::FEMGPU_OK | sys::FEMGPU_ERR_NOT_CONVERGED => Ok((result(&r), lmax)),
        _ => Err(SolverError::Linger(last_error())),
    }
}

/// FP32 inner PCG-Chebyshev with FP64 refinement (continued form) on a hex
/// operator pair (`op32 = op64.convert(FEMGPU_F32)`).
pub fn solve_ir_mixed(op64: &GpuOperator, op32: &GpuOperator, b: &[f64], x: &mut [f64], cfg: &SolverConfig)
                      -> Result<(SolveResult, u64), SolverError> {
    let ir = sys::femgpu_ir_config {
        inner_rtol: 1e-2,
        inner_max_iter: 0,
        max_outer: 50,
        cheb: sys::femgpu_cheb_config { degree: 4, ratio: 30.0, safety: 1.1, power_iters: 20, lmax: 0.0 },
        mode: 1,
        update_interval: 16,
    };
    let c = cfg_c(cfg);
    let mut r = sys::femgpu_solve_result::default();
    let mut inner = 0u64;
    let rc = unsafe {
        sys::femgpu_solve_ir_mixed(op64.raw, op32.raw, &ir, b.as_ptr().cast(), x.as_mut_ptr().cast(),
                                   sys::FEMGPU_HOST, &c, &mut r, &mut inner)
    };
    match rc {
        sys::FEMGPU_OK | sys::FEMGPU_ERR_NOT_CONVERGED => Ok((result(&r), inner)),
        _ => Err(SolverError::Linger(last_error())),
    }
}

// ------------------------------------------------------- backend selection
/// `OperatorBackend` with the GPU arm: `parse` accepts `"hip"` besides the
/// reference's `"native"` / `"reed"` (case-insensitive, backend.rs:21-28).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum Backend {
    Reference(OperatorBackend),
    Hip,
}

impl Backend {
    pub fn parse(s: &str) -> Option<Self> {
        if s.eq_ignore_ascii_case("hip") {
            return Some(Backend::Hip);
        }
        OperatorBackend::parse(s).map(Backend::Reference)
    }
    pub fn as_str(self) -> &'static str {
        match self {
            Backend::Reference(b) => b.as_str(),
            Backend::Hip => "hip",
        }
    }
}

/// The `/gpu/hip` resource of `FemCeed::from_backend_resource`
/// (reed/context.rs:83-100): `apply_mass_2d` / `apply_poisson_2d` on P1 / P2
/// triangles, returning a new vector as the reference does.  The Poisson form
/// is D = |det J| w J^-1 J^-T (the reference's FemCeed builds J^-T J^-1,
/// context.rs:282-284, which differs on non-normal J).
pub struct HipCeed {
    ctx: Arc<GpuContext>,
}

impl HipCeed {
    pub const RESOURCE: &'static str = "/gpu/hip";

    pub fn from_backend_resource(resource: &str, device: i32) -> Option<Result<Self, SolverError>> {
        if resource != Self::RESOURCE {
            return None; // the caller falls back to fem_assembly::reed::FemCeed
        }
        Some(GpuContext::new(device).map(|ctx| HipCeed { ctx }))
    }

    pub fn effective_resource(&self) -> &str {
        Self::RESOURCE
    }

    fn apply_2d<S: FESpace>(&self, space: &S, diffusion: bool, quad_order: u8, input: &[f64])
                            -> Result<Vec<f64>, SolverError> {
        if input.len() != space.n_dofs() {
            return Err(SolverError::DimensionMismatch { rows: space.n_dofs(), cols: space.n_dofs(), rhs: input.len() });
        }
        let op = GpuOperator::simplex_pa(&self.ctx, space, &1.0f64, diffusion, quad_order)?;
        let mut out = vec![0.0; input.len()];
        MatFreeOperator::apply_zero(&op, input, &mut out);
        Ok(out)
    }

    pub fn apply_mass_2d<S: FESpace>(&self, space: &S, quad_order: u8, input: &[f64]) -> Result<Vec<f64>, SolverError> {
        self.apply_2d(space, false, quad_order, input)
    }

    pub fn apply_poisson_2d<S: FESpace>(&self, space: &S, quad_order: u8, input: &[f64])
                                        -> Result<Vec<f64>, SolverError> {
        self.apply_2d(space, true, quad_order, input)
    }
}
