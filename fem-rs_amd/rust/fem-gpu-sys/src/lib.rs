//! Raw `extern "C"` bindings to libfemgpu: one declaration per entry point of
//! `include/femgpu.h`, with the same names, argument order and types.  Every
//! function returns a `femgpu_status` (0 = OK); `femgpu_last_error()` gives
//! the message of the last failing call on the calling thread.  The safe
//! wrappers that implement fem-rs's traits are in the `fem-gpu` crate.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

// ------------------------------------------------------------- status
pub const FEMGPU_OK: c_int = 0;
pub const FEMGPU_ERR_INVALID_ARG: c_int = 1;
pub const FEMGPU_ERR_DIM_MISMATCH: c_int = 2;
pub const FEMGPU_ERR_NOT_CONVERGED: c_int = 3;
pub const FEMGPU_ERR_BREAKDOWN: c_int = 4;
pub const FEMGPU_ERR_HIP: c_int = 5;
pub const FEMGPU_ERR_UNSUPPORTED: c_int = 6;
pub const FEMGPU_ERR_OUT_OF_MEMORY: c_int = 7;

pub type femgpu_dtype = c_int;
pub const FEMGPU_F64: femgpu_dtype = 0;
pub const FEMGPU_F32: femgpu_dtype = 1;

pub type femgpu_mem = c_int;
pub const FEMGPU_HOST: femgpu_mem = 0;
pub const FEMGPU_DEVICE: femgpu_mem = 1;

pub type femgpu_form = c_int;
pub const FEMGPU_MASS: femgpu_form = 1;
pub const FEMGPU_DIFFUSION: femgpu_form = 2;
pub const FEMGPU_DIFF_MASS: femgpu_form = 3;

pub const FEMGPU_COEF_CONST: c_int = 0;
pub const FEMGPU_COEF_SINPROD: c_int = 1;
pub const FEMGPU_COEF_GRF: c_int = 2;
pub const FEMGPU_COEF_QP_ARRAY: c_int = 3;

pub const FEMGPU_SOURCE_CONST: c_int = 0;
pub const FEMGPU_SOURCE_SINPROD: c_int = 1;

pub const FEMGPU_CYCLE_V: c_int = 0;
pub const FEMGPU_CYCLE_W: c_int = 1;
pub const FEMGPU_CYCLE_F: c_int = 2;

/// Bytes of the per-GPU scratch array of `femgpu_solve_pcg_jacobi_noalloc`.
pub const FEMGPU_NOALLOC_SCRATCH_BYTES: usize = 64 << 20;

// ------------------------------------------------------------- handles
#[repr(C)]
pub struct femgpu_ctx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct femgpu_op {
    _private: [u8; 0],
}
#[repr(C)]
pub struct femgpu_amg {
    _private: [u8; 0],
}
#[repr(C)]
pub struct femgpu_unet {
    _private: [u8; 0],
}

// ------------------------------------------------------------- structs
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct femgpu_coef {
    pub kind: c_int,
    pub value: f64,
    pub n_features: c_int,
    pub sigma: f64,
    pub length: f64,
    pub lo: f64,
    pub hi: f64,
    pub seed: u64,
    pub qp_values: *const f64,
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct femgpu_hex_desc {
    pub p: c_int,
    pub q: c_int,
    pub nx: c_int,
    pub ny: c_int,
    pub nz: c_int,
    pub form: c_int,
    pub dtype: femgpu_dtype,
    pub dirichlet: c_int,
    pub rho: f64,
    pub kappa: femgpu_coef,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct femgpu_amg_config {
    pub theta: f64,
    pub max_levels: c_int,
    pub coarse_max: u64,
    pub smoother_degree: c_int,
    pub smoother_ratio: f64,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct femgpu_solver_config {
    pub rtol: f64,
    pub atol: f64,
    pub max_iter: u64,
    pub verbose: c_int,
    pub print_level: c_int,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct femgpu_solve_result {
    pub converged: c_int,
    pub iterations: u64,
    pub final_residual: f64,
    pub final_residual_is_relative: c_int,
    pub error_estimate: f64,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct femgpu_cheb_config {
    pub degree: c_int,
    pub ratio: f64,
    pub safety: f64,
    pub power_iters: c_int,
    pub lmax: f64,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct femgpu_ir_config {
    pub inner_rtol: f64,
    pub inner_max_iter: u64,
    pub max_outer: u64,
    pub cheb: femgpu_cheb_config,
    pub mode: c_int,
    pub update_interval: c_int,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct femgpu_fcg_info {
    pub patience: c_int,
    pub truncation: c_int,
    pub fcg_iterations: u64,
    pub fallback_iterations: u64,
    pub fell_back: c_int,
    pub fallback_reason: c_int,
    pub net_corrections: u64,
}

/// `y = A x` for a callback operator; returns 0 on success.
pub type femgpu_apply_fn =
    Option<unsafe extern "C" fn(user: *mut c_void, x: *const c_void, y: *mut c_void, n: u64) -> c_int>;

extern "C" {
    // ---------------------------------------------------------- context
    pub fn femgpu_last_error() -> *const c_char;
    pub fn femgpu_version() -> *const c_char;
    pub fn femgpu_ctx_create(device: c_int, out: *mut *mut femgpu_ctx) -> c_int;
    pub fn femgpu_ctx_destroy(ctx: *mut femgpu_ctx) -> c_int;
    pub fn femgpu_ctx_synchronize(ctx: *mut femgpu_ctx) -> c_int;
    pub fn femgpu_malloc(ctx: *mut femgpu_ctx, bytes: usize, dptr: *mut *mut c_void) -> c_int;
    pub fn femgpu_free(ctx: *mut femgpu_ctx, dptr: *mut c_void) -> c_int;
    pub fn femgpu_memcpy(ctx: *mut femgpu_ctx, dst: *mut c_void, src: *const c_void, bytes: usize,
                         kind: c_int) -> c_int;
    pub fn femgpu_vec_axpby(ctx: *mut femgpu_ctx, dtype: femgpu_dtype, n: u64, alpha: f64, x: *const c_void,
                            beta: f64, y: *mut c_void) -> c_int;
    pub fn femgpu_vec_convert(ctx: *mut femgpu_ctx, n: u64, src_dtype: femgpu_dtype, src: *const c_void,
                              dst_dtype: femgpu_dtype, dst: *mut c_void) -> c_int;
    pub fn femgpu_timer_start(ctx: *mut femgpu_ctx) -> c_int;
    pub fn femgpu_timer_stop(ctx: *mut femgpu_ctx, ms: *mut f64) -> c_int;

    // -------------------------------------------------------- operators
    pub fn femgpu_op_apply(op: *mut femgpu_op, x: *const c_void, y: *mut c_void, accumulate: c_int,
                           mem: femgpu_mem) -> c_int;
    pub fn femgpu_op_n_dofs(op: *const femgpu_op, n: *mut u64) -> c_int;
    pub fn femgpu_op_dtype(op: *const femgpu_op, dtype: *mut femgpu_dtype) -> c_int;
    pub fn femgpu_op_diagonal(op: *mut femgpu_op, d: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_op_destroy(op: *mut femgpu_op) -> c_int;
    pub fn femgpu_op_time_apply(op: *mut femgpu_op, x: *const c_void, y: *mut c_void, reps: c_int,
                                ms_apply: *mut f64, ms_kernel: *mut f64) -> c_int;
    pub fn femgpu_sumfac_probe(ctx: *mut femgpu_ctx, p: c_int, q: c_int, dtype: femgpu_dtype, engine: c_int,
                               nblocks: c_int, reps: c_int, input: *const c_void, output: *mut c_void,
                               launches: c_int, ms: *mut f64) -> c_int;
    pub fn femgpu_op_create_callback(ctx: *mut femgpu_ctx, n: u64, dtype: femgpu_dtype, mem: femgpu_mem,
                                     apply: femgpu_apply_fn, user: *mut c_void, diag: *const c_void,
                                     out: *mut *mut femgpu_op) -> c_int;

    // -------------------------------------------------------------- CSR
    pub fn femgpu_csr_create(ctx: *mut femgpu_ctx, nrows: u64, ncols: u64, row_ptr: *const u64,
                             col_idx: *const u32, values: *const c_void, dtype: femgpu_dtype,
                             out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_csr_spmv_add(csr: *mut femgpu_op, alpha: f64, x: *const c_void, beta: f64,
                               y: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_csr_apply_dirichlet(csr: *mut femgpu_op, n_constrained: u64, dofs: *const u32,
                                      values: *const f64, rhs: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_csr_get_values(csr: *mut femgpu_op, values: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_csr_get_pattern(csr: *mut femgpu_op, nnz: *mut u64, row_ptr: *mut u64,
                                  col_idx: *mut u32, mem: femgpu_mem) -> c_int;

    // --------------------------------------------------------------- hex
    pub fn femgpu_pa_create_hex(ctx: *mut femgpu_ctx, desc: *const femgpu_hex_desc,
                                vertex_coords: *const f64, out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_hex_vertices(nx: c_int, ny: c_int, nz: c_int, amp: f64, seed: u64,
                               coords: *mut f64) -> c_int;
    pub fn femgpu_pa_hex_rhs_unit(op: *mut femgpu_op, b: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_pa_hex_rhs_random(op: *mut femgpu_op, seed: u64, b: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_pa_hex_qdata_per_point(op: *const femgpu_op, values: *mut c_int) -> c_int;
    pub fn femgpu_pa_hex_qdata_mode(op: *const femgpu_op, mode: *mut c_int) -> c_int;
    pub fn femgpu_pa_hex_element_kernel(op: *const femgpu_op, kind: *mut c_int) -> c_int;
    pub fn femgpu_pa_hex_convert(op: *mut femgpu_op, dtype: femgpu_dtype, out: *mut *mut femgpu_op) -> c_int;

    // ----------------------------------------------------------- simplex
    pub fn femgpu_pa_create_simplex(ctx: *mut femgpu_ctx, dim: c_int, order: c_int, n_nodes: u64,
                                    coords: *const f64, n_elem: u64, elem_nodes: *const u32,
                                    elem_dofs: *const u32, n_dofs: u64, form: c_int, coef: f64,
                                    quad_order: c_int, out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_pa_create_simplex_qp(ctx: *mut femgpu_ctx, dim: c_int, order: c_int, n_nodes: u64,
                                       coords: *const f64, n_elem: u64, elem_nodes: *const u32,
                                       elem_dofs: *const u32, n_dofs: u64, form: c_int,
                                       coef_qp: *const f64, quad_order: c_int,
                                       out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_simplex_qp_coords(dim: c_int, n_nodes: u64, coords: *const f64, n_elem: u64,
                                    elem_nodes: *const u32, quad_order: c_int, xq: *mut f64,
                                    nq: *mut u64) -> c_int;
    pub fn femgpu_csr_assemble_simplex(ctx: *mut femgpu_ctx, dim: c_int, order: c_int, n_nodes: u64,
                                       coords: *const f64, n_elem: u64, elem_nodes: *const u32,
                                       elem_dofs: *const u32, n_dofs: u64, form: c_int, coef: f64,
                                       quad_order: c_int, out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_csr_assemble_simplex_qp(ctx: *mut femgpu_ctx, dim: c_int, order: c_int, n_nodes: u64,
                                          coords: *const f64, n_elem: u64, elem_nodes: *const u32,
                                          elem_dofs: *const u32, n_dofs: u64, form: c_int,
                                          coef_qp: *const f64, quad_order: c_int,
                                          out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_rhs_assemble_simplex(ctx: *mut femgpu_ctx, dim: c_int, order: c_int, n_nodes: u64,
                                       coords: *const f64, n_elem: u64, elem_nodes: *const u32,
                                       elem_dofs: *const u32, n_dofs: u64, source: c_int, value: f64,
                                       quad_order: c_int, b: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_essential_create(inner: *mut femgpu_op, n_constrained: u64, dofs: *const u32,
                                   out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_essential_mask_rhs(op: *mut femgpu_op, b: *mut c_void, mem: femgpu_mem) -> c_int;
    pub fn femgpu_lumped_mass_create(ctx: *mut femgpu_ctx, dim: c_int, order: c_int, n_nodes: u64,
                                     coords: *const f64, n_elem: u64, elem_nodes: *const u32,
                                     elem_dofs: *const u32, n_dofs: u64, rho: f64, quad_order: c_int,
                                     out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_lumped_mass_create_qp(ctx: *mut femgpu_ctx, dim: c_int, order: c_int, n_nodes: u64,
                                        coords: *const f64, n_elem: u64, elem_nodes: *const u32,
                                        elem_dofs: *const u32, n_dofs: u64, rho_qp: *const f64,
                                        quad_order: c_int, out: *mut *mut femgpu_op) -> c_int;
    pub fn femgpu_lumped_mass_apply_inverse(op: *mut femgpu_op, x: *const c_void, y: *mut c_void,
                                            mem: femgpu_mem) -> c_int;

    // --------------------------------------------------------------- AMG
    pub fn femgpu_amg_create(csr: *mut femgpu_op, cfg: *const femgpu_amg_config,
                             out: *mut *mut femgpu_amg) -> c_int;
    pub fn femgpu_amg_destroy(amg: *mut femgpu_amg) -> c_int;
    pub fn femgpu_amg_n_levels(amg: *const femgpu_amg, n: *mut c_int) -> c_int;
    pub fn femgpu_amg_level(amg: *const femgpu_amg, level: c_int, which: c_int, nrows: *mut u64,
                            ncols: *mut u64, nnz: *mut u64, row_ptr: *mut u64, col_idx: *mut u32,
                            values: *mut f64) -> c_int;
    pub fn femgpu_amg_set_cycle(amg: *mut femgpu_amg, cycle: c_int) -> c_int;
    pub fn femgpu_amg_vcycle(amg: *mut femgpu_amg, b: *const c_void, x: *mut c_void, mem: femgpu_mem) -> c_int;

    // ----------------------------------------------------------- solvers
    pub fn femgpu_residual_history(ctx: *mut femgpu_ctx, out: *mut f64, cap: u64, len: *mut u64) -> c_int;
    pub fn femgpu_solve_cg_operator(op: *mut femgpu_op, b: *const c_void, x: *mut c_void, mem: femgpu_mem,
                                    cfg: *const femgpu_solver_config, res: *mut femgpu_solve_result) -> c_int;
    pub fn femgpu_solve_pcg_jacobi(op: *mut femgpu_op, b: *const c_void, x: *mut c_void, mem: femgpu_mem,
                                   cfg: *const femgpu_solver_config, res: *mut femgpu_solve_result) -> c_int;
    pub fn femgpu_solve_pcg_jacobi_noalloc(op: *mut femgpu_op, b: *const c_void, x: *mut c_void,
                                           cfg: *const femgpu_solver_config,
                                           res: *mut femgpu_solve_result) -> c_int;
    pub fn femgpu_solve_amg_cg(amg: *mut femgpu_amg, b: *const c_void, x: *mut c_void, mem: femgpu_mem,
                               cfg: *const femgpu_solver_config, res: *mut femgpu_solve_result) -> c_int;
    pub fn femgpu_solve_pcg_chebyshev(op: *mut femgpu_op, cheb: *const femgpu_cheb_config, b: *const c_void,
                                      x: *mut c_void, mem: femgpu_mem, cfg: *const femgpu_solver_config,
                                      res: *mut femgpu_solve_result, lmax_used: *mut f64) -> c_int;
    pub fn femgpu_chebyshev_smooth(op: *mut femgpu_op, x: *mut c_void, b: *const c_void, lambda_min: f64,
                                   lambda_max: f64, n_iter: c_int, mem: femgpu_mem) -> c_int;
    pub fn femgpu_estimate_spectral_radius(op: *mut femgpu_op, n_iter: c_int, rho: *mut f64) -> c_int;
    pub fn femgpu_solve_ir_mixed(op64: *mut femgpu_op, op32: *mut femgpu_op, ir: *const femgpu_ir_config,
                                 b: *const c_void, x: *mut c_void, mem: femgpu_mem,
                                 cfg: *const femgpu_solver_config, res: *mut femgpu_solve_result,
                                 inner_iterations: *mut u64) -> c_int;
    pub fn femgpu_chebyshev_jacobi_apply(op: *mut femgpu_op, cheb: *const femgpu_cheb_config, r: *const c_void,
                                         z: *mut c_void, mem: femgpu_mem) -> c_int;

    // ------------------------------------------- learned preconditioner
    pub fn femgpu_unet_create(ctx: *mut femgpu_ctx, nx: c_int, ny: c_int, nz: c_int, seed: u64,
                              out: *mut *mut femgpu_unet) -> c_int;
    pub fn femgpu_unet_create_coarse(ctx: *mut femgpu_ctx, nx: c_int, ny: c_int, nz: c_int, coarsen: c_int,
                                     seed: u64, out: *mut *mut femgpu_unet) -> c_int;
    pub fn femgpu_unet_grid(net: *const femgpu_unet, nx: *mut c_int, ny: *mut c_int, nz: *mut c_int) -> c_int;
    pub fn femgpu_unet_destroy(net: *mut femgpu_unet) -> c_int;
    pub fn femgpu_unet_n_params(net: *const femgpu_unet, n: *mut u64) -> c_int;
    pub fn femgpu_unet_layer_info(net: *const femgpu_unet, layer: c_int, cin: *mut c_int, cout: *mut c_int,
                                  kpad: *mut c_int, woff: *mut u64, boff: *mut u64) -> c_int;
    pub fn femgpu_unet_get_params(net: *mut femgpu_unet, host: *mut f32) -> c_int;
    pub fn femgpu_unet_set_params(net: *mut femgpu_unet, host: *const f32) -> c_int;
    pub fn femgpu_unet_set_aux(net: *mut femgpu_unet, aux: *const f32, mem: femgpu_mem) -> c_int;
    pub fn femgpu_unet_aux_from_hex(net: *mut femgpu_unet, hex: *mut femgpu_op) -> c_int;
    pub fn femgpu_unet_forward(net: *mut femgpu_unet, x0: *const f32, out: *mut f32, mem: femgpu_mem) -> c_int;
    pub fn femgpu_unet_backward(net: *mut femgpu_unet, dout: *const f32, grads: *mut f32,
                                mem: femgpu_mem) -> c_int;
    pub fn femgpu_unet_train_step(net: *mut femgpu_unet, r: *const f32, e: *const f32, mem: femgpu_mem,
                                  lr: f32, loss: *mut f64) -> c_int;
    pub fn femgpu_unet_train_step_energy(net: *mut femgpu_unet, op: *mut femgpu_op, r: *const f32,
                                         e: *const f32, mem: femgpu_mem, lr: f32, loss: *mut f64) -> c_int;
    pub fn femgpu_unet_apply(net: *mut femgpu_unet, r: *const c_void, z: *mut c_void, dtype: femgpu_dtype,
                             mem: femgpu_mem) -> c_int;
    pub fn femgpu_unet_save(net: *mut femgpu_unet, path: *const c_char) -> c_int;
    pub fn femgpu_unet_load(net: *mut femgpu_unet, path: *const c_char) -> c_int;
    pub fn femgpu_solve_fcg_learned(op: *mut femgpu_op, net: *mut femgpu_unet, cheb: *const femgpu_cheb_config,
                                    b: *const c_void, x: *mut c_void, mem: femgpu_mem,
                                    cfg: *const femgpu_solver_config, res: *mut femgpu_solve_result,
                                    info: *mut femgpu_fcg_info) -> c_int;
    pub fn femgpu_solve_ir_learned(op64: *mut femgpu_op, op32: *mut femgpu_op, net: *mut femgpu_unet,
                                   ir: *const femgpu_ir_config, b: *const c_void, x: *mut c_void,
                                   mem: femgpu_mem, cfg: *const femgpu_solver_config,
                                   res: *mut femgpu_solve_result, inner_iterations: *mut u64,
                                   info: *mut femgpu_fcg_info) -> c_int;
}
