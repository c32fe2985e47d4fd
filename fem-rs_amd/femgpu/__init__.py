"""Python mirror of the fem-rs solve-path API over the libfemgpu C ABI.

Every name here follows the reference crate it stands in for (fem-rs, Rust):

  CsrMatrix.spmv / spmv_add / diagonal     linalg/src/csr.rs:140-252
  PAMassOperator, PADiffusionOperator      assembly/src/partial.rs:45-209
  LumpedMassOperator                       assembly/src/partial.rs:217-274
  SolverConfig, SolveResult, SolverError   solver/src/lib.rs:47-142
  solve_cg_operator, solve_pcg_jacobi      solver/src/lib.rs:198-297
  chebyshev_smooth, estimate_spectral_radius  amg/src/lib.rs:53-123

All arithmetic runs in the HIP kernels of ``libfemgpu.so`` (gfx950); this
module only marshals arguments through ctypes.  There is no CPU fallback:
importing fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfemgpu.so")

F64, F32 = 0, 1
HOST, DEVICE = 0, 1
MASS, DIFFUSION, DIFF_MASS = 1, 2, 3
COEF_CONST, COEF_SINPROD, COEF_GRF, COEF_QP_ARRAY = 0, 1, 2, 3

OK, ERR_INVALID_ARG, ERR_DIM_MISMATCH, ERR_NOT_CONVERGED, ERR_BREAKDOWN, ERR_HIP, \
    ERR_UNSUPPORTED, ERR_OOM = range(8)

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libfemgpu.so not found at {LIB_PATH}: build it with `make -C fem-rs_amd` "
        "(or __graft_entry__.build()); there is no CPU fallback")

lib = C.CDLL(LIB_PATH)

# GPU objects still alive at interpreter exit are left to the process teardown:
# calling into the HIP runtime while it shuts down is not safe.
import atexit  # noqa: E402

_EXITING = False


def _mark_exit():
    global _EXITING
    _EXITING = True


atexit.register(_mark_exit)

P = C.c_void_p
I = C.c_int
D = C.c_double
U64 = C.c_uint64
Z = C.c_size_t


class _Coef(C.Structure):
    _fields_ = [("kind", I), ("value", D), ("n_features", I), ("sigma", D), ("length", D),
                ("lo", D), ("hi", D), ("seed", U64), ("qp_values", P)]


class _HexDesc(C.Structure):
    _fields_ = [("p", I), ("q", I), ("nx", I), ("ny", I), ("nz", I), ("form", I), ("dtype", I),
                ("dirichlet", I), ("rho", D), ("kappa", _Coef)]


class _SolverConfig(C.Structure):
    _fields_ = [("rtol", D), ("atol", D), ("max_iter", U64), ("verbose", I), ("print_level", I)]


class _SolveResult(C.Structure):
    _fields_ = [("converged", I), ("iterations", U64), ("final_residual", D),
                ("final_residual_is_relative", I), ("error_estimate", D)]


class _ChebConfig(C.Structure):
    _fields_ = [("degree", I), ("ratio", D), ("safety", D), ("power_iters", I), ("lmax", D)]


class _IrConfig(C.Structure):
    _fields_ = [("inner_rtol", D), ("inner_max_iter", U64), ("max_outer", U64),
                ("cheb", _ChebConfig), ("mode", I), ("update_interval", I)]


class _AmgConfig(C.Structure):
    _fields_ = [("theta", D), ("max_levels", I), ("coarse_max", U64), ("smoother_degree", I),
                ("smoother_ratio", D)]


def _sig(name, args, res=I):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("femgpu_last_error", [], C.c_char_p)
_sig("femgpu_version", [], C.c_char_p)
_sig("femgpu_ctx_create", [I, C.POINTER(P)])
_sig("femgpu_ctx_destroy", [P])
_sig("femgpu_ctx_synchronize", [P])
_sig("femgpu_malloc", [P, Z, C.POINTER(P)])
_sig("femgpu_free", [P, P])
_sig("femgpu_memcpy", [P, P, P, Z, I])
_sig("femgpu_timer_start", [P])
_sig("femgpu_vec_axpby", [P, I, U64, D, P, D, P])
_sig("femgpu_vec_convert", [P, U64, I, P, I, P])
_sig("femgpu_timer_stop", [P, C.POINTER(D)])
_sig("femgpu_op_apply", [P, P, P, I, I])
_sig("femgpu_op_n_dofs", [P, C.POINTER(U64)])
_sig("femgpu_op_dtype", [P, C.POINTER(I)])
_sig("femgpu_op_diagonal", [P, P, I])
_sig("femgpu_op_destroy", [P])
_sig("femgpu_op_time_apply", [P, P, P, I, C.POINTER(D), C.POINTER(D)])
_sig("femgpu_sumfac_probe", [P, I, I, I, I, I, I, P, P, I, C.POINTER(D)])
_APPLY_FN = C.CFUNCTYPE(I, P, P, P, U64)
_sig("femgpu_op_create_callback", [P, U64, I, I, _APPLY_FN, P, P, C.POINTER(P)])
_sig("femgpu_residual_history", [P, P, U64, C.POINTER(U64)])
_sig("femgpu_csr_create", [P, U64, U64, P, P, P, I, C.POINTER(P)])
_sig("femgpu_csr_spmv_add", [P, D, P, D, P, I])
_sig("femgpu_csr_apply_dirichlet", [P, U64, P, P, P, I])
_sig("femgpu_csr_get_values", [P, P, I])
_sig("femgpu_csr_get_pattern", [P, C.POINTER(U64), P, P, I])
_sig("femgpu_csr_assemble_simplex", [P, I, I, U64, P, U64, P, P, U64, I, D, I, C.POINTER(P)])
_sig("femgpu_rhs_assemble_simplex", [P, I, I, U64, P, U64, P, P, U64, I, D, I, P, I])
_sig("femgpu_pa_create_hex", [P, C.POINTER(_HexDesc), P, C.POINTER(P)])
_sig("femgpu_hex_vertices", [I, I, I, D, U64, P])
_sig("femgpu_pa_hex_rhs_unit", [P, P, I])
_sig("femgpu_pa_hex_rhs_random", [P, U64, P, I])
_sig("femgpu_pa_hex_convert", [P, I, C.POINTER(P)])
_sig("femgpu_pa_hex_qdata_per_point", [P, C.POINTER(I)])
_sig("femgpu_pa_hex_qdata_mode", [P, C.POINTER(I)])
_sig("femgpu_pa_hex_element_kernel", [P, C.POINTER(I)])
_sig("femgpu_pa_create_simplex", [P, I, I, U64, P, U64, P, P, U64, I, D, I, C.POINTER(P)])
_sig("femgpu_pa_create_simplex_qp", [P, I, I, U64, P, U64, P, P, U64, I, P, I, C.POINTER(P)])
_sig("femgpu_csr_assemble_simplex_qp", [P, I, I, U64, P, U64, P, P, U64, I, P, I, C.POINTER(P)])
_sig("femgpu_lumped_mass_create_qp", [P, I, I, U64, P, U64, P, P, U64, P, I, C.POINTER(P)])
_sig("femgpu_simplex_qp_coords", [I, U64, P, U64, P, I, P, C.POINTER(U64)])
_sig("femgpu_lumped_mass_create", [P, I, I, U64, P, U64, P, P, U64, D, I, C.POINTER(P)])
_sig("femgpu_lumped_mass_apply_inverse", [P, P, P, I])
_sig("femgpu_essential_create", [P, U64, P, C.POINTER(P)])
_sig("femgpu_amg_create", [P, P, C.POINTER(P)])
_sig("femgpu_amg_destroy", [P])
_sig("femgpu_amg_n_levels", [P, C.POINTER(I)])
_sig("femgpu_amg_level", [P, I, I, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64), P, P, P])
_sig("femgpu_amg_vcycle", [P, P, P, I])
_sig("femgpu_amg_set_cycle", [P, I])
_sig("femgpu_solve_amg_cg", [P, P, P, I, P, P])
_sig("femgpu_essential_mask_rhs", [P, P, I])
_sig("femgpu_solve_cg_operator", [P, P, P, I, C.POINTER(_SolverConfig), C.POINTER(_SolveResult)])
_sig("femgpu_solve_pcg_jacobi", [P, P, P, I, C.POINTER(_SolverConfig), C.POINTER(_SolveResult)])
_sig("femgpu_solve_pcg_jacobi_noalloc", [P, P, P, C.POINTER(_SolverConfig), C.POINTER(_SolveResult)])
_sig("femgpu_solve_pcg_chebyshev", [P, C.POINTER(_ChebConfig), P, P, I, C.POINTER(_SolverConfig),
                                    C.POINTER(_SolveResult), C.POINTER(D)])
_sig("femgpu_chebyshev_smooth", [P, P, P, D, D, I, I])
_sig("femgpu_estimate_spectral_radius", [P, I, C.POINTER(D)])
_sig("femgpu_solve_ir_mixed", [P, P, C.POINTER(_IrConfig), P, P, I, C.POINTER(_SolverConfig),
                               C.POINTER(_SolveResult), C.POINTER(U64)])


class _FcgInfo(C.Structure):
    _fields_ = [("patience", I), ("truncation", I), ("fcg_iterations", U64), ("fallback_iterations", U64),
                ("fell_back", I), ("fallback_reason", I), ("net_corrections", U64)]


_sig("femgpu_chebyshev_jacobi_apply", [P, C.POINTER(_ChebConfig), P, P, I])
_sig("femgpu_unet_create", [P, I, I, I, U64, C.POINTER(P)])
_sig("femgpu_unet_destroy", [P])
_sig("femgpu_unet_n_params", [P, C.POINTER(U64)])
_sig("femgpu_unet_layer_info", [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(U64),
                                C.POINTER(U64)])
_sig("femgpu_unet_get_params", [P, P])
_sig("femgpu_unet_set_params", [P, P])
_sig("femgpu_unet_set_aux", [P, P, I])
_sig("femgpu_unet_create_coarse", [P, I, I, I, I, U64, C.POINTER(P)])
_sig("femgpu_unet_grid", [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)])
_sig("femgpu_unet_aux_from_hex", [P, P])
_sig("femgpu_unet_forward", [P, P, P, I])
_sig("femgpu_unet_backward", [P, P, P, I])
_sig("femgpu_unet_train_step", [P, P, P, I, C.c_float, C.POINTER(D)])
_sig("femgpu_unet_train_step_energy", [P, P, P, P, I, C.c_float, C.POINTER(D)])
_sig("femgpu_unet_apply", [P, P, P, I, I])
_sig("femgpu_unet_save", [P, C.c_char_p])
_sig("femgpu_unet_load", [P, C.c_char_p])
_sig("femgpu_solve_fcg_learned", [P, P, C.POINTER(_ChebConfig), P, P, I, C.POINTER(_SolverConfig),
                                  C.POINTER(_SolveResult), C.POINTER(_FcgInfo)])
_sig("femgpu_solve_ir_learned", [P, P, P, C.POINTER(_IrConfig), P, P, I, C.POINTER(_SolverConfig),
                                   C.POINTER(_SolveResult), C.POINTER(U64), C.POINTER(_FcgInfo)])


# ---------------------------------------------------------------- errors --
class FemGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[femgpu {code}] {msg}")
        self.code = code


class SolverError(FemGpuError):
    """solver/src/lib.rs:47-55"""


class ConvergenceFailed(SolverError):
    def __init__(self, max_iter, residual, result=None):
        RuntimeError.__init__(self, f"solver did not converge in {max_iter} iterations "
                                    f"(residual = {residual:.3e})")
        self.code = ERR_NOT_CONVERGED
        self.max_iter, self.residual, self.result = max_iter, residual, result


class DimensionMismatch(SolverError):
    def __init__(self, rows, cols, rhs):
        RuntimeError.__init__(self, f"dimension mismatch: matrix is {rows}x{cols}, rhs has length {rhs}")
        self.code = ERR_DIM_MISMATCH
        self.rows, self.cols, self.rhs = rows, cols, rhs


class Linger(SolverError):
    def __init__(self, msg):
        RuntimeError.__init__(self, f"linger error: {msg}")
        self.code = ERR_BREAKDOWN
        self.msg = msg


def _check(rc):
    if rc != OK:
        raise FemGpuError(rc, lib.femgpu_last_error().decode())


def version():
    return lib.femgpu_version().decode()


# --------------------------------------------------------------- context --
class Context:
    """One HIP stream + solver scratch on one GPU (femgpu_ctx)."""

    def __init__(self, device=0):
        h = P()
        _check(lib.femgpu_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def synchronize(self):
        _check(lib.femgpu_ctx_synchronize(self.h))

    def timer_start(self):
        _check(lib.femgpu_timer_start(self.h))

    def timer_stop(self):
        ms = D()
        _check(lib.femgpu_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            if not _EXITING:
                lib.femgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("FEMGPU_DEVICE", "0")))
    return _default_ctx


class DeviceArray:
    """A 1-D vector resident in HBM (for FEMGPU_DEVICE calls); zero-filled
    unless data is given."""

    def __init__(self, ctx, n, dtype=np.float64, data=None):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        h = P()
        _check(lib.femgpu_malloc(ctx.h, self.nbytes, C.byref(h)))
        self.ptr = h
        if data is not None:
            self.upload(data)

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        _check(lib.femgpu_memcpy(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes, 0))
        return self

    def download(self):
        out = np.empty(self.n, self.dtype)
        _check(lib.femgpu_memcpy(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def copy_from(self, other):
        _check(lib.femgpu_memcpy(self.ctx.h, self.ptr, other.ptr, self.nbytes, 2))

    def _dt(self):
        return F64 if self.dtype == np.float64 else F32

    def axpby(self, alpha, x, beta):
        """self = alpha x + beta self on the device (femgpu_vec_axpby); returns self."""
        assert x.n == self.n and x.dtype == self.dtype
        _check(lib.femgpu_vec_axpby(self.ctx.h, self._dt(), self.n, float(alpha), x.ptr, float(beta), self.ptr))
        return self

    def astype(self, dtype):
        """A new DeviceArray with the values converted (f64 <-> f32) on the device."""
        out = DeviceArray(self.ctx, self.n, dtype)
        _check(lib.femgpu_vec_convert(self.ctx.h, self.n, self._dt(), self.ptr, out._dt(), out.ptr))
        return out

    def view(self, offset, n):
        """A non-owning DeviceArray over entries [offset, offset + n) of this
        one, as a caller hands the library a sub-range of its own buffer (the
        view keeps this array alive and frees nothing)."""
        if offset < 0 or n < 0 or offset + n > self.n:
            raise ValueError(f"view [{offset}, {offset + n}) outside an array of {self.n}")
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.n, v.dtype = self.ctx, int(n), self.dtype
        v.nbytes = v.n * v.dtype.itemsize
        v.ptr = P(self.ptr.value + offset * self.dtype.itemsize)
        v._base = self
        return v

    def free(self):
        if getattr(self, "_base", None) is not None:  # a view owns nothing
            self.ptr = None
            return
        if getattr(self, "ptr", None):
            if not _EXITING:
                lib.femgpu_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _npdt(dtype):
    return np.float64 if dtype == F64 else np.float32


def _vec(a, dtype, n, what):
    if isinstance(a, DeviceArray):
        if a.n != n:
            raise DimensionMismatch(n, n, a.n)
        return a.ptr, DEVICE
    if a.dtype != _npdt(dtype) or not a.flags.c_contiguous:
        raise TypeError(f"{what}: expected contiguous {np.dtype(_npdt(dtype))} array")
    return a.ctypes.data, HOST


# ------------------------------------------------------------- operators --
class Operator:
    """A square device operator (femgpu_op)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle
        n = U64()
        _check(lib.femgpu_op_n_dofs(handle, C.byref(n)))
        dt = I()
        _check(lib.femgpu_op_dtype(handle, C.byref(dt)))
        self._n, self.dtype = n.value, dt.value

    def n_dofs(self):
        return self._n

    @property
    def nrows(self):
        return self._n

    @property
    def ncols(self):
        return self._n

    def _apply(self, x, y, accumulate):
        if not isinstance(x, DeviceArray) and len(x) != self._n:
            raise DimensionMismatch(self._n, self._n, len(x))
        if not isinstance(y, DeviceArray) and len(y) != self._n:
            raise DimensionMismatch(self._n, self._n, len(y))
        xp, mx = _vec(x, self.dtype, self._n, "x")
        yp, my = _vec(y, self.dtype, self._n, "y")
        if mx != my:
            raise TypeError("x and y must both be host arrays or both DeviceArrays")
        _check(lib.femgpu_op_apply(self.h, xp, yp, 1 if accumulate else 0, mx))

    def apply(self, x, y):
        """MatFreeOperator::apply: y += A x (partial.rs:45-47)."""
        self._apply(x, y, True)

    def apply_zero(self, x, y):
        """MatFreeOperator::apply_zero: y = A x (partial.rs:50-53)."""
        self._apply(x, y, False)

    def matvec(self, x):
        y = np.zeros(self._n, _npdt(self.dtype))
        self.apply_zero(np.ascontiguousarray(x, _npdt(self.dtype)), y)
        return y

    def diagonal(self):
        d = np.zeros(self._n, _npdt(self.dtype))
        _check(lib.femgpu_op_diagonal(self.h, d.ctypes.data, HOST))
        return d

    def time_apply(self, x, y, reps=20):
        """(ms per apply, ms per main compute kernel) over reps device applies."""
        a, k = D(), D()
        _check(lib.femgpu_op_time_apply(self.h, x.ptr, y.ptr, reps, C.byref(a), C.byref(k)))
        return a.value, k.value

    def close(self):
        if getattr(self, "h", None):
            if not _EXITING:
                lib.femgpu_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CallbackOperator(Operator):
    """An operator whose apply is a Python function fn(x, y) setting y = A x
    (the `apply` closure of solve_cg_operator, solver/src/lib.rs:198-208).
    mem=HOST (default): x and y are numpy views of the library's host staging
    buffers; mem=DEVICE: fn(x_ptr, y_ptr) gets device addresses (ints) on the
    context's GPU.  diag: optional diagonal (for Jacobi / Chebyshev).  An
    exception in fn aborts the solve with FemGpuError."""

    def __init__(self, n, fn, dtype=F64, mem=HOST, diag=None, ctx=None):
        ctx = ctx or default_context()
        npt = _npdt(dtype)
        self._fn, self._mem, self.error = fn, mem, None

        def tramp(user, xp, yp, m):
            try:
                if mem == HOST:
                    x = np.ctypeslib.as_array(C.cast(xp, C.POINTER(C.c_double if npt == np.float64 else C.c_float)),
                                              shape=(m,))
                    y = np.ctypeslib.as_array(C.cast(yp, C.POINTER(C.c_double if npt == np.float64 else C.c_float)),
                                              shape=(m,))
                    fn(x, y)
                else:
                    fn(xp, yp)
                return 0
            except Exception as e:  # reported by the failing solve
                self.error = e
                return 1

        self._cb = _APPLY_FN(tramp)  # kept alive with the operator
        d = None if diag is None else np.ascontiguousarray(diag, npt)
        h = P()
        _check(lib.femgpu_op_create_callback(ctx.h, int(n), dtype, mem, self._cb, None,
                                             None if d is None else d.ctypes.data, C.byref(h)))
        super().__init__(ctx, h)


def residual_history(ctx=None):
    """Residuals of the last solve on ctx that recorded them (SolverConfig
    with verbose=True or print_level >= 2): entry 0 is the initial residual."""
    ctx = ctx or default_context()
    ln = U64()
    _check(lib.femgpu_residual_history(ctx.h, None, 0, C.byref(ln)))
    out = np.zeros(ln.value)
    _check(lib.femgpu_residual_history(ctx.h, out.ctypes.data if ln.value else None, ln.value, C.byref(ln)))
    return out


class CsrMatrix(Operator):
    """CsrMatrix<T> (linalg/src/csr.rs:81-93) resident on the GPU."""

    def __init__(self, row_ptr, col_idx, values, ncols=None, dtype=F64, ctx=None):
        ctx = ctx or default_context()
        rp = np.ascontiguousarray(row_ptr, np.uint64)
        ci = np.ascontiguousarray(col_idx, np.uint32)
        v = np.ascontiguousarray(values, _npdt(dtype))
        nrows = len(rp) - 1
        ncols = nrows if ncols is None else ncols
        h = P()
        _check(lib.femgpu_csr_create(ctx.h, nrows, ncols, rp.ctypes.data, ci.ctypes.data,
                                     v.ctypes.data, dtype, C.byref(h)))
        super().__init__(ctx, h)
        self._nnz = int(rp[-1])

    def nnz(self):
        return self._nnz

    def spmv(self, x, y):
        """y = A x (csr.rs:140-164); a length mismatch raises (the reference panics)."""
        self.apply_zero(x, y)

    def spmv_add(self, alpha, x, beta, y):
        """y = alpha A x + beta y (csr.rs:167-193)."""
        if not isinstance(x, DeviceArray) and (len(x) != self._n or len(y) != self._n):
            raise DimensionMismatch(self._n, self._n, len(x))
        xp, m = _vec(x, self.dtype, self._n, "x")
        yp, _ = _vec(y, self.dtype, self._n, "y")
        _check(lib.femgpu_csr_spmv_add(self.h, alpha, xp, beta, yp, m))

    def apply_dirichlet_row_zeroing(self, dofs, values, rhs=None):
        """apply_dirichlet_row_zeroing (csr.rs:297-312) for every (dof, value)
        in order, on the device copy; rhs (host array or DeviceArray) updated."""
        d = np.ascontiguousarray(dofs, np.uint32)
        v = np.ascontiguousarray(values, np.float64)
        if d.shape != v.shape:
            raise ValueError("constrained_dofs and values must have the same length")
        rp, m = (None, HOST) if rhs is None else _vec(rhs, self.dtype, self._n, "rhs")
        _check(lib.femgpu_csr_apply_dirichlet(self.h, len(d), d.ctypes.data, v.ctypes.data, rp, m))

    def values(self):
        """The nnz matrix values (CsrMatrix::values) as a host array."""
        out = np.empty(self._nnz, _npdt(self.dtype))
        _check(lib.femgpu_csr_get_values(self.h, out.ctypes.data, HOST))
        return out

    def pattern(self):
        """(row_ptr u64[n+1], col_idx u32[nnz]) as host arrays (csr.rs:81-93)."""
        nnz = U64()
        rp = np.empty(self._n + 1, np.uint64)
        ci = np.empty(self._nnz, np.uint32)
        _check(lib.femgpu_csr_get_pattern(self.h, C.byref(nnz), rp.ctypes.data, ci.ctypes.data, HOST))
        return rp, ci

    @classmethod
    def _adopt(cls, ctx, h):
        m = cls.__new__(cls)
        Operator.__init__(m, ctx, h)
        nnz = U64()
        _check(lib.femgpu_csr_get_pattern(h, C.byref(nnz), None, None, HOST))
        m._nnz = int(nnz.value)
        return m


def _mesh_args(dim, order, coords, elem_nodes, elem_dofs):
    coords = np.ascontiguousarray(coords, np.float64).ravel()
    nodes = np.ascontiguousarray(elem_nodes, np.uint32).ravel()
    dofs = np.ascontiguousarray(elem_dofs, np.uint32).ravel()
    ne = len(nodes) // (dim + 1)
    return coords, nodes, dofs, ne


def assemble_bilinear(dim, order, coords, elem_nodes, elem_dofs, n_dofs, form=DIFFUSION, coef=1.0,
                      quad_order=3, ctx=None):
    """Assembler::assemble_bilinear (assembly/src/assembler.rs:147-249) with the
    diffusion or mass integrator, then CooMatrix::into_csr, on the GPU;
    returns the device CsrMatrix.  coef: a constant, an array (n_elem, nq) of
    point values, or a function of the physical quadrature points."""
    ctx = ctx or default_context()
    q = _coef_qp(coef, dim, coords, elem_nodes, quad_order)
    coords, nodes, dofs, ne = _mesh_args(dim, order, coords, elem_nodes, elem_dofs)
    h = P()
    if q is None:
        _check(lib.femgpu_csr_assemble_simplex(ctx.h, dim, order, len(coords) // dim, coords.ctypes.data, ne,
                                               nodes.ctypes.data, dofs.ctypes.data, n_dofs, form, coef,
                                               quad_order, C.byref(h)))
    else:
        _check(lib.femgpu_csr_assemble_simplex_qp(ctx.h, dim, order, len(coords) // dim, coords.ctypes.data,
                                                  ne, nodes.ctypes.data, dofs.ctypes.data, n_dofs, form,
                                                  q.ctypes.data, quad_order, C.byref(h)))
    return CsrMatrix._adopt(ctx, h)


SOURCE_CONST, SOURCE_SINPROD = 0, 1


def assemble_linear(dim, order, coords, elem_nodes, elem_dofs, n_dofs, source=SOURCE_CONST, value=1.0,
                    quad_order=3, out=None, ctx=None):
    """Assembler::assemble_linear with DomainSourceIntegrator
    (assembly/src/standard/source.rs:44-50) on the GPU: f = value, or
    value * prod_d sin(pi x_d) for SOURCE_SINPROD.  out: host array (default,
    returned) or DeviceArray."""
    ctx = ctx or default_context()
    coords, nodes, dofs, ne = _mesh_args(dim, order, coords, elem_nodes, elem_dofs)
    if out is None:
        out = np.zeros(n_dofs)
    ptr, m = _vec(out, F64, n_dofs, "b")
    _check(lib.femgpu_rhs_assemble_simplex(ctx.h, dim, order, len(coords) // dim, coords.ctypes.data, ne,
                                           nodes.ctypes.data, dofs.ctypes.data, n_dofs, source, value,
                                           quad_order, ptr, m))
    return out


def apply_dirichlet(mat, rhs, constrained_dofs, values):
    """apply_dirichlet (space/src/constraints.rs:31-42): row zeroing of the
    constrained rows of `mat`, rhs set to the prescribed values."""
    if len(constrained_dofs) != len(values):
        raise ValueError("constrained_dofs and values must have the same length")
    mat.apply_dirichlet_row_zeroing(constrained_dofs, values, rhs)


def _coef_struct(kappa, qp_keep):
    if isinstance(kappa, _Coef):
        return kappa
    if kappa is None:
        return _Coef(COEF_CONST, 1.0, 0, 0.0, 0.0, 0.0, 0.0, 0, None)
    if isinstance(kappa, str):
        if kappa != "sinprod":
            raise ValueError(f"unknown coefficient {kappa!r}")
        return _Coef(COEF_SINPROD, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0, None)
    if np.isscalar(kappa):
        return _Coef(COEF_CONST, float(kappa), 0, 0.0, 0.0, 0.0, 0.0, 0, None)
    if isinstance(kappa, dict) and kappa.get("kind") == "grf":
        return _Coef(COEF_GRF, 0.0, int(kappa.get("n_features", 256)), float(kappa.get("sigma", 1.5)),
                     float(kappa.get("length", 0.1)), float(kappa.get("lo", 1e-2)),
                     float(kappa.get("hi", 1e2)), int(kappa.get("seed", 0)), None)
    arr = np.ascontiguousarray(kappa, np.float64)
    qp_keep.append(arr)
    return _Coef(COEF_QP_ARRAY, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0, arr.ctypes.data)


def grf(seed, n_features=256, sigma=1.5, length=0.1, lo=1e-2, hi=1e2):
    """Gaussian-random-field coefficient kappa = clamp(exp(g), lo, hi)."""
    return dict(kind="grf", seed=seed, n_features=n_features, sigma=sigma, length=length, lo=lo, hi=hi)


def hex_vertices(nx, ny, nz, amp=0.0, seed=0):
    c = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    _check(lib.femgpu_hex_vertices(nx, ny, nz, amp, seed, c.ctypes.data))
    return c


class HexPAOperator(Operator):
    """Sum-factorised PA mass / diffusion / diffusion+mass on a structured hex
    Q_p mesh (the tensor-product form of PADiffusionOperator / PAMassOperator)."""

    def __init__(self, p, q, n, form=DIFFUSION, kappa=1.0, rho=1.0, dirichlet=False, dtype=F64,
                 vertices=None, ctx=None, _handle=None):
        ctx = ctx or default_context()
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        self.p, self.q, self.shape, self.form = p, q, (nx, ny, nz), form
        if _handle is None:
            keep = []
            desc = _HexDesc(p, q, nx, ny, nz, form, dtype, 1 if dirichlet else 0, float(rho),
                            _coef_struct(kappa, keep))
            v = None if vertices is None else np.ascontiguousarray(vertices, np.float64)
            h = P()
            _check(lib.femgpu_pa_create_hex(ctx.h, C.byref(desc), None if v is None else v.ctypes.data,
                                             C.byref(h)))
            _handle = h
        super().__init__(ctx, _handle)

    @property
    def n_elements(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    def qdata_per_point(self):
        """Quadrature-data values stored per Gauss point: 1 (w kappa) with
        the geometry formed on the fly (diffusion on a mesh that is not an
        axis-aligned box, the default), 1 (+ 1 with the mass term) on an
        axis-aligned box mesh with per-axis factors, 6 (+ 1) for the stored
        general form (FEMGPU_HEX_QDATA=general at creation)."""
        v = I()
        _check(lib.femgpu_pa_hex_qdata_per_point(self.h, C.byref(v)))
        return v.value

    def qdata_mode(self):
        """"general" (D = w kappa |J| J^-1 J^-T per point; mass-only
        operators, or FEMGPU_HEX_QDATA=general at creation), "cartesian"
        (axis-aligned box mesh) or "geometry" (w kappa per point, J formed
        from the vertices by every apply; the default elsewhere)."""
        v = I()
        _check(lib.femgpu_pa_hex_qdata_mode(self.h, C.byref(v)))
        return {0: "general", 1: "cartesian", 2: "geometry"}[v.value]

    def element_kernel(self):
        """"workgroup" (LDS hand-offs between phases), "thread-per-element"
        (Q2 diffusion with Cartesian qdata) or "quad-per-element" (Q3 diffusion,
        4^3 points, Cartesian qdata)."""
        v = I()
        _check(lib.femgpu_pa_hex_element_kernel(self.h, C.byref(v)))
        return {1: "thread-per-element", 2: "quad-per-element"}.get(v.value, "workgroup")

    @property
    def cartesian(self):
        return bool(self.form & DIFFUSION) and self.qdata_mode() == "cartesian"

    def convert(self, dtype):
        h = P()
        _check(lib.femgpu_pa_hex_convert(self.h, dtype, C.byref(h)))
        return HexPAOperator(self.p, self.q, self.shape, self.form, ctx=self.ctx, _handle=h)

    def rhs_unit(self, out=None):
        out = np.zeros(self._n, _npdt(self.dtype)) if out is None else out
        ptr, m = _vec(out, self.dtype, self._n, "b")
        _check(lib.femgpu_pa_hex_rhs_unit(self.h, ptr, m))
        return out

    def rhs_random(self, seed, out=None):
        out = np.zeros(self._n, _npdt(self.dtype)) if out is None else out
        ptr, m = _vec(out, self.dtype, self._n, "b")
        _check(lib.femgpu_pa_hex_rhs_random(self.h, seed, ptr, m))
        return out


def _simplex_create(fn, ctx, dim, order, coords, elem_nodes, elem_dofs, n_dofs, *rest):
    coords = np.ascontiguousarray(coords, np.float64)
    en = np.ascontiguousarray(elem_nodes, np.uint32)
    ed = np.ascontiguousarray(elem_dofs, np.uint32)
    h = P()
    _check(fn(ctx.h, dim, order, len(coords) // dim, coords.ctypes.data, len(en) // (dim + 1),
              en.ctypes.data, ed.ctypes.data, n_dofs, *rest, C.byref(h)))
    return h


def simplex_qp_coords(dim, coords, elem_nodes, quad_order):
    """Physical quadrature points (n_elem, nq, dim): x = x_0 + J xi, the points
    where ScalarCoeff::eval sees the coefficient (coefficient.rs:67-132)."""
    coords = np.ascontiguousarray(coords, np.float64)
    en = np.ascontiguousarray(elem_nodes, np.uint32)
    ne = len(en) // (dim + 1)
    nq = U64()
    _check(lib.femgpu_simplex_qp_coords(dim, len(coords) // dim, coords.ctypes.data, ne, en.ctypes.data,
                                        quad_order, None, C.byref(nq)))
    xq = np.zeros((ne, nq.value, dim))
    _check(lib.femgpu_simplex_qp_coords(dim, len(coords) // dim, coords.ctypes.data, ne, en.ctypes.data,
                                        quad_order, xq.ctypes.data, C.byref(nq)))
    return xq


def _coef_qp(coef, dim, coords, elem_nodes, quad_order):
    """None for a constant coefficient; else its values (n_elem * nq, f64) at
    the quadrature points: coef is an array (n_elem, nq) or a function of the
    physical points (n_elem, nq, dim) -> (n_elem, nq), like a ScalarCoeff."""
    if np.isscalar(coef):
        return None
    if callable(coef):
        coef = coef(simplex_qp_coords(dim, coords, elem_nodes, quad_order))
    return np.ascontiguousarray(np.ravel(coef), np.float64)


class PAMassOperator(Operator):
    """PAMassOperator::new(space, rho, quad_order) (partial.rs:65-131), simplex P1..P3.
    rho: a constant, an array (n_elem, nq) of point values, or a function of
    the physical quadrature points (evaluated once, at set-up)."""

    def __init__(self, dim, order, coords, elem_nodes, elem_dofs, n_dofs, rho=1.0, quad_order=3,
                 ctx=None):
        ctx = ctx or default_context()
        q = _coef_qp(rho, dim, coords, elem_nodes, quad_order)
        if q is None:
            h = _simplex_create(lib.femgpu_pa_create_simplex, ctx, dim, order, coords, elem_nodes,
                                elem_dofs, n_dofs, MASS, float(rho), quad_order)
        else:
            h = _simplex_create(lib.femgpu_pa_create_simplex_qp, ctx, dim, order, coords, elem_nodes,
                                elem_dofs, n_dofs, MASS, q.ctypes.data, quad_order)
        super().__init__(ctx, h)


class PADiffusionOperator(Operator):
    """PADiffusionOperator::new(space, kappa, quad_order) (partial.rs:138-209), simplex P1..P3.
    kappa: a constant, an array (n_elem, nq) of point values, or a function
    of the physical quadrature points (partial.rs:137-147, evaluated once)."""

    def __init__(self, dim, order, coords, elem_nodes, elem_dofs, n_dofs, kappa=1.0, quad_order=3,
                 ctx=None):
        ctx = ctx or default_context()
        q = _coef_qp(kappa, dim, coords, elem_nodes, quad_order)
        if q is None:
            h = _simplex_create(lib.femgpu_pa_create_simplex, ctx, dim, order, coords, elem_nodes,
                                elem_dofs, n_dofs, DIFFUSION, float(kappa), quad_order)
        else:
            h = _simplex_create(lib.femgpu_pa_create_simplex_qp, ctx, dim, order, coords, elem_nodes,
                                elem_dofs, n_dofs, DIFFUSION, q.ctypes.data, quad_order)
        super().__init__(ctx, h)


class EssentialBcOperator(Operator):
    """The matrix-free essential-BC operator of examples/src/maxwell.rs:142-197:
    A applied to x with the constrained entries zeroed, identity rows on the
    constrained DOFs.  Keeps a reference to `inner`."""

    def __init__(self, inner, constrained_dofs):
        d = np.ascontiguousarray(constrained_dofs, np.uint32)
        h = P()
        _check(lib.femgpu_essential_create(inner.h, len(d), d.ctypes.data, C.byref(h)))
        super().__init__(inner.ctx, h)
        self.inner = inner

    def mask_rhs(self, b):
        """b with its constrained entries set to zero (in place; host array or DeviceArray)."""
        ptr, m = _vec(b, self.dtype, self._n, "b")
        _check(lib.femgpu_essential_mask_rhs(self.h, ptr, m))
        return b


class LumpedMassOperator(Operator):
    """LumpedMassOperator::assemble(space, rho, quad_order) (partial.rs:217-274)."""

    def __init__(self, dim, order, coords, elem_nodes, elem_dofs, n_dofs, rho=1.0, quad_order=3,
                 ctx=None):
        ctx = ctx or default_context()
        q = _coef_qp(rho, dim, coords, elem_nodes, quad_order)
        if q is None:
            h = _simplex_create(lib.femgpu_lumped_mass_create, ctx, dim, order, coords, elem_nodes,
                                elem_dofs, n_dofs, float(rho), quad_order)
        else:
            h = _simplex_create(lib.femgpu_lumped_mass_create_qp, ctx, dim, order, coords, elem_nodes,
                                elem_dofs, n_dofs, q.ctypes.data, quad_order)
        super().__init__(ctx, h)

    @property
    def diag(self):
        return self.diagonal()

    def apply_inverse(self, x, y):
        xp, m = _vec(x, self.dtype, self._n, "x")
        yp, _ = _vec(y, self.dtype, self._n, "y")
        _check(lib.femgpu_lumped_mass_apply_inverse(self.h, xp, yp, m))


# --------------------------------------------------------------- solvers --
@dataclass
class SolverConfig:
    """solver/src/lib.rs:97-113"""
    rtol: float = 1e-8
    atol: float = 0.0
    max_iter: int = 1000
    verbose: bool = False
    print_level: int = 0

    def _c(self):
        return _SolverConfig(self.rtol, self.atol, self.max_iter, 1 if self.verbose else 0,
                             self.print_level)


@dataclass
class SolveResult:
    """solver/src/lib.rs:71-75"""
    converged: bool
    iterations: int
    final_residual: float


def _result(r):
    return SolveResult(bool(r.converged), int(r.iterations), float(r.final_residual))


def _solve_vectors(op, b, x):
    bp, mb = _vec(b, op.dtype, op.n_dofs(), "b")
    xp, mx = _vec(x, op.dtype, op.n_dofs(), "x")
    if mb != mx:
        raise TypeError("b and x must both be host arrays or both DeviceArrays")
    return bp, xp, mb


def _len(v):
    return v.n if isinstance(v, DeviceArray) else len(v)


SUMFAC_ENGINES = ("valu", "valu_even_odd", "mfma")


def sumfac_probe(p, q, dtype, engine, nblocks, reps, x_in, y_out, launches=5, ctx=None):
    """Time one batched sum-factorisation sweep (femgpu_sumfac_probe): 256
    lines of p + 1 values per workgroup, Out = [B u | G u] at q Gauss points,
    reps times per launch, on the vector unit ("valu", "valu_even_odd") or the
    matrix cores ("mfma").  x_in: nblocks * 256 * (p + 1) device values,
    y_out: nblocks * 256 * 2q.  Returns ms per launch."""
    ctx = ctx or default_context()
    eng = SUMFAC_ENGINES.index(engine)
    if x_in.n != nblocks * 256 * (p + 1) or y_out.n != nblocks * 256 * 2 * q:
        raise ValueError(f"sumfac_probe: sizes {x_in.n}, {y_out.n} do not fit {nblocks} blocks")
    ms = D()
    _check(lib.femgpu_sumfac_probe(ctx.h, p, q, dtype, eng, nblocks, reps, x_in.ptr, y_out.ptr, launches,
                                   C.byref(ms)))
    return ms.value


def solve_cg_operator(nrows, ncols, op, b, x, cfg=None):
    """solve_cg_operator (solver/src/lib.rs:198-278): x is the initial guess
    and is overwritten with the solution; returns SolveResult or raises
    ConvergenceFailed / Linger (breakdown) / DimensionMismatch.  op is a
    device operator or, as the reference's `apply: F`, a function
    apply(x, y) setting y = A x on numpy arrays."""
    cfg = cfg or SolverConfig()
    if nrows != ncols or _len(b) != nrows or _len(x) != ncols:
        raise DimensionMismatch(nrows, ncols, _len(b))
    if not isinstance(op, Operator):
        ctx = b.ctx if isinstance(b, DeviceArray) else None
        cop = CallbackOperator(nrows, op, ctx=ctx)
        try:
            return solve_cg_operator(nrows, ncols, cop, b, x, cfg)
        except FemGpuError:
            if cop.error is not None:
                raise cop.error
            raise
        finally:
            cop.close()
    if op.n_dofs() != nrows:
        raise DimensionMismatch(nrows, ncols, _len(b))
    bp, xp, m = _solve_vectors(op, b, x)
    r = _SolveResult()
    rc = lib.femgpu_solve_cg_operator(op.h, bp, xp, m, C.byref(cfg._c()), C.byref(r))
    if rc == ERR_NOT_CONVERGED:
        raise ConvergenceFailed(cfg.max_iter, r.final_residual, _result(r))
    if rc == ERR_BREAKDOWN:
        raise Linger("CG operator breakdown: p^T A p is near zero")
    _check(rc)
    return _result(r)


def _check_dims(a, b, x):
    # check_dims (solver/src/lib.rs:1074-1083)
    if a.nrows != _len(b) or a.ncols != _len(x):
        raise DimensionMismatch(a.nrows, a.ncols, _len(b))


def solve_pcg_jacobi(a, b, x, cfg=None):
    """solve_pcg_jacobi (solver/src/lib.rs:281-297) with par_solve_pcg_jacobi
    semantics (parallel/src/par_solver.rs:262-346)."""
    cfg = cfg or SolverConfig()
    _check_dims(a, b, x)
    bp, xp, m = _solve_vectors(a, b, x)
    r = _SolveResult()
    rc = lib.femgpu_solve_pcg_jacobi(a.h, bp, xp, m, C.byref(cfg._c()), C.byref(r))
    if rc == ERR_NOT_CONVERGED:
        return _result(r)  # par_solve_pcg_jacobi: Ok(converged = false)
    _check(rc)
    return _result(r)


def solve_pcg_jacobi_noalloc(a, b, x, cfg=None):
    """solve_pcg_jacobi for callers that cannot allocate: same result as
    solve_pcg_jacobi, with the work vectors in the library's fixed per-GPU
    scratch array (no device allocation in the call).  a is an assembled CSR
    matrix; b and x are DeviceArrays; at most 838,848 unknowns in FP64."""
    cfg = cfg or SolverConfig()
    _check_dims(a, b, x)
    if not (isinstance(b, DeviceArray) and isinstance(x, DeviceArray)):
        raise TypeError("solve_pcg_jacobi_noalloc takes DeviceArrays only")
    r = _SolveResult()
    rc = lib.femgpu_solve_pcg_jacobi_noalloc(a.h, b.ptr, x.ptr, C.byref(cfg._c()), C.byref(r))
    if rc == ERR_NOT_CONVERGED:
        return _result(r)
    _check(rc)
    return _result(r)


CYCLE_V, CYCLE_W, CYCLE_F = 0, 1, 2


class AmgSolver:
    """AmgSolver (amg/src/lib.rs:157-207): a smoothed-aggregation hierarchy
    built once (host) for the device CsrMatrix `a`, then PCG with one cycle
    (V by default, see with_cycle) per iteration on the device.  Keeps a
    reference to `a`."""

    def __init__(self, a, theta=0.08, max_levels=10, coarse_max=50, smoother_degree=2, smoother_ratio=10.0):
        cfg = _AmgConfig(theta, max_levels, coarse_max, smoother_degree, smoother_ratio)
        h = P()
        _check(lib.femgpu_amg_create(a.h, C.byref(cfg), C.byref(h)))
        self.h, self.a, self.ctx = h, a, a.ctx

    @classmethod
    def setup(cls, a, **config):
        return cls(a, **config)

    def with_cycle(self, cycle):
        """AmgSolver::with_cycle (amg/src/lib.rs:176-179): "V", "W" or "F"
        (or CYCLE_V / CYCLE_W / CYCLE_F); returns self."""
        c = {"V": CYCLE_V, "W": CYCLE_W, "F": CYCLE_F}.get(cycle, cycle) if isinstance(cycle, str) else cycle
        _check(lib.femgpu_amg_set_cycle(self.h, int(c) if isinstance(c, int) else -1))
        return self

    def n_levels(self):
        n = I()
        _check(lib.femgpu_amg_n_levels(self.h, C.byref(n)))
        return n.value

    def level(self, level, which="A"):
        """(row_ptr, col_idx, values, (nrows, ncols)) of a level's operator
        ("A") or prolongator to the next level ("P")."""
        w = 0 if which == "A" else 1
        nr, nc, nz = U64(), U64(), U64()
        _check(lib.femgpu_amg_level(self.h, level, w, C.byref(nr), C.byref(nc), C.byref(nz), None, None, None))
        rp = np.empty(nr.value + 1, np.uint64)
        ci = np.empty(nz.value, np.uint32)
        v = np.empty(nz.value)
        _check(lib.femgpu_amg_level(self.h, level, w, C.byref(nr), C.byref(nc), C.byref(nz), rp.ctypes.data,
                                    ci.ctypes.data, v.ctypes.data))
        return rp, ci, v, (nr.value, nc.value)

    def vcycle(self, b, x=None):
        """x = one cycle (V unless with_cycle chose another) applied to b
        (host arrays or DeviceArrays)."""
        n = self.a.n_dofs()
        if x is None:
            x = np.zeros(n)
        bp, m = _vec(b, F64, n, "b")
        xp, m2 = _vec(x, F64, n, "x")
        if m != m2:
            raise TypeError("b and x must both be host arrays or both DeviceArrays")
        _check(lib.femgpu_amg_vcycle(self.h, bp, xp, m))
        return x

    def solve(self, b, x, cfg=None):
        """PCG with the cycle as preconditioner; stopping rule of
        par_solve_pcg_jacobi (converged = False at max_iter)."""
        cfg = cfg or SolverConfig()
        _check_dims(self.a, b, x)
        bp, xp, m = _solve_vectors(self.a, b, x)
        r = _SolveResult()
        rc = lib.femgpu_solve_amg_cg(self.h, bp, xp, m, C.byref(cfg._c()), C.byref(r))
        if rc not in (OK, ERR_NOT_CONVERGED):
            _check(rc)
        return _result(r)

    def close(self):
        if getattr(self, "h", None):
            if not _EXITING:
                lib.femgpu_amg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_amg_cg(a, b, x, cfg=None, **amg_config):
    """solve_amg_cg (amg/src/lib.rs:137-155): build the hierarchy, then AMG-PCG."""
    s = AmgSolver(a, **amg_config)
    try:
        return s.solve(b, x, cfg)
    finally:
        s.close()


def solve_pcg_chebyshev(a, b, x, cfg=None, degree=4, ratio=30.0, safety=1.1, power_iters=20,
                        lmax=0.0):
    """PCG with a textbook Chebyshev-Jacobi preconditioner; returns
    (SolveResult, lmax used)."""
    cfg = cfg or SolverConfig()
    _check_dims(a, b, x)
    bp, xp, m = _solve_vectors(a, b, x)
    r = _SolveResult()
    lm = D()
    cc = _ChebConfig(degree, ratio, safety, power_iters, lmax)
    rc = lib.femgpu_solve_pcg_chebyshev(a.h, C.byref(cc), bp, xp, m, C.byref(cfg._c()),
                                        C.byref(r), C.byref(lm))
    if rc not in (OK, ERR_NOT_CONVERGED):
        _check(rc)
    return _result(r), lm.value


def chebyshev_smooth(a, x, b, lambda_min, lambda_max, n_iter):
    """chebyshev_smooth (amg/src/lib.rs:53-103); x updated in place."""
    xp, m = _vec(x, a.dtype, a.n_dofs(), "x")
    bp, _ = _vec(b, a.dtype, a.n_dofs(), "b")
    _check(lib.femgpu_chebyshev_smooth(a.h, xp, bp, lambda_min, lambda_max, n_iter, m))


def estimate_spectral_radius(a, n_iter):
    """estimate_spectral_radius (amg/src/lib.rs:108-123)."""
    rho = D()
    _check(lib.femgpu_estimate_spectral_radius(a.h, n_iter, C.byref(rho)))
    return rho.value


def solve_ir_mixed(op64, op32, b, x, cfg=None, inner_rtol=None, inner_max_iter=0, max_outer=50,
                   degree=4, ratio=30.0, safety=1.1, power_iters=20, lmax=0.0, mode=1, update_interval=16):
    """FP32-inner / FP64-outer iterative refinement; returns (SolveResult with
    outer iterations / residual replacements, total inner iterations).
    mode 1 (default): one FP32 PCG whose residual is replaced by the FP64
    residual each time it has dropped by inner_rtol (default 1e-2), at most
    update_interval iterations apart; mode 0: restarted inner solves to
    inner_rtol (default 1e-4)."""
    cfg = cfg or SolverConfig()
    if inner_rtol is None:
        inner_rtol = 1e-2 if mode == 1 else 1e-4
    _check_dims(op64, b, x)
    bp, xp, m = _solve_vectors(op64, b, x)
    r = _SolveResult()
    inner = U64()
    irc = _IrConfig(inner_rtol, inner_max_iter, max_outer,
                    _ChebConfig(degree, ratio, safety, power_iters, lmax), mode, update_interval)
    rc = lib.femgpu_solve_ir_mixed(op64.h, op32.h, C.byref(irc), bp, xp, m, C.byref(cfg._c()),
                                   C.byref(r), C.byref(inner))
    if rc not in (OK, ERR_NOT_CONVERGED):
        _check(rc)
    return _result(r), inner.value


def solve_ir_learned(op64, op32, net, b, x, cfg=None, inner_rtol=1e-4, inner_max_iter=0, max_outer=50,
                     degree=4, ratio=30.0, safety=1.1, power_iters=20, patience=20, truncation=8,
                     lmax=0.0):
    """FP64 refinement around the FP32 flexible CG with the network; returns
    (SolveResult with outer iterations, total inner iterations, FcgInfo totals)."""
    cfg = cfg or SolverConfig()
    _check_dims(op64, b, x)
    bp, xp, m = _solve_vectors(op64, b, x)
    r = _SolveResult()
    inner = U64()
    info = _FcgInfo(patience, truncation, 0, 0, 0, 0, 0)
    irc = _IrConfig(inner_rtol, inner_max_iter, max_outer,
                    _ChebConfig(degree, ratio, safety, power_iters, lmax), 0, 0)
    rc = lib.femgpu_solve_ir_learned(op64.h, op32.h, net.h, C.byref(irc), bp, xp, m, C.byref(cfg._c()),
                                     C.byref(r), C.byref(inner), C.byref(info))
    if rc not in (OK, ERR_NOT_CONVERGED):
        _check(rc)
    return _result(r), inner.value, FcgInfo(int(info.fcg_iterations), int(info.fallback_iterations),
                                            bool(info.fell_back), int(info.fallback_reason),
                                            int(info.net_corrections))


def chebyshev_jacobi_apply(op, r, z, degree=4, ratio=30.0, safety=1.1, power_iters=20, lmax=0.0):
    """z = p(D^-1 A) D^-1 r, the textbook Chebyshev-Jacobi polynomial."""
    rp, m = _vec(r, op.dtype, op.n_dofs(), "r")
    zp, _ = _vec(z, op.dtype, op.n_dofs(), "z")
    cc = _ChebConfig(degree, ratio, safety, power_iters, lmax)
    _check(lib.femgpu_chebyshev_jacobi_apply(op.h, C.byref(cc), rp, zp, m))


# ------------------------------------------------- learned preconditioner --
class UNet:
    """3-level 3-D conv U-Net (16/32/64 channels, BF16 on MFMA) on an
    nx*ny*nz grid; see include/femgpu.h for the parameter layout."""

    def __init__(self, shape, seed=0, ctx=None, coarsen=1):
        """shape: the DOF grid; coarsen > 1 places the network on every
        coarsen-th node of it (femgpu_unet_create_coarse).  self.V is the
        network grid's size, self.n_dofs the DOF grid's."""
        self.ctx = ctx or default_context()
        nx, ny, nz = shape
        self.n_dofs = nx * ny * nz
        self.coarsen = coarsen
        h = P()
        if coarsen > 1:
            _check(lib.femgpu_unet_create_coarse(self.ctx.h, nx, ny, nz, coarsen, seed, C.byref(h)))
        else:
            _check(lib.femgpu_unet_create(self.ctx.h, nx, ny, nz, seed, C.byref(h)))
        self.h = h
        g = (I(), I(), I())
        _check(lib.femgpu_unet_grid(h, *(C.byref(v) for v in g)))
        self.shape = tuple(v.value for v in g)
        self.V = self.shape[0] * self.shape[1] * self.shape[2]
        n = U64()
        _check(lib.femgpu_unet_n_params(h, C.byref(n)))
        self.n_params = n.value

    @classmethod
    def for_operator(cls, op, seed=0, coarsen=1):
        n1 = op.p
        shape = tuple(s * n1 + 1 for s in op.shape)
        net = cls(shape, seed, op.ctx, coarsen=coarsen)
        _check(lib.femgpu_unet_aux_from_hex(net.h, op.h))
        return net

    def layer_info(self, layer):
        ci, co, kp = I(), I(), I()
        wo, bo = U64(), U64()
        _check(lib.femgpu_unet_layer_info(self.h, layer, C.byref(ci), C.byref(co), C.byref(kp),
                                          C.byref(wo), C.byref(bo)))
        return dict(cin=ci.value, cout=co.value, kpad=kp.value, woff=wo.value, boff=bo.value)

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(lib.femgpu_unet_get_params(self.h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(lib.femgpu_unet_set_params(self.h, p.ctypes.data))

    def set_aux(self, aux):
        a = np.ascontiguousarray(aux, np.float32).ravel()
        assert a.size == self.V
        _check(lib.femgpu_unet_set_aux(self.h, a.ctypes.data, HOST))

    def forward(self, x0):
        x0 = np.ascontiguousarray(x0, np.float32).ravel()
        out = np.zeros(self.V, np.float32)
        _check(lib.femgpu_unet_forward(self.h, x0.ctypes.data, out.ctypes.data, HOST))
        return out

    def backward(self, dout):
        dout = np.ascontiguousarray(dout, np.float32).ravel()
        g = np.zeros(self.n_params, np.float32)
        _check(lib.femgpu_unet_backward(self.h, dout.ctypes.data, g.ctypes.data, HOST))
        return g

    def train_step(self, r, e, lr=1e-3):
        loss = D()
        if isinstance(r, DeviceArray):
            _check(lib.femgpu_unet_train_step(self.h, r.ptr, e.ptr, DEVICE, lr, C.byref(loss)))
        else:
            r = np.ascontiguousarray(r, np.float32); e = np.ascontiguousarray(e, np.float32)
            _check(lib.femgpu_unet_train_step(self.h, r.ctypes.data, e.ctypes.data, HOST, lr,
                                              C.byref(loss)))
        return loss.value

    def train_step_energy(self, op, r, e, lr=1e-3):
        """Adam step on |e - g N(r)|_A^2 / |e|_A^2 (g the Galerkin step); returns
        that relative energy-norm error before the step."""
        loss = D()
        if isinstance(r, DeviceArray):
            _check(lib.femgpu_unet_train_step_energy(self.h, op.h, r.ptr, e.ptr, DEVICE, lr, C.byref(loss)))
            return loss.value
        r = np.ascontiguousarray(r, np.float32); e = np.ascontiguousarray(e, np.float32)
        _check(lib.femgpu_unet_train_step_energy(self.h, op.h, r.ctypes.data, e.ctypes.data, HOST, lr,
                                                 C.byref(loss)))
        return loss.value

    def aux_from(self, op):
        """aux channel = log(kappa)/log(100) at the DOF nodes of hex operator op."""
        _check(lib.femgpu_unet_aux_from_hex(self.h, op.h))

    def apply(self, r, dtype=F64):
        r = np.ascontiguousarray(r, _npdt(dtype))
        z = np.zeros_like(r)
        _check(lib.femgpu_unet_apply(self.h, r.ctypes.data, z.ctypes.data, dtype, HOST))
        return z

    def save(self, path):
        _check(lib.femgpu_unet_save(self.h, str(path).encode()))

    def load(self, path):
        _check(lib.femgpu_unet_load(self.h, str(path).encode()))

    def close(self):
        if getattr(self, "h", None):
            if not _EXITING:
                lib.femgpu_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class FcgInfo:
    fcg_iterations: int
    fallback_iterations: int
    fell_back: bool
    fallback_reason: int  # 1 stagnation, 2 non-finite, 3 p^T A p <= 0
    net_corrections: int  # preconditioner applications that used the network


def solve_fcg_learned(a, net, b, x, cfg=None, degree=4, ratio=30.0, safety=1.1, power_iters=20,
                      lmax=0.0, patience=20, truncation=8):
    """Flexible CG whose preconditioner is Chebyshev-Jacobi smoothing followed by
    the network's Galerkin-scaled correction of the remaining residual; falls
    back to plain PCG-Chebyshev on stagnation.  Returns (SolveResult, FcgInfo)."""
    cfg = cfg or SolverConfig()
    _check_dims(a, b, x)
    bp, xp, m = _solve_vectors(a, b, x)
    r = _SolveResult()
    info = _FcgInfo(patience, truncation, 0, 0, 0, 0, 0)
    cc = _ChebConfig(degree, ratio, safety, power_iters, lmax)
    rc = lib.femgpu_solve_fcg_learned(a.h, net.h, C.byref(cc), bp, xp, m, C.byref(cfg._c()),
                                      C.byref(r), C.byref(info))
    if rc not in (OK, ERR_NOT_CONVERGED):
        _check(rc)
    return _result(r), FcgInfo(int(info.fcg_iterations), int(info.fallback_iterations),
                               bool(info.fell_back), int(info.fallback_reason), int(info.net_corrections))
