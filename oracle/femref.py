"""ctypes bindings for the CPU oracle (oracle/_build/libfemref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker and by ``bench.py``'s cpu_baseline leg.  The product library
never imports this module.  Every wrapped function restates a fem-rs
reference algorithm; the citations live in ``femref.h`` / ``femref.c``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfemref.so")

FR_MASS, FR_DIFFUSION, FR_DIFF_MASS = 1, 2, 3
FR_COEF_CONST, FR_COEF_SINPROD, FR_COEF_GRF = 0, 1, 2

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Coef(C.Structure):
    _fields_ = [("kind", C.c_int), ("value", C.c_double), ("n_features", C.c_int),
                ("sigma", C.c_double), ("length", C.c_double), ("lo", C.c_double),
                ("hi", C.c_double), ("seed", C.c_uint64)]


class HexDesc(C.Structure):
    _fields_ = [("p", C.c_int), ("q", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("kind", C.c_int), ("rho", C.c_double), ("kappa", Coef),
                ("dirichlet", C.c_int)]


class Csr(C.Structure):
    _fields_ = [("nrows", C.c_size_t), ("ncols", C.c_size_t), ("row_ptr", C.c_void_p),
                ("col_idx", C.c_void_p), ("values", C.c_void_p)]


class SolverConfig(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("max_iter", C.c_uint64)]


class SolveResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_uint64),
                ("final_residual", C.c_double), ("status", C.c_int)]


class HexCtx(C.Structure):
    _fields_ = [("d", C.POINTER(HexDesc)), ("D", C.c_void_p), ("M", C.c_void_p),
                ("nthreads", C.c_int)]


APPLY_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))

P = C.c_void_p
D = C.c_double
I = C.c_int
Z = C.c_size_t
U64 = C.c_uint64


def _declare(L):
    sig = {
        "fr_uniform": (D, [U64, C.c_uint32, U64]),
        "fr_gauss_legendre": (I, [I, P, P]),
        "fr_tri_rule": (I, [I, P, P]),
        "fr_tet_rule": (I, [I, P, P]),
        "fr_lagrange_1d": (None, [I, I, P, P, P]),
        "fr_simplex_basis": (I, [I, I, P, P, P]),
        "fr_hex_vertices": (None, [I, I, I, D, U64, P]),
        "fr_grf_features": (None, [P, P, P]),
        "fr_coef_eval": (D, [P, P, P, P]),
        "fr_hex_ndofs": (Z, [P]),
        "fr_hex_qdata": (None, [P, P, P, P]),
        "fr_hex_apply_refloop": (None, [P, P, P, P, I]),
        "fr_hex_apply_refloop_mt": (None, [P, P, P, P, I, I]),
        "fr_hex_apply_sumfac": (None, [P, P, P, P, P, I, I]),
        "fr_hex_apply_sumfac_f32": (None, [P, P, P, P, P, I, I]),
        "fr_hex_diagonal": (None, [P, P, P, P]),
        "fr_hex_assemble_csr": (Z, [P, P, P, P, P]),
        "fr_hex_rhs_unit": (None, [P, P, P]),
        "fr_csr_spmv": (None, [P, P, P]),
        "fr_csr_spmv_add": (None, [P, D, P, D, P]),
        "fr_csr_spmv_f32": (None, [Z, P, P, P, P, P]),
        "fr_csr_diagonal": (None, [P, P]),
        "fr_unit_square_tri": (None, [I, P, P]),
        "fr_simplex_assemble_csr": (Z, [I, I, Z, P, Z, P, P, Z, I, D, P, I, P, P, P]),
        "fr_simplex_qp_coords": (I, [I, P, Z, P, I, P]),
        "fr_p1_source_rhs": (None, [I, P, P, I, P]),
        "fr_unit_square_boundary": (Z, [I, P]),
        "fr_csr_dirichlet_row_zeroing": (None, [Z, P, P, P, Z, P, P, P]),
        "fr_simplex_pa_apply": (None, [I, I, P, Z, P, P, I, D, P, I, P, P, I]),
        "fr_simplex_lumped_mass": (None, [I, I, P, Z, P, P, Z, D, P, I, P]),
        "fr_simplex_dofs": (Z, [I, I, Z, Z, P, P]),
        "fr_solve_cg_operator": (I, [Z, P, P, P, P, P, P]),
        "fr_pcg_jacobi": (I, [Z, P, P, P, P, P, P, P]),
        "fr_chebyshev_smooth_ref": (None, [P, P, P, D, D, I]),
        "fr_estimate_spectral_radius": (D, [P, I]),
        "fr_estimate_lmax_jacobi": (D, [Z, P, P, P, I]),
        "fr_chebyshev_jacobi_apply": (None, [Z, P, P, P, I, D, D, P, P]),
        "fr_pcg_chebyshev": (I, [Z, P, P, P, I, D, D, P, P, P, P]),
        "fr_conv3d_fwd": (None, [I, I, I, I, I, P, P, P, P]),
        "fr_conv3d_bwd": (None, [I, I, I, I, I, P, P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------- basics --
def uniform(seed, stream, index):
    return lib().fr_uniform(seed, stream, index)


def gauss_legendre(n):
    x = np.zeros(n); w = np.zeros(n)
    if lib().fr_gauss_legendre(n, _p(x), _p(w)) != 0:
        raise ValueError(f"unsupported n={n}")
    return x, w


def tri_rule(order):
    p = np.zeros(24); w = np.zeros(12)
    n = lib().fr_tri_rule(order, _p(p), _p(w))
    return p[: 2 * n].reshape(n, 2), w[:n]


def tet_rule(order):
    p = np.zeros(120); w = np.zeros(40)
    n = lib().fr_tet_rule(order, _p(p), _p(w))
    return p[: 3 * n].reshape(n, 3), w[:n]


def lagrange_1d(p, xq):
    xq = _f64(xq)
    B = np.zeros((len(xq), p + 1)); G = np.zeros((len(xq), p + 1))
    lib().fr_lagrange_1d(p, len(xq), _p(xq), _p(B), _p(G))
    return B, G


def simplex_basis(dim, order, xi):
    xi = _f64(xi)
    phi = np.zeros(20); grad = np.zeros(60)
    n = lib().fr_simplex_basis(dim, order, _p(xi), _p(phi), _p(grad))
    return phi[:n], grad[: n * dim].reshape(n, dim)


# ------------------------------------------------------------------- hex --
def coef_const(v):
    return Coef(FR_COEF_CONST, float(v), 0, 0.0, 0.0, 0.0, 0.0, 0)


def coef_sinprod():
    return Coef(FR_COEF_SINPROD, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0)


def coef_grf(seed, n_features=256, sigma=1.5, length=0.1, lo=1e-2, hi=1e2):
    return Coef(FR_COEF_GRF, 0.0, n_features, sigma, length, lo, hi, seed)


def hex_desc(p, q, n, kind=FR_DIFFUSION, kappa=None, rho=1.0, dirichlet=False):
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    return HexDesc(p, q, nx, ny, nz, kind, rho, kappa if kappa is not None else coef_const(1.0),
                   1 if dirichlet else 0)


def hex_vertices(nx, ny, nz, amp=0.0, seed=0):
    c = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    lib().fr_hex_vertices(nx, ny, nz, amp, seed, _p(c))
    return c


def grf_features(coef):
    w = np.zeros(3 * coef.n_features); b = np.zeros(coef.n_features)
    lib().fr_grf_features(C.byref(coef), _p(w), _p(b))
    return w, b


def coef_eval(coef, x):
    w, b = grf_features(coef) if coef.kind == FR_COEF_GRF else (np.zeros(3), np.zeros(1))
    x = _f64(x)
    return lib().fr_coef_eval(C.byref(coef), _p(w), _p(b), _p(x))


def hex_ndofs(d):
    return lib().fr_hex_ndofs(C.byref(d))


def hex_qdata(d, coords):
    E = d.nx * d.ny * d.nz; Q = d.q ** 3
    Dq = np.zeros((E, Q, 6)); Mq = np.zeros((E, Q))
    lib().fr_hex_qdata(C.byref(d), _p(_f64(coords)), _p(Dq), _p(Mq))
    return Dq, Mq


def hex_apply_refloop(d, coords, x, y=None, nthreads=0):
    """Reference element loop; nthreads > 0 runs the parity-coloured threaded
    form (same operator, another summation order into y)."""
    x = _f64(x)
    acc = y is not None
    y = np.zeros(hex_ndofs(d)) if y is None else _f64(y).copy()
    if nthreads > 0:
        lib().fr_hex_apply_refloop_mt(C.byref(d), _p(_f64(coords)), _p(x), _p(y), 1 if acc else 0, nthreads)
    else:
        lib().fr_hex_apply_refloop(C.byref(d), _p(_f64(coords)), _p(x), _p(y), 1 if acc else 0)
    return y


def hex_apply_sumfac(d, Dq, Mq, x, y=None, nthreads=1):
    x = _f64(x)
    acc = y is not None
    y = np.zeros(hex_ndofs(d)) if y is None else _f64(y).copy()
    lib().fr_hex_apply_sumfac(C.byref(d), _p(_f64(Dq)), _p(_f64(Mq)), _p(x), _p(y),
                              1 if acc else 0, nthreads)
    return y


def hex_apply_sumfac_f32(d, Dq, Mq, x, nthreads=1):
    Dq = np.ascontiguousarray(Dq, np.float32); Mq = np.ascontiguousarray(Mq, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(hex_ndofs(d), np.float32)
    lib().fr_hex_apply_sumfac_f32(C.byref(d), _p(Dq), _p(Mq), _p(x), _p(y), 0, nthreads)
    return y


def hex_diagonal(d, Dq, Mq):
    out = np.zeros(hex_ndofs(d))
    lib().fr_hex_diagonal(C.byref(d), _p(_f64(Dq)), _p(_f64(Mq)), _p(out))
    return out


def hex_assemble_csr(d, coords):
    N = hex_ndofs(d)
    rp = np.zeros(N + 1, np.uint64)
    c = _f64(coords)
    nnz = lib().fr_hex_assemble_csr(C.byref(d), _p(c), _p(rp), None, None)
    ci = np.zeros(nnz, np.uint32); v = np.zeros(nnz)
    lib().fr_hex_assemble_csr(C.byref(d), _p(c), _p(rp), _p(ci), _p(v))
    return rp, ci, v


def hex_rhs_unit(d, coords):
    b = np.zeros(hex_ndofs(d))
    lib().fr_hex_rhs_unit(C.byref(d), _p(_f64(coords)), _p(b))
    return b


# ------------------------------------------------------------------- CSR --
class CsrRef:
    """Keeps numpy arrays alive behind an fr_csr struct."""

    def __init__(self, row_ptr, col_idx, values, ncols=None):
        self.row_ptr = np.ascontiguousarray(row_ptr, np.uint64)
        self.col_idx = np.ascontiguousarray(col_idx, np.uint32)
        self.values = _f64(values)
        self.nrows = len(self.row_ptr) - 1
        self.ncols = self.nrows if ncols is None else ncols
        self.s = Csr(self.nrows, self.ncols, _p(self.row_ptr), _p(self.col_idx), _p(self.values))

    def spmv(self, x):
        y = np.zeros(self.nrows)
        lib().fr_csr_spmv(C.byref(self.s), _p(_f64(x)), _p(y))
        return y

    def spmv_add(self, alpha, x, beta, y):
        y = _f64(y).copy()
        lib().fr_csr_spmv_add(C.byref(self.s), alpha, _p(_f64(x)), beta, _p(y))
        return y

    def diagonal(self):
        d = np.zeros(self.nrows)
        lib().fr_csr_diagonal(C.byref(self.s), _p(d))
        return d

    def dense(self):
        A = np.zeros((self.nrows, self.ncols))
        for r in range(self.nrows):
            for k in range(int(self.row_ptr[r]), int(self.row_ptr[r + 1])):
                A[r, self.col_idx[k]] += self.values[k]
        return A


def csr_spmv_f32(row_ptr, col_idx, values, x):
    values = np.ascontiguousarray(values, np.float32); x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(len(row_ptr) - 1, np.float32)
    lib().fr_csr_spmv_f32(len(row_ptr) - 1, _p(np.ascontiguousarray(row_ptr, np.uint64)),
                          _p(np.ascontiguousarray(col_idx, np.uint32)), _p(values), _p(x), _p(y))
    return y


# --------------------------------------------------------------- simplex --
def unit_square_tri(n):
    c = np.zeros(2 * (n + 1) ** 2); conn = np.zeros(3 * 2 * n * n, np.uint32)
    lib().fr_unit_square_tri(n, _p(c), _p(conn))
    return c, conn


def simplex_dofs(dim, order, n_nodes, conn):
    conn = np.ascontiguousarray(conn, np.uint32)
    nv = dim + 1
    ne = len(conn) // nv
    nl = (order + 1) * (order + 2) // 2 if dim == 2 else (order + 1) * (order + 2) * (order + 3) // 6
    ed = np.zeros(ne * nl, np.uint32)
    nd = lib().fr_simplex_dofs(dim, order, n_nodes, ne, _p(conn), _p(ed))
    return nd, ed


def simplex_qp_coords(dim, coords, conn, quad_order):
    """Physical quadrature points (n_elem, nq, dim), x = x_0 + J xi."""
    conn = np.ascontiguousarray(conn, np.uint32)
    ne = len(conn) // (dim + 1)
    nq = lib().fr_simplex_qp_coords(dim, _p(_f64(coords)), ne, _p(conn), quad_order, None)
    xq = np.zeros((ne, nq, dim))
    lib().fr_simplex_qp_coords(dim, _p(_f64(coords)), ne, _p(conn), quad_order, _p(xq))
    return xq


def simplex_assemble_csr(dim, order, coords, conn, edofs, n_dofs, kind, coef, quad_order, coef_qp=None):
    """coef_qp: optional coefficient values (n_elem, nq) at the quadrature points."""
    coords = _f64(coords); conn = np.ascontiguousarray(conn, np.uint32)
    edofs = np.ascontiguousarray(edofs, np.uint32)
    ne = len(conn) // (dim + 1)
    rp = np.zeros(n_dofs + 1, np.uint64)
    keep = None if coef_qp is None else _f64(np.ravel(coef_qp))
    args = (dim, order, len(coords) // dim, _p(coords), ne, _p(conn), _p(edofs), n_dofs, kind,
            coef, None if keep is None else _p(keep), quad_order)
    nnz = lib().fr_simplex_assemble_csr(*args, _p(rp), None, None)
    ci = np.zeros(nnz, np.uint32); v = np.zeros(nnz)
    lib().fr_simplex_assemble_csr(*args, _p(rp), _p(ci), _p(v))
    return rp, ci, v


def p1_source_rhs(n, coords, conn, quad_order=3):
    rhs = np.zeros((n + 1) ** 2)
    lib().fr_p1_source_rhs(n, _p(_f64(coords)), _p(np.ascontiguousarray(conn, np.uint32)),
                           quad_order, _p(rhs))
    return rhs


def unit_square_boundary(n):
    b = np.zeros(4 * n + 4, np.uint32)
    k = lib().fr_unit_square_boundary(n, _p(b))
    return b[:k]


def dirichlet_row_zeroing(row_ptr, col_idx, values, bnd, rhs, bval=None):
    values = _f64(values).copy(); rhs = _f64(rhs).copy()
    bnd = np.ascontiguousarray(bnd, np.uint32)
    lib().fr_csr_dirichlet_row_zeroing(len(row_ptr) - 1, _p(np.ascontiguousarray(row_ptr, np.uint64)),
                                       _p(np.ascontiguousarray(col_idx, np.uint32)), _p(values),
                                       len(bnd), _p(bnd), _p(None if bval is None else _f64(bval)),
                                       _p(rhs))
    return values, rhs


def poisson_p1(n, quad_order=3):
    """Config-1 system: P1 diffusion CSR, source RHS, row-zeroed Dirichlet
    (examples/mfem_ex1_poisson.rs:71-97)."""
    coords, conn = unit_square_tri(n)
    N = (n + 1) ** 2
    edofs = conn.copy()
    rp, ci, v = simplex_assemble_csr(2, 1, coords, conn, edofs, N, FR_DIFFUSION, 1.0, quad_order)
    rhs = p1_source_rhs(n, coords, conn, quad_order)
    bnd = unit_square_boundary(n)
    v, rhs = dirichlet_row_zeroing(rp, ci, v, bnd, rhs)
    return rp, ci, v, rhs, coords, conn


def simplex_pa_apply(dim, order, coords, conn, edofs, kind, coef, quad_order, x, y=None, coef_qp=None):
    x = _f64(x)
    acc = y is not None
    y = np.zeros(len(x)) if y is None else _f64(y).copy()
    conn = np.ascontiguousarray(conn, np.uint32)
    keep = None if coef_qp is None else _f64(np.ravel(coef_qp))
    lib().fr_simplex_pa_apply(dim, order, _p(_f64(coords)), len(conn) // (dim + 1), _p(conn),
                              _p(np.ascontiguousarray(edofs, np.uint32)), kind, coef,
                              None if keep is None else _p(keep), quad_order, _p(x), _p(y), 1 if acc else 0)
    return y


def simplex_lumped_mass(dim, order, coords, conn, edofs, n_dofs, rho, quad_order, rho_qp=None):
    d = np.zeros(n_dofs)
    conn = np.ascontiguousarray(conn, np.uint32)
    keep = None if rho_qp is None else _f64(np.ravel(rho_qp))
    lib().fr_simplex_lumped_mass(dim, order, _p(_f64(coords)), len(conn) // (dim + 1), _p(conn),
                                 _p(np.ascontiguousarray(edofs, np.uint32)), n_dofs, rho,
                                 None if keep is None else _p(keep), quad_order, _p(d))
    return d


# --------------------------------------------------------------- solvers --
def _cb_from_python(fn, n):
    def cb(_ctx, xp, yp):
        x = np.ctypeslib.as_array(xp, shape=(n,))
        y = np.ctypeslib.as_array(yp, shape=(n,))
        y[:] = fn(x.copy())
    return APPLY_FN(cb)


class _Op:
    """An apply callback (C function + context) for the oracle solvers."""

    def __init__(self, n, fn=None, csr=None, hexop=None):
        self.n = n
        self._keep = []
        if csr is not None:
            self.fn = C.cast(lib().fr_csr_apply_cb, C.c_void_p)
            self.ctx = C.cast(C.pointer(csr.s), C.c_void_p)
            self._keep.append(csr)
        elif hexop is not None:
            d, Dq, Mq, nth = hexop
            Dq = _f64(Dq); Mq = _f64(Mq)
            self._keep += [d, Dq, Mq]
            self.hc = HexCtx(C.pointer(d), _p(Dq), _p(Mq), nth)
            self.fn = C.cast(lib().fr_hex_apply_cb, C.c_void_p)
            self.ctx = C.cast(C.pointer(self.hc), C.c_void_p)
        else:
            self.cb = _cb_from_python(fn, n)
            self.fn = C.cast(self.cb, C.c_void_p)
            self.ctx = None


def solve_cg_operator(op, b, x0, rtol=1e-8, atol=0.0, max_iter=1000):
    x = _f64(x0).copy()
    cfg = SolverConfig(rtol, atol, max_iter); res = SolveResult()
    lib().fr_solve_cg_operator(op.n, op.fn, op.ctx, _p(_f64(b)), _p(x), C.byref(cfg), C.byref(res))
    return x, res


def pcg_jacobi(op, diag, b, x0, rtol=1e-8, atol=0.0, max_iter=1000):
    x = _f64(x0).copy()
    cfg = SolverConfig(rtol, atol, max_iter); res = SolveResult()
    lib().fr_pcg_jacobi(op.n, op.fn, op.ctx, _p(_f64(diag)), _p(_f64(b)), _p(x), C.byref(cfg),
                        C.byref(res))
    return x, res


def chebyshev_smooth_ref(csr, x, b, lmin, lmax, n_iter):
    x = _f64(x).copy()
    lib().fr_chebyshev_smooth_ref(C.byref(csr.s), _p(x), _p(_f64(b)), lmin, lmax, n_iter)
    return x


def estimate_spectral_radius(csr, n_iter):
    return lib().fr_estimate_spectral_radius(C.byref(csr.s), n_iter)


def estimate_lmax_jacobi(op, diag, n_iter):
    return lib().fr_estimate_lmax_jacobi(op.n, op.fn, op.ctx, _p(_f64(diag)), n_iter)


def chebyshev_jacobi_apply(op, diag, degree, lmin, lmax, r):
    z = np.zeros(op.n)
    lib().fr_chebyshev_jacobi_apply(op.n, op.fn, op.ctx, _p(_f64(diag)), degree, lmin, lmax,
                                    _p(_f64(r)), _p(z))
    return z


def pcg_chebyshev(op, diag, degree, lmin, lmax, b, x0, rtol=1e-8, atol=0.0, max_iter=1000):
    x = _f64(x0).copy()
    cfg = SolverConfig(rtol, atol, max_iter); res = SolveResult()
    lib().fr_pcg_chebyshev(op.n, op.fn, op.ctx, _p(_f64(diag)), degree, lmin, lmax, _p(_f64(b)),
                           _p(x), C.byref(cfg), C.byref(res))
    return x, res


# ---------------------------------------------------------------- conv3d --
def conv3d_fwd(x, w, b):
    ci, nz, ny, nx = x.shape
    co = w.shape[0]
    x = np.ascontiguousarray(x, np.float32); w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    y = np.zeros((co, nz, ny, nx), np.float32)
    lib().fr_conv3d_fwd(ci, co, nz, ny, nx, _p(x), _p(w), _p(b), _p(y))
    return y


def conv3d_bwd(x, w, dy, need_dx=True):
    ci, nz, ny, nx = x.shape
    co = w.shape[0]
    x = np.ascontiguousarray(x, np.float32); w = np.ascontiguousarray(w, np.float32)
    dy = np.ascontiguousarray(dy, np.float32)
    dx = np.zeros_like(x) if need_dx else None
    dw = np.zeros_like(w); db = np.zeros(co, np.float32)
    lib().fr_conv3d_bwd(ci, co, nz, ny, nx, _p(x), _p(w), _p(dy), _p(dx), _p(dw), _p(db))
    return dx, dw, db
