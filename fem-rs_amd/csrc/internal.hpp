// Internal declarations of libfemgpu (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/femgpu.h"
#include "csr_row.hpp"

namespace femgpu {

// ------------------------------------------------------------- errors --
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define FG_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::femgpu::Error(FEMGPU_ERR_HIP, std::string(#call) + ": " +               \
                                                      hipGetErrorString(e_));               \
    } while (0)

#define FG_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                    \
        if (!(cond)) throw ::femgpu::Error((code), (msg));                                  \
    } while (0)

// Wraps a C-ABI body: converts exceptions to status codes + last error.
template <class F>
int guard(F &&f) {
    try {
        set_last_error("");
        return f();
    } catch (const Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(FEMGPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(FEMGPU_ERR_INVALID_ARG, e.what());
    }
}

// ------------------------------------------------------- device buffer --
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        release();
        ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t b) {
        release();
        if (b == 0) return;
        hipError_t e = hipMalloc(&ptr, b);
        if (e != hipSuccess) {
            ptr = nullptr;
            throw Error(FEMGPU_ERR_OUT_OF_MEMORY,
                        "hipMalloc(" + std::to_string(b) + " bytes): " + hipGetErrorString(e));
        }
        bytes = b;
    }
    void ensure(size_t b) { if (b > bytes) alloc(b); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr; bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// solvers.hip: device address of the allocation-free solver's scratch array
// on `device` (which must be current); looked up once per device.
void *noalloc_scratch(int device);

// amg.hip: one V-cycle z = V r on the AMG's stream; its level-0 operator.
void amg_apply(const femgpu_amg *h, const double *r, double *z, const int *skip);
femgpu_op *amg_operator(const femgpu_amg *h);

// A square f64 CsrOp owning device arrays row_ptr u64[n+1], col_idx u32[nnz],
// values f64[nnz] (csr.hip).
femgpu_op *csr_adopt(femgpu_ctx *c, uint64_t n, uint64_t nnz, DevBuf &&rp, DevBuf &&ci, DevBuf &&vals);

// Reduction partials: a fixed grid gives a deterministic summation order.
constexpr int kRedBlocks = 1024;
// Device scalar slots per context (and the pinned host mirror): 0..63 solver
// scalars, 64..255 results of batched dot products.
constexpr int kScalarSlots = 256;
constexpr int kMultiDotBase = 64, kMultiDotMax = kScalarSlots - kMultiDotBase;
// Operator workspace slots 0..199 belong to the solvers; the network's
// energy-loss training uses two from this one on.
constexpr int kUnetOpSlot = 200;
constexpr int kRedThreads = 256;
// Longest residual history a solve records (entries, iteration 0 included).
constexpr uint64_t kHistMax = 1u << 20;

}  // namespace femgpu

// ------------------------------------------------------------- context --
struct femgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    femgpu::DevBuf partials;  // kRedBlocks * 8 doubles
    femgpu::DevBuf scalars;   // solver scalars (see solvers)
    femgpu::DevBuf flags;     // int flags: [0] stop (kernels skip work when set)
    double *pinned = nullptr; // host-visible scalars for read-back
    hipEvent_t t0 = nullptr, t1 = nullptr;
    void *noalloc_scratch = nullptr;  // this device's scratch array (solvers.hip)
    // residual history of the last solve that recorded one (print level >= 2)
    femgpu::DevBuf hist;
    std::vector<double> hist_host;
    uint64_t hist_len = 0;
};

// ------------------------------------------------------------ operator --
namespace femgpu {
// An instantiated HIP graph of a chunk of solver iterations, valid for the
// device pointers and parameters in key.
struct IterGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> key;
    std::vector<double> params;
    IterGraph() = default;
    IterGraph(const IterGraph &) = delete;
    IterGraph &operator=(const IterGraph &) = delete;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
        params.clear();
    }
    ~IterGraph() { reset(); }
};
}  // namespace femgpu

struct femgpu_op {
    femgpu_ctx *ctx = nullptr;
    femgpu_dtype dtype = FEMGPU_F64;
    uint64_t n = 0;
    // Staging for host-memory calls.
    femgpu::DevBuf stage_x, stage_y;
    // Solver workspace vectors, grown on demand and reused across solves.
    std::vector<femgpu::DevBuf> ws;
    bool diag_ready = false;  // workspace slot of the diagonal holds diag(A)
    femgpu::IterGraph pcg_graph;  // cached iteration chunk of the fused PCG-Jacobi
    void *wsvec(size_t i, size_t bytes) {
        if (ws.size() <= i) ws.resize(i + 1);
        ws[i].ensure(bytes);
        return ws[i].ptr;
    }

    // Optional profiling (femgpu_op_time_apply): when set, apply_dev records
    // prof_begin and prof_end around its kernels, and prof_mid right after its
    // main compute kernel when it has a second one (the hex PA gather).
    hipEvent_t prof_begin = nullptr, prof_mid = nullptr, prof_end = nullptr;

    virtual ~femgpu_op() = default;
    // y (+)= A x on device pointers of type dtype, enqueued on ctx->stream.
    // skip: optional device flag; when it reads non-zero the kernels do no work.
    virtual void apply_dev(const void *x, void *y, bool accumulate, const int *skip) = 0;
    virtual void diagonal_dev(void *d) = 0;
    virtual const char *name() const = 0;
    // Optional fused forms used by the solvers; false = not provided, and the
    // caller runs apply_dev followed by its own vector kernel.  Both must give
    // the results of that pair bit for bit (same per-entry arithmetic, and the
    // partial sums over the fixed red_blocks(n) x 256 grid-stride grid).
    //  apply_cheb_dev: t = A x; d = c1 d + c2 inv (r - t); z += d; with last,
    //                  part[block] = sum r.z  (x may alias z).  With first,
    //                  z equals d on entry and is not read (x = d): z = d_old + d,
    //                  and d_old = (1/theta) inv r (inv_theta = 1/theta) may be
    //                  recomputed instead of read
    //  apply_dot_dev:  y = A x; part[block] = sum x.y
    virtual bool apply_cheb_dev(const void *x, const void *r, const void *inv, void *z, void *d,
                                double c1, double c2, int last, double *part, const int *skip, int first,
                                double inv_theta) {
        (void)x; (void)r; (void)inv; (void)z; (void)d; (void)c1; (void)c2; (void)last; (void)part; (void)skip;
        (void)first; (void)inv_theta;
        return false;
    }
    virtual bool apply_dot_dev(const void *x, void *y, double *part, const int *skip) {
        (void)x; (void)y; (void)part; (void)skip;
        return false;
    }
    // Assembled CSR operators whose SpMV runs one row per thread in the
    // reference's summation order expose their arrays here, so that the
    // PCG-Jacobi iteration can form the search direction inside the SpMV.
    virtual bool csr_view(femgpu::CsrView *out) const {
        (void)out;
        return false;
    }
    // The device arrays of any assembled CSR operator, whatever kernel its
    // SpMV runs (set-up code that copies the matrix, e.g. AMG).
    virtual bool csr_arrays(femgpu::CsrView *out) const {
        (void)out;
        return false;
    }
    size_t esize() const { return dtype == FEMGPU_F64 ? 8 : 4; }
};

namespace femgpu {

// blas.hip
template <class T> void fill(hipStream_t s, T *y, T v, size_t n, const int *skip = nullptr);
template <class T> void copy(hipStream_t s, T *y, const T *x, size_t n);
// y = a*x + b*y
template <class T> void axpby(hipStream_t s, T *y, double a, const T *x, double b, size_t n);
// inv[i] = 1/d[i] (1 if |d| <= tiny)
template <class T> void inv_diag(hipStream_t s, T *inv, const T *d, size_t n, double tiny);
// Deterministic dot products in double: writes partials then finalises into out[0].
template <class T> void dot(femgpu_ctx *c, const T *a, const T *b, size_t n, double *out_dev);
double dot_host(femgpu_ctx *c, const void *a, const void *b, size_t n, femgpu_dtype dt);
// Up to kPack vectors (and coefficients) passed by value to one kernel.
constexpr int kPack = 8;
template <class T> struct VecPack {
    const T *v[kPack];
    double c[kPack];
    int m;
};
// out[j] = a . v[j], j < m (m <= kMultiDotMax), one pass over a per 8 vectors,
// one synchronisation.
template <class T>
void multidot_host(femgpu_ctx *c, const T *a, const T *const *v, int m, size_t n, double *out);
// y = x0 + sum_j coef[j] v[j] (y may be x0).
template <class T>
void multiaxpy(hipStream_t s, T *y, const T *x0, const T *const *v, const double *coef, int m, size_t n);
// m dot products a[j].b[j] (m <= kMultiDotMax) enqueued back to back, read
// back with one copy and one synchronisation.
void dots_host(femgpu_ctx *c, int m, const void *const *a, const void *const *b, size_t n,
               femgpu_dtype dt, double *out);
void convert_f64_to_f32(hipStream_t s, float *y, const double *x, size_t n, double scale);
void convert_f32_to_f64_axpy(hipStream_t s, double *y, const float *x, size_t n, double scale);

// Helpers for host/device vector arguments.
struct VecArg {
    void *dev = nullptr;   // device pointer to use
    void *host = nullptr;  // host pointer (if staged)
    size_t bytes = 0;
};

// hex operator helpers used by the learned preconditioner
void hex_log_kappa_nodes(femgpu_op *op, float *out_dev);
bool hex_grid_dims(femgpu_op *op, int *Nx, int *Ny, int *Nz);
// true when every boundary DOF of the hex operator is essential
bool hex_dirichlet(femgpu_op *op);

// learned preconditioner (unet.hip): z = |r| N(r/|r|) on device vectors
bool unet_apply_dev(femgpu_unet *u, const void *r, void *z, femgpu_dtype dt, double rnorm);
// the coarse placement's correction with |R r|^2 kept in *nrm2 on the device (no read-back)
void unet_apply_coarse_dev(femgpu_unet *u, const void *r, void *z, femgpu_dtype dt, double *nrm2);
long long unet_voxels(const femgpu_unet *u);
// true when the network runs on a coarsened copy of the DOF grid
bool unet_coarse(const femgpu_unet *u);
femgpu_ctx *unet_ctx(const femgpu_unet *u);

// hex mesh generator (host) shared by the C ABI and set-up code
void hex_vertices(int nx, int ny, int nz, double amp, uint64_t seed, double *coords);
double philox_uniform(uint64_t seed, uint32_t stream, uint64_t index);

}  // namespace femgpu
