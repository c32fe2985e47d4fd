// Host-callback operator: a linear operator whose apply is a function of the
// caller's, so that the solvers take an operator this library does not own.
//
// Reference behaviour: solve_cg_operator<F: Fn(&[f64], &mut [f64])>
// (solver/src/lib.rs:198-208) calls `apply(x, y)` with overwrite semantics
// (y <- A x, lib.rs:194) on host slices; its real caller is a matrix-free
// operator outside the solver crate (examples/src/maxwell.rs:173-200).  Here
// the solvers keep their vectors on the device, so each apply hands the
// callback either host copies (FEMGPU_HOST: x is copied into a pinned host
// buffer, the callback fills a second one, which is copied back) or the device
// pointers themselves (FEMGPU_DEVICE).  MatFreeOperator::apply's accumulate
// form (partial.rs:45-47) adds the callback's result to y on the device.
#include "internal.hpp"

#include <cstring>

namespace femgpu {

template <class T>
__global__ void k_add_into(size_t n, const T *__restrict__ t, T *__restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] += t[i];
}

static unsigned cb_grid(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

struct CallbackOp final : femgpu_op {
    femgpu_apply_fn fn = nullptr;
    void *user = nullptr;
    femgpu_mem mem = FEMGPU_HOST;
    DevBuf dg;              // optional diagonal (n values of dtype)
    bool has_diag = false;
    DevBuf tmp;             // device result of an accumulating apply
    void *hx = nullptr, *hy = nullptr;  // pinned host staging (FEMGPU_HOST)
    int *hflag = nullptr;   // pinned copy of the solvers' stop flag

    const char *name() const override { return "callback"; }

    ~CallbackOp() override {
        if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (hx) (void)hipHostFree(hx);
        if (hy) (void)hipHostFree(hy);
        if (hflag) (void)hipHostFree(hflag);
    }

    void apply_dev(const void *x, void *y, bool acc, const int *skip) override {
        hipStream_t s = ctx->stream;
        const size_t bytes = n * esize();
        if (n == 0) return;
        if (skip) FG_HIP(hipMemcpyAsync(hflag, skip, sizeof(int), hipMemcpyDeviceToHost, s));
        if (mem == FEMGPU_HOST) FG_HIP(hipMemcpyAsync(hx, x, bytes, hipMemcpyDeviceToHost, s));
        // the callback runs on the host after everything before it on the stream
        FG_HIP(hipStreamSynchronize(s));
        if (skip && *hflag) return;  // the solver has stopped: later kernels skip too
        void *dst = acc ? tmp.ptr : y;
        int rc;
        if (mem == FEMGPU_HOST) {
            rc = fn(user, hx, hy, n);
            if (rc == 0) FG_HIP(hipMemcpyAsync(dst, hy, bytes, hipMemcpyHostToDevice, s));
        } else {
            rc = fn(user, x, dst, n);
        }
        FG_REQUIRE(rc == 0, FEMGPU_ERR_INVALID_ARG, "callback operator: apply returned " + std::to_string(rc));
        if (acc) {
            if (dtype == FEMGPU_F64)
                hipLaunchKernelGGL(k_add_into<double>, dim3(cb_grid(n)), dim3(256), 0, s, n,
                                   (const double *)tmp.ptr, (double *)y);
            else
                hipLaunchKernelGGL(k_add_into<float>, dim3(cb_grid(n)), dim3(256), 0, s, n,
                                   (const float *)tmp.ptr, (float *)y);
            FG_HIP(hipGetLastError());
        }
    }

    void diagonal_dev(void *d) override {
        FG_REQUIRE(has_diag, FEMGPU_ERR_UNSUPPORTED,
                   "callback operator: no diagonal was given at creation (needed by Jacobi / Chebyshev)");
        FG_HIP(hipMemcpyAsync(d, dg.ptr, n * esize(), hipMemcpyDeviceToDevice, ctx->stream));
    }
};

}  // namespace femgpu

using namespace femgpu;

extern "C" {

int femgpu_op_create_callback(femgpu_ctx *c, uint64_t n, femgpu_dtype dtype, femgpu_mem mem,
                              femgpu_apply_fn apply, void *user, const void *diag, femgpu_op **out) {
    return guard([&] {
        FG_REQUIRE(c && apply && out, FEMGPU_ERR_INVALID_ARG, "op_create_callback: NULL argument");
        FG_REQUIRE(dtype == FEMGPU_F64 || dtype == FEMGPU_F32, FEMGPU_ERR_INVALID_ARG,
                   "op_create_callback: unknown dtype");
        FG_REQUIRE(mem == FEMGPU_HOST || mem == FEMGPU_DEVICE, FEMGPU_ERR_INVALID_ARG,
                   "op_create_callback: unknown memory space");
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        FG_HIP(hipSetDevice(c->device));
        auto *op = new CallbackOp();
        try {
            op->ctx = c;
            op->dtype = dtype;
            op->n = n;
            op->fn = apply;
            op->user = user;
            op->mem = mem;
            const size_t bytes = std::max<size_t>(n * op->esize(), 8);
            FG_HIP(hipHostMalloc(&op->hflag, sizeof(int), hipHostMallocDefault));
            op->tmp.alloc(bytes);
            if (mem == FEMGPU_HOST) {
                FG_HIP(hipHostMalloc(&op->hx, bytes, hipHostMallocDefault));
                FG_HIP(hipHostMalloc(&op->hy, bytes, hipHostMallocDefault));
            }
            if (diag) {
                op->dg.alloc(bytes);
                FG_HIP(hipMemcpyAsync(op->dg.ptr, diag, n * op->esize(), hipMemcpyHostToDevice, c->stream));
                FG_HIP(hipStreamSynchronize(c->stream));
                op->has_diag = true;
            }
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
        return 0;
    });
}

}  // extern "C"
