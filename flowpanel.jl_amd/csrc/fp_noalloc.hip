// fp_noalloc.hip -- Uind! at device-resident targets without any device allocation.
//
// fp_uind (the reference's Uind!) stages the targets and the output on the
// device and, when the targets are too few to fill the GPU, splits the panels
// over grid.y with per-split partial sums -- all in buffers the library
// allocates per call.  fp_uind_noalloc runs the same kernels on caller-owned
// device arrays and takes its staging and partial sums from one file-scope
// device array instead, so that it can be called where hipMalloc cannot (under
// stream capture, in tight time-stepping loops).
#include "fp_internal.h"

#include <algorithm>
#include <cstdio>
#include <mutex>

namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;   // per GPU
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr size_t NOALLOC_ALIGN = 256;
}  // namespace

// The scratch of fp_uind_noalloc: SoA staging of the targets (3 x nt doubles)
// followed by the partial sums of the split field launch.  Every GPU the
// library runs on holds its own copy until the process ends.  In host code the
// name is only a stand-in: kernels receive the device address returned by
// hipGetSymbolAddress for the current device.
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
std::mutex g_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES];   // device address of `scratch` on each GPU
int g_scratch_state[NOALLOC_MAX_DEVICES];   // 0 not yet queried, 1 ready, -1 query failed
hipError_t g_scratch_err[NOALLOC_MAX_DEVICES];

__global__ void k_aos_to_soa3(const double *__restrict__ aos, int64_t n, double *__restrict__ soa)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    soa[i] = aos[3 * i];
    soa[n + i] = aos[3 * i + 1];
    soa[2 * n + i] = aos[3 * i + 2];
}

bool device_array(const void *p, int device)
{
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) { (void)hipGetLastError(); return false; }
    return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) && a.device == device;
}
}  // namespace

// Set-up of the scratch address, once per GPU, with that GPU current (called by
// fp_ctx_create).  A failed query only disables fp_uind_noalloc on that GPU.
void fp_noalloc_init(fp_ctx *ctx)
{
    const int d = ctx->device;
    if (d < 0 || d >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch_state[d] != 0) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess && ptr) {
        g_scratch_ptr[d] = ptr;
        g_scratch_state[d] = 1;
    } else {
        g_scratch_err[d] = e == hipSuccess ? hipErrorInvalidSymbol : e;
        g_scratch_state[d] = -1;
        (void)hipGetLastError();
    }
}

extern "C" int fp_uind_noalloc(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *d_targets, int64_t nt,
                               double *d_out)
{
    if (!ctx) return FP_ERR_ARG;
    if (!bodies || nb <= 0) return fp_fail(ctx, FP_ERR_ARG, "no bodies given");
    for (int k = 0; k < nb; ++k) {
        if (!bodies[k] || bodies[k]->ctx != ctx) return fp_fail(ctx, FP_ERR_ARG, "invalid body handle");
        if (!bodies[k]->solved)
            return fp_fail(ctx, FP_ERR_UNSOLVED,
                           "Source body hasn't been solved yet. Please call `solve(...)` function first.");
    }
    if (nt < 0 || nt > FP_UIND_NOALLOC_MAX_TARGETS) {
        char buf[160];
        snprintf(buf, sizeof buf, "fp_uind_noalloc takes 0..%lld targets, got %lld",
                 (long long)FP_UIND_NOALLOC_MAX_TARGETS, (long long)nt);
        return fp_fail(ctx, FP_ERR_ARG, buf);
    }
    if (nt == 0) return FP_OK;
    if (!d_targets || !d_out) return fp_fail(ctx, FP_ERR_ARG, "null targets/out");
    const int d = ctx->device;
    if (d < 0 || d >= NOALLOC_MAX_DEVICES) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "device index too large");
    FP_HIP(ctx, hipSetDevice(d));
    void *base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_scratch_state[d] != 1) {
            char buf[160];
            snprintf(buf, sizeof buf, "fp_uind_noalloc: no scratch on device %d (hipGetSymbolAddress: %s)", d,
                     g_scratch_state[d] == 0 ? "not queried" : hipGetErrorString(g_scratch_err[d]));
            return fp_fail(ctx, FP_ERR_DEVICE, buf);
        }
        base = g_scratch_ptr[d];
    }
    if (!device_array(d_targets, d) || !device_array(d_out, d))
        return fp_fail(ctx, FP_ERR_ARG, "fp_uind_noalloc: targets and out must be device memory of the context's GPU");
    // scratch layout: [targets SoA | partial sums]
    const size_t stage = ((size_t)3 * nt * sizeof(double) + NOALLOC_ALIGN - 1) / NOALLOC_ALIGN * NOALLOC_ALIGN;
    size_t need = 0;
    for (int k = 0; k < nb; ++k) need = std::max(need, fp_field_partial_bytes(nt, bodies[k]->N));
    if (stage + need > NOALLOC_SCRATCH_BYTES) {
        char buf[160];
        snprintf(buf, sizeof buf, "fp_uind_noalloc: %zu scratch bytes needed, %zu available", stage + need,
                 NOALLOC_SCRATCH_BYTES);
        return fp_fail(ctx, FP_ERR_ARG, buf);
    }
    double *soa = (double *)base;
    double *partial = (double *)((char *)base + stage);
    hipLaunchKernelGGL(k_aos_to_soa3, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream, d_targets, nt, soa);
    FP_HIP(ctx, hipGetLastError());
    TargetView tv{};
    tv.x = soa;
    tv.y = soa + nt;
    tv.z = soa + 2 * nt;
    tv.M = nt;
    // no profiling events here: they would be created on first use
    const bool prof = ctx->prof;
    ctx->prof = false;
    int st = FP_OK;
    for (int k = 0; k < nb && !st; ++k)
        st = fp_launch_field(ctx, bodies[k], bodies[k]->d_strength, tv, 0, d_out, partial,
                             NOALLOC_SCRATCH_BYTES - stage);
    ctx->prof = prof;
    return st;
}
