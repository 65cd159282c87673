// fp_abi.hip -- implementation of the C ABI declared in include/flowpanel_hip.h.
//
// Host-side orchestration only: argument checks that mirror the reference's
// error() cases, device allocations, and the solver drivers (LU, FP32 LU +
// FP64 refinement, block-Jacobi GMRES inside FP64 refinement, FP64 GMRES).
// All arithmetic runs in the kernels of fp_influence.hip and fp_dense.hip.
#include "fp_internal.h"

#include <thread>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

// vector helpers from fp_dense.hip
int fp_mdot(fp_ctx *ctx, const double *V, int64_t ldv, int nv, const double *w, int64_t n, double *part,
            double *out);
int fp_vupdate(fp_ctx *ctx, const double *V, int64_t ldv, int nv, const double *h, double *w, int64_t n);
int fp_axpby(fp_ctx *ctx, double a, const double *x, double b, double *y, int64_t n);
int fp_vcombine(fp_ctx *ctx, const double *V, int64_t ldv, int nv, const double *z, double *x, int64_t n);
int fp_arnoldi_start(fp_ctx *ctx, double *g, int m, double beta);
int fp_arnoldi_step(fp_ctx *ctx, const double *hbuf, int m, int j, double bnorm, double *H, double *cs, double *sn,
                    double *g, double *relh, double *hnv, const double *w, double *vnext, int64_t n);
int fp_arnoldi_solve(fp_ctx *ctx, const double *H, int m, int jc, const double *g, double *y);

int fp_fail(fp_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int fp_hip_fail(fp_ctx *ctx, hipError_t e, const char *what)
{
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
    return fp_fail(ctx, e == hipErrorOutOfMemory ? FP_ERR_NOMEM : FP_ERR_DEVICE, buf);
}

namespace {

// device buffer owned by a host scope; large ones go back to their context's
// pool when released
struct DBuf {
    void *p = nullptr;
    fp_ctx *owner = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), owner(o.owner), bytes(o.bytes) { o.p = nullptr; }
    ~DBuf() { reset(); }
    void reset()
    {
        if (p) {
            if (owner && bytes >= FP_POOL_MIN) fp_pool_put(owner, p, bytes);
            else (void)hipFree(p);
        }
        p = nullptr;
    }
    template <typename T> T *as() const { return (T *)p; }
};

int dalloc(fp_ctx *ctx, DBuf &b, size_t bytes)
{
    b.reset();
    if (bytes == 0) bytes = 16;
    if (bytes >= FP_POOL_MIN) {
        for (size_t i = 0; i < ctx->pool.size(); ++i)
            if (ctx->pool[i].bytes == bytes) {
                b.p = ctx->pool[i].p;
                ctx->pool.erase(ctx->pool.begin() + (std::ptrdiff_t)i);
                ctx->pool_bytes -= bytes;
                b.owner = ctx;
                b.bytes = bytes;
                return FP_OK;
            }
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess && !ctx->pool.empty()) {
        // out of memory with buffers held for re-use: release them and try once more
        (void)hipGetLastError();
        fp_pool_trim(ctx);
        e = hipMalloc(&b.p, bytes);
    }
    if (e != hipSuccess) { b.p = nullptr; return fp_hip_fail(ctx, e, "hipMalloc"); }
    b.owner = ctx;
    b.bytes = bytes;
    return FP_OK;
}

int64_t round_ld(int64_t n) { return std::max<int64_t>(64, (n + 63) / 64 * 64); }

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st)
    {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, s);
    }
    double stop()
    {
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    void restart() { (void)hipEventRecord(a, s); }
    // stop without waiting: the elapsed time is read by elapsed() once the stream has
    // passed b (a stop() would hold the host until the GPU gets there, and the launches
    // queued after it would reach an idle GPU at the host's pace)
    void mark() { (void)hipEventRecord(b, s); }
    double elapsed()
    {
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

bool check_unit(const double *d) { return std::fabs(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] - 1.0) <= 2 * DBL_EPSILON; }

// Residual bookkeeping that never hides a NaN: the worst of two residuals is
// NaN if either is, and a residual counts as converged only when it compares
// <= tol (false for NaN).
inline double worst_residual(double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : std::max(a, b); }
inline bool converged(double rel, double tol) { return rel <= tol; }
int not_converged(fp_ctx *ctx, double rel, const char *what)
{
    char buf[160];
    if (!std::isfinite(rel)) snprintf(buf, sizeof buf, "%s produced a non-finite residual", what);
    else snprintf(buf, sizeof buf, "%s did not reach the tolerance (relative residual %.3e)", what, rel);
    return fp_fail(ctx, FP_ERR_NOT_CONVERGED, buf);
}

// Combined control points + normals of several bodies (SoA, M = sum N).
struct Targets {
    DBuf buf;
    TargetView tv{};
    int64_t M = 0;
};

int gather_targets(fp_ctx *ctx, fp_body *const *bodies, int nb, Targets &t)
{
    int64_t M = 0;
    for (int b = 0; b < nb; ++b) M += bodies[b]->N;
    t.M = M;
    if (nb == 1) {
        const fp_body *b = bodies[0];
        const int64_t N = b->N;
        t.tv.nx = b->d_geo; t.tv.ny = b->d_geo + N; t.tv.nz = b->d_geo + 2 * N;
        t.tv.x = b->d_geo + 3 * N; t.tv.y = b->d_geo + 4 * N; t.tv.z = b->d_geo + 5 * N;
        t.tv.M = N;
        return FP_OK;
    }
    int st = dalloc(ctx, t.buf, 6 * M * sizeof(double));
    if (st) return st;
    double *base = t.buf.as<double>();
    int64_t off = 0;
    for (int b = 0; b < nb; ++b) {
        const fp_body *bd = bodies[b];
        for (int c = 0; c < 6; ++c)
            FP_HIP(ctx, hipMemcpyAsync(base + c * M + off, bd->d_geo + c * bd->N, bd->N * sizeof(double),
                                       hipMemcpyDeviceToDevice, ctx->stream));
        off += bd->N;
    }
    t.tv.nx = base; t.tv.ny = base + M; t.tv.nz = base + 2 * M;
    t.tv.x = base + 3 * M; t.tv.y = base + 4 * M; t.tv.z = base + 5 * M;
    t.tv.M = M;
    return FP_OK;
}

int check_bodies(fp_ctx *ctx, fp_body *const *bodies, int nb)
{
    if (!ctx) return FP_ERR_ARG;
    if (!bodies || nb <= 0) return fp_fail(ctx, FP_ERR_ARG, "no bodies given");
    for (int b = 0; b < nb; ++b)
        if (!bodies[b] || bodies[b]->ctx != ctx) return fp_fail(ctx, FP_ERR_ARG, "invalid body handle");
    return FP_OK;
}

// ----------------------------------------------------------------------------
// operators for the iterative solvers
struct Operator {
    virtual int apply(const double *x, double *y) = 0;  // y = A x
    virtual ~Operator() {}
};

template <typename T>
struct DenseOp : Operator {
    fp_ctx *ctx; const T *A; int64_t n, ld;
    DenseOp(fp_ctx *c, const T *a, int64_t n_, int64_t ld_) : ctx(c), A(a), n(n_), ld(ld_) {}
    int apply(const double *x, double *y) override { return fp_gemv<T>(ctx, A, n, ld, x, y, 1.0, 0.0); }
};

// y = G x recomputed from the geometry in FP64 (the normal velocity induced
// at every control point by strengths x): no stored matrix.
struct MatrixFreeOp : Operator {
    fp_ctx *ctx; std::vector<fp_body *> bodies; TargetView tv; DBuf part; size_t part_bytes = 0; int64_t n;
    MatrixFreeOp(fp_ctx *c, const std::vector<fp_body *> &b, const TargetView &t) : ctx(c), bodies(b), tv(t), n(t.M) {}
    int init()
    {
        for (auto *b : bodies) part_bytes = std::max(part_bytes, fp_field_partial_bytes(tv.M, b->N));
        return part_bytes ? dalloc(ctx, part, part_bytes) : FP_OK;
    }
    int apply(const double *x, double *y) override
    {
        int st0 = fp_fill(ctx, y, n, 0.0);
        if (st0) return st0;
        int64_t col = 0;
        for (auto *b : bodies) {
            int st = fp_launch_field(ctx, b, x + col, tv, 1, y, part.as<double>(), part_bytes);
            if (st) return st;
            col += b->N;
        }
        return FP_OK;
    }
};

struct Precond {
    virtual int apply(const double *x, double *y) = 0;
    virtual ~Precond() {}
};
// Block-Jacobi: the explicit inverses of the diagonal blocks (block <= 1024)
// are formed once, so that each application is one batched GEMV.
// With coarse_block > 0 a Galerkin coarse correction on chunk indicators Z is
// applied first (two-level, multiplicative):  y = Z c + B (v - A Z c),
// c = (Z^T A Z)^-1 Z^T v.
// coarse spaces up to this many vectors are queued ahead of the block inverses (BJPrecond::setup)
constexpr int COARSE_FIRST_MAX = 1024;

struct BJPrecond : Precond {
    fp_ctx *ctx; int64_t n; int block;
    DBuf inv, gjs, AZ, Einv, work, tc, r2, cinfo, binfo;
    hipEvent_t fork = nullptr, join = nullptr, pre = nullptr, bj_done = nullptr, asm_done = nullptr;
    int cs = 0;
    bool early = false;   // the block inverses were started beside the assembly (assemble_early, invert_early)
    BJPrecond(fp_ctx *c, int64_t n_, int bl) : ctx(c), n(n_), block(bl) {}
    ~BJPrecond()
    {
        // side-stream work (coarse set-up, look-ahead Gauss-Jordan passes, early
        // inverses) must be over before its buffers can be handed out again
        if ((fork || asm_done) && ctx->side) (void)hipStreamSynchronize(ctx->side);
        if (ctx->side2) (void)hipStreamSynchronize(ctx->side2);
        if (early && ctx->side3) (void)hipStreamSynchronize(ctx->side3);
        for (hipEvent_t e : {fork, join, pre, bj_done, asm_done})
            if (e) (void)hipEventDestroy(e);
    }
    // Block inverses beside the assembly: the diagonal blocks of the body's AIC are
    // assembled into `inv` first (a few per cent of the entries, on the context
    // stream: assemble_early), then, once the whole matrix's assembly is queued
    // behind them, inverted on the fourth stream while it runs (invert_early: the
    // host queues the Gauss-Jordan passes while the assembly already runs).
    // finish_early adds the coarse space (it needs the whole matrix) and joins
    // both streams into the context stream.
    int assemble_early(const fp_body *b, const TargetView &tv)
    {
        int st;
        if (block > 1024) return fp_fail(ctx, FP_ERR_ARG, "block-Jacobi block larger than 1024");
        const int64_t nb = (n + block - 1) / block;
        if ((st = dalloc(ctx, inv, (size_t)nb * block * block * sizeof(float)))) return st;
        const size_t sf = fp_bj_gj_scratch_floats(n, block);
        if (sf && (st = dalloc(ctx, gjs, sf * sizeof(float)))) return st;
        if ((st = dalloc(ctx, binfo, sizeof(int)))) return st;
        if ((st = fp_launch_assemble_diag(ctx, b, tv, block, inv.as<float>()))) return st;
        if (!ctx->side3) FP_HIP(ctx, hipStreamCreateWithFlags(&ctx->side3, hipStreamNonBlocking));
        if (!ctx->side2) FP_HIP(ctx, hipStreamCreateWithFlags(&ctx->side2, hipStreamNonBlocking));
        FP_HIP(ctx, hipEventCreateWithFlags(&pre, hipEventDisableTiming));
        FP_HIP(ctx, hipEventCreateWithFlags(&bj_done, hipEventDisableTiming));
        FP_HIP(ctx, hipEventRecord(pre, ctx->stream));
        early = true;
        return FP_OK;
    }
    int invert_early()
    {
        int st;
        const size_t sf = fp_bj_gj_scratch_floats(n, block);
        FP_HIP(ctx, hipStreamWaitEvent(ctx->side3, pre, 0));
        if (!ctx->gj_fork) FP_HIP(ctx, hipEventCreateWithFlags(&ctx->gj_fork, hipEventDisableTiming));
        if (!ctx->gj_join) FP_HIP(ctx, hipEventCreateWithFlags(&ctx->gj_join, hipEventDisableTiming));
        fp_ctx sc;   // borrows the context's streams and events; owns none of them
        sc.device = ctx->device;
        sc.num_cu = ctx->num_cu;
        sc.stream = ctx->side3;
        sc.side2 = ctx->side2;
        sc.gj_fork = ctx->gj_fork;
        sc.gj_join = ctx->gj_join;
        st = fp_bj_gj_inverse<float>(&sc, inv.as<float>(), n, block, block, inv.as<float>(), binfo.as<int>(),
                                     sf ? gjs.as<float>() : nullptr, nullptr, true, true);
        sc.side2 = nullptr;
        sc.stream = nullptr;
        sc.gj_fork = sc.gj_join = nullptr;
        if (st) return fp_fail(ctx, st, sc.err.c_str());
        FP_HIP(ctx, hipEventRecord(bj_done, ctx->side3));
        return FP_OK;
    }
    int finish_early(const void *A, int prec, int64_t ld, int coarse)
    {
        int st;
        if (coarse > 0) {
            cs = coarse;
            const int nc = (int)((n + cs - 1) / cs);
            if ((st = dalloc(ctx, AZ, (size_t)nc * n * sizeof(float))) ||
                (st = dalloc(ctx, Einv, (size_t)nc * nc * sizeof(double))) ||
                (st = dalloc(ctx, work, fp_coarse_work_bytes(nc))) ||
                (st = dalloc(ctx, tc, fp_coarse_tc_bytes(n, nc))) || (st = dalloc(ctx, r2, n * sizeof(double))) ||
                (st = dalloc(ctx, cinfo, sizeof(int))))
                return st;
            if (!ctx->side) FP_HIP(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
            FP_HIP(ctx, hipEventCreateWithFlags(&asm_done, hipEventDisableTiming));
            FP_HIP(ctx, hipEventCreateWithFlags(&join, hipEventDisableTiming));
            FP_HIP(ctx, hipEventRecord(asm_done, ctx->stream));
            FP_HIP(ctx, hipStreamWaitEvent(ctx->side, asm_done, 0));
            fp_ctx sc;
            sc.device = ctx->device;
            sc.num_cu = ctx->num_cu;
            sc.stream = ctx->side;
            st = prec == FP_F64 ? fp_coarse_setup<double>(&sc, (const double *)A, n, ld, cs, AZ.as<float>(),
                                                          Einv.as<double>(), work.as<double>(), cinfo.as<int>())
                                : fp_coarse_setup<float>(&sc, (const float *)A, n, ld, cs, AZ.as<float>(),
                                                         Einv.as<double>(), work.as<double>(), cinfo.as<int>());
            sc.stream = nullptr;
            if (st) return fp_fail(ctx, st, sc.err.c_str());
            FP_HIP(ctx, hipEventRecord(join, ctx->side));
            FP_HIP(ctx, hipStreamWaitEvent(ctx->stream, join, 0));
        }
        FP_HIP(ctx, hipStreamWaitEvent(ctx->stream, bj_done, 0));
        return FP_OK;
    }
    int setup(const void *A, int prec, int64_t ld, int coarse, int *info_dev)
    {
        int st;
        if (block > 1024) return fp_fail(ctx, FP_ERR_ARG, "block-Jacobi block larger than 1024");
        const int64_t nb = (n + block - 1) / block;
        if ((st = dalloc(ctx, inv, (size_t)nb * block * block * sizeof(float)))) return st;
        const size_t sf = fp_bj_gj_scratch_floats(n, block);
        if (sf && (st = dalloc(ctx, gjs, sf * sizeof(float)))) return st;
        float *sp = sf ? gjs.as<float>() : nullptr;
        int nc = 0;
        if (coarse > 0) {
            cs = coarse;
            nc = (int)((n + cs - 1) / cs);
            if ((st = dalloc(ctx, AZ, (size_t)nc * n * sizeof(float)))) return st;
            if ((st = dalloc(ctx, Einv, (size_t)nc * nc * sizeof(double)))) return st;
            if ((st = dalloc(ctx, work, fp_coarse_work_bytes(nc)))) return st;
            if ((st = dalloc(ctx, tc, fp_coarse_tc_bytes(n, nc)))) return st;
            if ((st = dalloc(ctx, r2, n * sizeof(double)))) return st;
            if ((st = dalloc(ctx, cinfo, sizeof(int)))) return st;
            if (!ctx->side) FP_HIP(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
            FP_HIP(ctx, hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            FP_HIP(ctx, hipEventCreateWithFlags(&join, hipEventDisableTiming));
        }
        // the coarse space does not depend on the block inverses: it is built on the side
        // stream beside them.  The host spends a few microseconds per launch: a short coarse
        // chain (up to COARSE_FIRST_MAX vectors) is queued first so that it starts at once;
        // a long one (nc / 64 Gauss-Jordan chunks of five launches and two events each) is
        // queued after the block inverses, which would otherwise wait for the host to finish
        // queueing it (3,000 vectors: the block inverses began 4.5 ms late)
        auto queue_coarse = [&]() -> int {
            // the coarse inverse's look-ahead passes on the fourth stream, with events of their own
            // (the block inverses' look-ahead uses the third stream and gj_fork / gj_join meanwhile)
            if (!ctx->side3) FP_HIP(ctx, hipStreamCreateWithFlags(&ctx->side3, hipStreamNonBlocking));
            if (!ctx->cgj_fork) FP_HIP(ctx, hipEventCreateWithFlags(&ctx->cgj_fork, hipEventDisableTiming));
            if (!ctx->cgj_join) FP_HIP(ctx, hipEventCreateWithFlags(&ctx->cgj_join, hipEventDisableTiming));
            fp_ctx sc;
            sc.device = ctx->device;
            sc.num_cu = ctx->num_cu;
            sc.stream = ctx->side;
            sc.side2 = ctx->side3;
            sc.gj_fork = ctx->cgj_fork;
            sc.gj_join = ctx->cgj_join;
            int cst = prec == FP_F64
                          ? fp_coarse_setup<double>(&sc, (const double *)A, n, ld, cs, AZ.as<float>(), Einv.as<double>(),
                                                    work.as<double>(), cinfo.as<int>(), true)
                          : fp_coarse_setup<float>(&sc, (const float *)A, n, ld, cs, AZ.as<float>(), Einv.as<double>(),
                                                   work.as<double>(), cinfo.as<int>(), true);
            sc.stream = nullptr;
            sc.side2 = nullptr;
            sc.gj_fork = sc.gj_join = nullptr;
            if (cst) return fp_fail(ctx, cst, sc.err.c_str());
            FP_HIP(ctx, hipEventRecord(join, ctx->side));
            return FP_OK;
        };
        bool coarse_first = nc <= COARSE_FIRST_MAX;
        if (const char *cf = getenv("FLOWPANEL_COARSE_FIRST")) coarse_first = cf[0] == '1';
        if (coarse > 0) {
            FP_HIP(ctx, hipEventRecord(fork, ctx->stream));
            FP_HIP(ctx, hipStreamWaitEvent(ctx->side, fork, 0));
            if (coarse_first && (st = queue_coarse())) return st;
        }
        st = prec == FP_F64 ? fp_bj_gj_inverse<double>(ctx, (const double *)A, n, ld, block, inv.as<float>(), info_dev,
                                                       sp)
                            : fp_bj_gj_inverse<float>(ctx, (const float *)A, n, ld, block, inv.as<float>(), info_dev,
                                                      sp);
        if (st) return st;
        if (coarse > 0 && !coarse_first && (st = queue_coarse())) return st;
        if (coarse > 0) FP_HIP(ctx, hipStreamWaitEvent(ctx->stream, join, 0));
        return FP_OK;
    }
    int apply(const double *x, double *y) override
    {
        if (cs > 0) {
            int st = fp_coarse_apply(ctx, x, n, cs, AZ.as<float>(), Einv.as<double>(), tc.as<double>(), y,
                                     r2.as<double>());
            if (st) return st;
            return fp_bj_apply_inv(ctx, inv.as<float>(), n, block, r2.as<double>(), y, 1);
        }
        return fp_bj_apply_inv(ctx, inv.as<float>(), n, block, x, y, 0);
    }
};

double dnorm_host(fp_ctx *ctx, const double *x, int64_t n, double *part, double *dout, int *st)
{
    *st = fp_mdot(ctx, x, n, 1, x, n, part, dout);
    if (*st) return 0;
    double v = 0;
    hipError_t e = hipMemcpyAsync(&v, dout, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { *st = fp_hip_fail(ctx, e, "norm download"); return 0; }
    return std::sqrt(v);
}

// Restarted GMRES with right preconditioning (x = 0 on entry).  Stops when
// ||b - A x|| <= tol * ||b|| (Arnoldi estimate, confirmed by the true residual
// at the end of a cycle) or after maxit iterations.  CGS with one
// re-orthogonalisation.  The Hessenberg column, the Givens rotations and the
// residual estimate are updated on the device; the host reads the estimates
// only every few iterations, at a spacing predicted from their decay, so the
// iterations queue back to back.  Iterations run past the converged one are
// discarded (their columns are never used).
struct Gmres {
    fp_ctx *ctx; int64_t n; int m;
    DBuf V, Z, w, h, part, tmp, arn, cgs;
    int init(fp_ctx *c, int64_t n_, int m_)
    {
        ctx = c; n = n_; m = m_;
        int st;
        if ((st = dalloc(ctx, cgs, fp_cgs_work_doubles(n, m) * sizeof(double)))) return st;
        if ((st = dalloc(ctx, V, (size_t)n * (m + 1) * sizeof(double)))) return st;
        if ((st = dalloc(ctx, w, n * sizeof(double)))) return st;
        if ((st = dalloc(ctx, tmp, n * sizeof(double)))) return st;
        if ((st = dalloc(ctx, Z, n * sizeof(double)))) return st;
        if ((st = dalloc(ctx, h, (m + 2) * 2 * sizeof(double)))) return st;
        if ((st = dalloc(ctx, part, (size_t)256 * (m + 2) * sizeof(double)))) return st;
        // H (m+1) x m, cs, sn, g (m+1), relh, hn, y
        if ((st = dalloc(ctx, arn, ((size_t)(m + 1) * m + 6 * (m + 1)) * sizeof(double)))) return st;
        return FP_OK;
    }
    // solve A x = b; x overwritten. Returns iterations via *its, final relative residual estimate via *rel.
    // bnorm_known >= 0: ||b|| as the caller already computed it (same reduction)
    int solve(Operator &A, Precond *M, const double *b, double *x, double tol, int maxit, int *its, double *rel,
              double bnorm_known = -1.0)
    {
        int st = FP_OK;
        double *Vp = V.as<double>(), *wp = w.as<double>(), *hp = h.as<double>(), *pp = part.as<double>();
        double *tp = tmp.as<double>(), *zp = Z.as<double>();
        double *dH = arn.as<double>(), *dcs = dH + (size_t)(m + 1) * m, *dsn = dcs + (m + 1), *dg = dsn + (m + 1);
        double *drel = dg + (m + 1), *dhn = drel + (m + 1), *dy = dhn + (m + 1);
        const double bnorm = bnorm_known >= 0 ? bnorm_known : dnorm_host(ctx, b, n, pp, hp, &st);
        if (st) return st;
        *its = 0;
        *rel = 0;
        FP_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
        if (!std::isfinite(bnorm)) { *rel = NAN; return not_converged(ctx, NAN, "GMRES (right-hand side)"); }
        if (bnorm == 0) return FP_OK;
        std::vector<double> relh(m), hnh(m);
        bool first = true;
        for (;;) {
            // r = b - A x
            if (first) {
                FP_HIP(ctx, hipMemcpyAsync(wp, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            } else {
                if ((st = A.apply(x, wp))) return st;
                if ((st = fp_residual_update(ctx, b, wp, n))) return st;
            }
            // (first cycle: w = b, whose norm is bnorm)
            const double beta = first ? bnorm : dnorm_host(ctx, wp, n, pp, hp, &st);
            first = false;
            if (st) return st;
            *rel = beta / bnorm;
            if (!std::isfinite(*rel)) return not_converged(ctx, *rel, "GMRES");
            if (*rel <= tol || *its >= maxit) return FP_OK;
            if ((st = fp_axpby(ctx, 1.0 / beta, wp, 0.0, Vp, n))) return st;
            if ((st = fp_arnoldi_start(ctx, dg, m, beta))) return st;
            int j = 0, jc = -1, checked = 0, next_check = 4;
            double last_rel = *rel;
            int last_j = 0;
            bool zready = false;  // M^-1 v_j already queued (before a host check)
            while (j < m && *its < maxit) {
                const double *vj = Vp + (int64_t)j * n;
                const double *zin = vj;
                if (M) {
                    if (!zready && (st = M->apply(vj, zp))) return st;
                    zin = zp;
                }
                zready = false;
                if ((st = A.apply(zin, wp))) return st;
                // CGS2 and the norm of what is left: two fused passes (restart <= FP_CGS_MAXV - 1), or
                // separate dot, update and norm launches
                if (m + 1 <= FP_CGS_MAXV) {
                    double *c1 = cgs.as<double>(), *c2 = c1 + (size_t)FP_CGS_PARTS * (m + 2);
                    double *cn = c2 + (size_t)FP_CGS_PARTS * (m + 2);
                    if ((st = fp_cgs_pass(ctx, Vp, n, j + 1, wp, n, c1, nullptr))) return st;
                    if ((st = fp_cgs_pass(ctx, Vp, n, j + 1, wp, n, c2, cn))) return st;
                    if ((st = fp_arnoldi_step2(ctx, c1, c2, cn, m, j, bnorm, dH, dcs, dsn, dg, drel, dhn, wp,
                                               Vp + (int64_t)(j + 1) * n, n)))
                        return st;
                } else {
                    if ((st = fp_mdot(ctx, Vp, n, j + 1, wp, n, pp, hp))) return st;
                    if ((st = fp_vupdate(ctx, Vp, n, j + 1, hp, wp, n))) return st;
                    if ((st = fp_mdot(ctx, Vp, n, j + 1, wp, n, pp, hp + (m + 2)))) return st;
                    if ((st = fp_vupdate(ctx, Vp, n, j + 1, hp + (m + 2), wp, n))) return st;
                    if ((st = fp_mdot(ctx, wp, n, 1, wp, n, pp, hp + 2 * (m + 2) - 1))) return st;
                    if ((st = fp_arnoldi_step(ctx, hp, m, j, bnorm, dH, dcs, dsn, dg, drel, dhn, wp,
                                              Vp + (int64_t)(j + 1) * n, n)))
                        return st;
                }
                ++j;
                ++*its;
                if (j < next_check && j < m && *its < maxit) continue;
                // the next basis vector's preconditioner runs on the device while the
                // host waits for the estimates (discarded if this check ends the cycle)
                if (M && j < m && *its < maxit) {
                    if ((st = M->apply(Vp + (int64_t)j * n, zp))) return st;
                    zready = true;
                }
                const int cnt = j - checked;
                FP_HIP(ctx, hipMemcpyAsync(relh.data() + checked, drel + checked, cnt * sizeof(double),
                                           hipMemcpyDeviceToHost, ctx->stream));
                FP_HIP(ctx, hipMemcpyAsync(hnh.data() + checked, dhn + checked, cnt * sizeof(double),
                                           hipMemcpyDeviceToHost, ctx->stream));
                FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
                for (int q = checked; q < j; ++q)
                    if (relh[q] <= tol || hnh[q] == 0 || !std::isfinite(relh[q])) { jc = q + 1; break; }
                if (jc > 0) {
                    *its -= j - jc;
                    break;
                }
                // next check where the decay of the estimate says tol is reached
                const double r = relh[j - 1];
                int K = 8;
                if (j > last_j && r < last_rel && r > 0) {
                    const double rate = std::log(r / last_rel) / (j - last_j);
                    const double remaining = std::log(tol / r) / rate;
                    K = (int)std::max(1.0, std::min(16.0, std::floor(0.8 * remaining)));
                }
                last_rel = r;
                last_j = j;
                checked = j;
                next_check = j + K;
            }
            if (jc < 0) jc = j;
            *rel = relh[jc - 1];
            if (!std::isfinite(*rel)) return not_converged(ctx, *rel, "GMRES");
            if ((st = fp_arnoldi_solve(ctx, dH, m, jc, dg, dy))) return st;
            if ((st = fp_vcombine(ctx, Vp, n, jc, dy, tp, n))) return st;
            if (M) {
                if ((st = M->apply(tp, zp))) return st;
                if ((st = fp_axpby(ctx, 1.0, zp, 1.0, x, n))) return st;
            } else {
                if ((st = fp_axpby(ctx, 1.0, tp, 1.0, x, n))) return st;
            }
            if (*rel <= tol || *its >= maxit) {
                // the true residual of the final iterate decides; restart if the
                // Arnoldi estimate was optimistic and iterations are left
                if ((st = A.apply(x, wp))) return st;
                if ((st = fp_residual_update(ctx, b, wp, n))) return st;
                const double rn = dnorm_host(ctx, wp, n, pp, hp, &st);
                if (st) return st;
                *rel = rn / bnorm;
                if (!std::isfinite(*rel)) return not_converged(ctx, *rel, "GMRES");
                if (*rel <= tol || *its >= maxit) return FP_OK;
            }
        }
    }
};

// Row permutation (perm[i] = original row at row i) -> LAPACK swap sequence
// (0-based), on the host: only needed to report factors in LAPACK form.
static std::vector<int> perm_to_ipiv(const std::vector<int> &perm)
{
    const int64_t n = (int64_t)perm.size();
    std::vector<int> cur(n), where(n), ipiv(n);
    for (int64_t i = 0; i < n; ++i) { cur[i] = (int)i; where[i] = (int)i; }
    for (int64_t i = 0; i < n; ++i) {
        const int want = perm[i];
        const int p = where[want];
        ipiv[i] = p;
        const int other = cur[i];
        cur[p] = other; where[other] = p;
        cur[i] = want; where[want] = (int)i;
    }
    return ipiv;
}

template <typename T>
int factor(fp_ctx *ctx, T *A, int64_t n, int64_t ld, int *d_perm, int *info_out)
{
    DBuf work, info;
    int st;
    if ((st = dalloc(ctx, work, fp_getrf_work_bytes(n, sizeof(T))))) return st;
    if ((st = dalloc(ctx, info, sizeof(int)))) return st;
    if ((st = fp_getrf<T>(ctx, A, n, ld, d_perm, info.as<int>(), work.p, fp_getrf_work_bytes(n, sizeof(T)))))
        return st;
    int hinfo = 0;
    FP_HIP(ctx, hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (info_out) *info_out = hinfo;
    if (hinfo) {
        char buf[128];
        snprintf(buf, sizeof buf, "singular matrix: zero pivot at column %d", hinfo);
        return fp_fail(ctx, FP_ERR_SINGULAR, buf);
    }
    return FP_OK;
}

// Core solver over an operator for the FP64 residual and (optionally) a
// stored matrix of the right precision.  b, x device FP64 vectors.
// ---- preconditioner blocks along a spatial order --------------------------
// The block-Jacobi blocks and coarse chunks are consecutive index ranges.  On
// an unstructured mesh consecutive panels need not be neighbours, so with
// FP_ORDER_SPATIAL the preconditioner is built for P A P^T, P the permutation
// that sorts the control points along a Morton curve, and applied as
// P^T M(P A P^T) P: the vector is gathered into that order, preconditioned and
// scattered back.  The operator and the FP64 residual keep the bodies' order.
__global__ void k_perm_gather(const double *__restrict__ in, const int *__restrict__ perm, int64_t n,
                              double *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = in[perm[q]];
}

__global__ void k_perm_scatter(const double *__restrict__ in, const int *__restrict__ perm, int64_t n,
                               double *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[perm[q]] = in[q];
}

// Ap[q][c] = A[perm[q]][perm[c]] (row-major, padding columns zero): one grid row
// per output row (rows beyond gridDim.y in a loop), a row of A read from L2.
template <typename T>
__global__ __launch_bounds__(256) void k_perm_matrix(const T *__restrict__ A, int64_t ld, const int *__restrict__ perm,
                                                     int64_t n, T *__restrict__ Ap)
{
    for (int64_t q = blockIdx.y; q < n; q += gridDim.y) {
        const T *row = A + (int64_t)perm[q] * ld;
        T *out = Ap + q * ld;
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ld; c += (int64_t)gridDim.x * blockDim.x)
            out[c] = c < n ? row[perm[c]] : T(0);
    }
}

// Control points sorted along a Morton curve (10 bits per axis, each axis
// scaled to its own extent; ties keep the index order): perm[q] = the panel at
// position q.
static int spatial_perm(fp_ctx *ctx, const TargetView &tv, DBuf &perm)
{
    const int64_t n = tv.M;
    std::vector<double> X((size_t)3 * n);
    FP_HIP(ctx, hipMemcpyAsync(X.data(), tv.x, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(X.data() + n, tv.y, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(X.data() + 2 * n, tv.z, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = *std::min_element(X.begin() + d * n, X.begin() + (d + 1) * n);
        hi[d] = *std::max_element(X.begin() + d * n, X.begin() + (d + 1) * n);
    }
    std::vector<uint32_t> key(n);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t k = 0;
        for (int d = 0; d < 3; ++d) {
            const double u = (X[d * n + i] - lo[d]) / std::max(hi[d] - lo[d], 1e-300);
            uint32_t q = (uint32_t)std::min(1023.0, std::max(0.0, std::floor(u * 1023.0)));
            for (int bit = 0; bit < 10; ++bit) k |= ((q >> bit) & 1u) << (3 * bit + d);
        }
        key[i] = k;
    }
    std::vector<int> p(n);
    for (int64_t i = 0; i < n; ++i) p[i] = (int)i;
    std::stable_sort(p.begin(), p.end(), [&](int x, int y) { return key[x] < key[y]; });
    int st;
    if ((st = dalloc(ctx, perm, n * sizeof(int)))) return st;
    FP_HIP(ctx, hipMemcpyAsync(perm.p, p.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));   // p goes out of scope
    return FP_OK;
}

struct PermPrecond : Precond {
    fp_ctx *ctx;
    Precond *inner;
    const int *perm;
    int64_t n;
    DBuf xp, yp;
    PermPrecond(fp_ctx *c, Precond *m, const int *p, int64_t nn) : ctx(c), inner(m), perm(p), n(nn) {}
    int init() { int st = dalloc(ctx, xp, n * sizeof(double)); return st ? st : dalloc(ctx, yp, n * sizeof(double)); }
    int apply(const double *x, double *y) override
    {
        const dim3 g((unsigned)((n + 255) / 256));
        hipLaunchKernelGGL(k_perm_gather, g, dim3(256), 0, ctx->stream, x, perm, n, xp.as<double>());
        FP_HIP(ctx, hipGetLastError());
        int st = inner->apply(xp.as<double>(), yp.as<double>());
        if (st) return st;
        hipLaunchKernelGGL(k_perm_scatter, g, dim3(256), 0, ctx->stream, (const double *)yp.as<double>(), perm, n, y);
        FP_HIP(ctx, hipGetLastError());
        return FP_OK;
    }
};

struct SolveArgs {
    int solver;
    int64_t n, ld;
    const TargetView *geo = nullptr;   // control points (FP_ORDER_SPATIAL)
    int64_t nrhs = 1;                  // right-hand sides solved together (choice of the LU32 correction)
    double *A64 = nullptr;   // FP64 matrix (LU path, or residual for uploaded matrices)
    float *A32 = nullptr;    // FP32 matrix (LU32_IR / GMRES_IR)
    Operator *resid = nullptr;  // FP64 residual operator
    const double *R64 = nullptr;  // the stored FP64 matrix behind `resid`, if any (batched residuals)
};

// A solver prepared once for a matrix (factorisation or preconditioner) and
// then applied to any number of right-hand sides.
// Seed of the random butterflies of the FP32 no-pivot LU (any fixed value: the
// refinement, not the butterflies, decides the result).
constexpr uint64_t RBT_SEED = 0x46504C5533324952ull;

struct Solver {
    fp_ctx *ctx;
    SolveArgs a;
    fp_solve_opts o;
    DBuf perm, work, info, r, d, part, dout;
    std::unique_ptr<Gmres> g;
    std::unique_ptr<BJPrecond> M;
    std::unique_ptr<PermPrecond> PM;   // M in a spatial order (FP_ORDER_SPATIAL)
    DBuf order, Aperm;
    std::unique_ptr<DenseOp<float>> A32op;
    Precond *prec() { return PM ? (Precond *)PM.get() : (Precond *)M.get(); }
    // block inverses started beside the assembly (BJPrecond::setup_early)
    void adopt(std::unique_ptr<BJPrecond> pre) { M = std::move(pre); }
    // FP32 LU of the randomized matrix U^T A V without pivoting (FP_SOLVER_LU32_IR):
    // Ar (n4 x ld4), the butterflies, identity row order for getrs, an n4 x k work vector
    bool rbt = false;
    int64_t n4 = 0, ld4 = 0;
    DBuf Ar, Wrbt, perm4, work4, t4;
    // explicit inverses of the no-pivot factors (-L^-1, -U^-1): a correction is two GEMVs
    bool inv = false;
    DBuf YL, YU, w4;
    // the getrs workspace whose diagonal-block inverses belong to the current factors
    const void *diag_ready = nullptr;
    Solver(fp_ctx *c, const SolveArgs &args, const fp_solve_opts &opts) : ctx(c), a(args), o(opts) {}

    // Randomize and factor without pivoting; false (and nothing kept) when a
    // diagonal block meets a zero or non-finite pivot: the caller then pivots.
    int rbt_setup(bool *ok)
    {
        const int64_t n = a.n;
        int st;
        *ok = false;
        diag_ready = nullptr;
        n4 = fp_rbt_size(n);
        ld4 = round_ld(n4);
        DBuf nw;
        if ((st = dalloc(ctx, Ar, (size_t)n4 * ld4 * sizeof(float))) ||
            (st = dalloc(ctx, Wrbt, fp_rbt_weights_bytes(n))) || (st = dalloc(ctx, perm4, n4 * sizeof(int))) ||
            (st = dalloc(ctx, work4, fp_getrs_work_bytes(n4, 1))) || (st = dalloc(ctx, t4, n4 * sizeof(double))) ||
            (st = dalloc(ctx, nw, fp_getrf_nopiv_work_bytes())) || (st = dalloc(ctx, info, sizeof(int))))
            return st;
        if ((st = fp_rbt_weights(ctx, n, RBT_SEED, Wrbt.as<double>()))) return st;
        st = a.R64 ? fp_rbt_matrix<double>(ctx, a.R64, n, a.ld, Ar.as<float>(), ld4, Wrbt.as<double>())
                   : fp_rbt_matrix<float>(ctx, a.A32, n, a.ld, Ar.as<float>(), ld4, Wrbt.as<double>());
        if (st) return st;
        // the inverses cost about twice the factorisation ((4/3) n^3 against (2/3) n^3 on MFMA) and
        // pay off when many right-hand sides are corrected together; FLOWPANEL_LU32_INVERSE=1/0 forces them
        const char *gi = getenv("FLOWPANEL_LU32_INVERSE");
        inv = gi && (gi[0] == '0' || gi[0] == '1') ? gi[0] == '1' : a.nrhs >= 4;
        if (inv) {
            if ((st = dalloc(ctx, YL, (size_t)n4 * ld4 * sizeof(float))) ||
                (st = dalloc(ctx, YU, (size_t)n4 * ld4 * sizeof(float))) || (st = dalloc(ctx, w4, n4 * sizeof(double))) ||
                (st = fp_fill_f32(ctx, YL.as<float>(), n4 * ld4, 0.f)) ||
                (st = fp_fill_f32(ctx, YU.as<float>(), n4 * ld4, 0.f)))
                return st;
        }
        if ((st = fp_getrf_nopiv_f32(ctx, Ar.as<float>(), n4, ld4, info.as<int>(), nw.p,
                                     inv ? YL.as<float>() : nullptr, inv ? YU.as<float>() : nullptr)))
            return st;
        std::vector<int> id(n4);
        for (int64_t i = 0; i < n4; ++i) id[i] = (int)i;
        int hinfo = 0;
        FP_HIP(ctx, hipMemcpyAsync(perm4.p, id.data(), n4 * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (hinfo) {
            Ar.reset();
            YL.reset();
            YU.reset();
            inv = false;
            return FP_OK;
        }
        if (inv) {
            // the triangular inverses take Ar's place; the scratch T reuses Ar after them
            DBuf T;
            if ((st = dalloc(ctx, T, (size_t)n4 * ld4 * sizeof(float)))) return st;
            if ((st = fp_trtri_nopiv_f32(ctx, Ar.as<float>(), n4, ld4, YL.as<float>(), YU.as<float>(), T.as<float>())))
                return st;
            Ar.reset();
        }
        *ok = true;
        rbt = true;
        return FP_OK;
    }

    // the pivoted FP32 LU of the refinement path
    int lu32_pivoted(fp_solve_stats &stats)
    {
        int st;
        rbt = false;
        inv = false;
        diag_ready = nullptr;
        Ar.reset();
        YL.reset();
        YU.reset();
        if ((st = dalloc(ctx, perm, a.n * sizeof(int)))) return st;
        if ((st = factor<float>(ctx, a.A32, a.n, a.ld, perm.as<int>(), &stats.info))) return st;
        return dalloc(ctx, work, fp_getrs_work_bytes(a.n, 1));
    }

    // d (n x k, ld n) ~ A^-1 r (n x k, ld n) by the FP32 factors
    int lu32_correct(const double *rr, double *dd, int64_t k, void *wk, size_t wbytes)
    {
        const int64_t n = a.n;
        int st;
        if (rbt) {
            DBuf tk;
            double *t = t4.as<double>();
            if (k > 1) {
                if ((st = dalloc(ctx, tk, (size_t)n4 * k * sizeof(double)))) return st;
                t = tk.as<double>();
            }
            if ((st = fp_rbt_vec(ctx, rr, n, t, n4, n, k, Wrbt.as<double>(), 1))) return st;
            if (inv) {
                // t <- (-U^-1)(-L^-1) t
                DBuf wk2;
                double *w = w4.as<double>();
                if (k > 1) {
                    if ((st = dalloc(ctx, wk2, (size_t)n4 * k * sizeof(double)))) return st;
                    w = wk2.as<double>();
                    if ((st = fp_gemv_multi<float>(ctx, YL.as<float>(), n4, ld4, t, n4, k, w, n4))) return st;
                    if ((st = fp_gemv_multi<float>(ctx, YU.as<float>(), n4, ld4, w, n4, k, t, n4))) return st;
                } else {
                    if ((st = fp_gemv<float>(ctx, YL.as<float>(), n4, ld4, t, w, 1.0, 0.0))) return st;
                    if ((st = fp_gemv<float>(ctx, YU.as<float>(), n4, ld4, w, t, 1.0, 0.0))) return st;
                }
            } else {
                const bool reuse = wk == diag_ready;
                diag_ready = wk;
                if ((st = fp_getrs<float>(ctx, Ar.as<float>(), n4, ld4, perm4.as<int>(), t, k, n4, wk, wbytes, reuse)))
                    return st;
            }
            return fp_rbt_vec(ctx, t, n4, dd, n, n, k, Wrbt.as<double>(), 0);
        }
        FP_HIP(ctx, hipMemcpyAsync(dd, rr, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        const bool reuse = wk == diag_ready;
        diag_ready = wk;
        return fp_getrs<float>(ctx, a.A32, n, a.ld, perm.as<int>(), dd, k, n, wk, wbytes, reuse);
    }

    int bj_setup(const void *A, int prec, int blk)
    {
        int st;
        if (M && M->early) {
            if ((st = M->finish_early(A, prec, a.ld, o.coarse_block))) return st;
            return bj_check(M->binfo.as<int>(), blk);
        }
        if ((st = dalloc(ctx, info, sizeof(int)))) return st;
        if (o.precond_order == FP_ORDER_SPATIAL) {
            if (!a.geo || a.geo->M != a.n)
                return fp_fail(ctx, FP_ERR_ARG, "a spatial preconditioner order needs the bodies' control points");
            if ((st = spatial_perm(ctx, *a.geo, order))) return st;
            const size_t es = prec == FP_F64 ? sizeof(double) : sizeof(float);
            if ((st = dalloc(ctx, Aperm, (size_t)a.n * a.ld * es))) return st;
            const dim3 g((unsigned)std::min<int64_t>((a.ld + 255) / 256, 8), (unsigned)std::min<int64_t>(a.n, 65535));
            if (prec == FP_F64)
                hipLaunchKernelGGL(k_perm_matrix<double>, g, dim3(256), 0, ctx->stream, (const double *)A, a.ld,
                                   (const int *)order.as<int>(), a.n, Aperm.as<double>());
            else
                hipLaunchKernelGGL(k_perm_matrix<float>, g, dim3(256), 0, ctx->stream, (const float *)A, a.ld,
                                   (const int *)order.as<int>(), a.n, Aperm.as<float>());
            FP_HIP(ctx, hipGetLastError());
            A = Aperm.p;
            PM.reset();
        } else if (o.precond_order != FP_ORDER_INDEX) {
            return fp_fail(ctx, FP_ERR_ARG, "unknown preconditioner order");
        }
        M.reset(new BJPrecond(ctx, a.n, blk));
        if ((st = M->setup(A, prec, a.ld, o.coarse_block, info.as<int>()))) return st;
        return bj_check(info.as<int>(), blk);
    }
    // zero pivots of the block inverses and of the coarse Galerkin matrix (side
    // streams, joined into the main stream by the set-up)
    int bj_check(const int *binfo_dev, int blk)
    {
        int st;
        int hinfo[2] = {0, 0};
        FP_HIP(ctx, hipMemcpyAsync(&hinfo[0], binfo_dev, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        if (M->cinfo.p)
            FP_HIP(ctx, hipMemcpyAsync(&hinfo[1], M->cinfo.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        char buf[192];
        if (hinfo[0]) {
            const int64_t row = hinfo[0] - 1;
            snprintf(buf, sizeof buf,
                     "block-Jacobi preconditioner: zero pivot in diagonal block %lld (global row %lld)",
                     (long long)(row / blk), (long long)row);
            return fp_fail(ctx, FP_ERR_SINGULAR, buf);
        }
        if (hinfo[1]) {
            snprintf(buf, sizeof buf, "coarse correction: zero pivot in the Galerkin matrix at row %d of %lld",
                     hinfo[1] - 1, (long long)((a.n + o.coarse_block - 1) / o.coarse_block));
            return fp_fail(ctx, FP_ERR_SINGULAR, buf);
        }
        if (o.precond_order == FP_ORDER_SPATIAL) {
            PM.reset(new PermPrecond(ctx, M.get(), order.as<int>(), a.n));
            if ((st = PM->init())) return st;
        }
        return FP_OK;
    }

    // factorisation / preconditioner set-up (stats.ms_factor, stats.info)
    int setup(fp_solve_stats &stats)
    {
        const int64_t n = a.n;
        int st;
        if ((st = dalloc(ctx, part, 256 * 4 * sizeof(double)))) return st;
        if ((st = dalloc(ctx, dout, 4 * sizeof(double)))) return st;
        Timer tm(ctx->stream);
        if (a.solver == FP_SOLVER_LU) {
            if (!a.A64) return fp_fail(ctx, FP_ERR_ARG, "internal: no FP64 matrix");
            if ((st = dalloc(ctx, perm, n * sizeof(int)))) return st;
            if ((st = factor<double>(ctx, a.A64, n, a.ld, perm.as<int>(), &stats.info))) return st;
            if ((st = dalloc(ctx, work, fp_getrs_work_bytes(n, 1)))) return st;
        } else if (a.solver == FP_SOLVER_GMRES) {
            // FP64 GMRES on the FP64 operator (solve_gmres!), optional block-Jacobi
            g.reset(new Gmres());
            if ((st = g->init(ctx, n, std::max(1, o.restart)))) return st;
            if (o.precond_block > 0) {
                const int blk = (int)std::min<int64_t>(o.precond_block, n);
                st = a.A64 ? bj_setup(a.A64, FP_F64, blk) : bj_setup(a.A32, FP_F32, blk);
                if (st) return st;
            }
        } else if (a.solver == FP_SOLVER_LU32_IR || a.solver == FP_SOLVER_GMRES_IR) {
            if (!a.A32) return fp_fail(ctx, FP_ERR_ARG, "internal: no FP32 matrix");
            if (a.solver == FP_SOLVER_LU32_IR) {
                bool ok = false;
                const char *pv = getenv("FLOWPANEL_LU32_PIVOTED");
                if (!(pv && pv[0] == '1') && (st = rbt_setup(&ok))) return st;
                if (!ok && (st = lu32_pivoted(stats))) return st;
            } else {
                const int blk = (int)std::min<int64_t>(o.precond_block > 0 ? o.precond_block : 512, n);
                if ((st = bj_setup(a.A32, FP_F32, blk))) return st;
                A32op.reset(new DenseOp<float>(ctx, a.A32, n, a.ld));
                g.reset(new Gmres());
                if ((st = g->init(ctx, n, std::max(1, o.restart)))) return st;
            }
            if ((st = dalloc(ctx, r, n * sizeof(double)))) return st;
            if ((st = dalloc(ctx, d, n * sizeof(double)))) return st;
        } else {
            return fp_fail(ctx, FP_ERR_ARG, "unknown solver");
        }
        stats.ms_factor += tm.stop();
        return FP_OK;
    }

    // x = A^-1 b for one right-hand side (stats: ms_solve, iterations and
    // refinements accumulate; residual is the worst over the calls)
    int solve(const double *b, double *x, fp_solve_stats &stats)
    {
        const int64_t n = a.n;
        int st;
        Timer tm(ctx->stream);
        if (a.solver == FP_SOLVER_LU) {
            FP_HIP(ctx, hipMemcpyAsync(x, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            if ((st = fp_getrs<double>(ctx, a.A64, n, a.ld, perm.as<int>(), x, 1, n, work.p,
                                       fp_getrs_work_bytes(n, 1))))
                return st;
            stats.ms_solve += tm.stop();
            if (a.resid) {  // report the FP64 residual
                DBuf rr;
                if ((st = dalloc(ctx, rr, n * sizeof(double)))) return st;
                if ((st = a.resid->apply(x, rr.as<double>()))) return st;
                if ((st = fp_residual_update(ctx, b, rr.as<double>(), n))) return st;
                const double bn = dnorm_host(ctx, b, n, part.as<double>(), dout.as<double>(), &st);
                const double rn = dnorm_host(ctx, rr.as<double>(), n, part.as<double>(), dout.as<double>(), &st);
                if (st) return st;
                const double rel = bn > 0 ? rn / bn : (bn == 0 ? 0.0 : NAN);
                stats.residual = worst_residual(stats.residual, rel);
                // a direct solve has no tolerance; a NaN in its result is still an error
                if (!std::isfinite(rel)) return not_converged(ctx, rel, "LU solve");
            }
            return FP_OK;
        }
        if (a.solver == FP_SOLVER_GMRES) {
            int its = 0;
            double rel = 0;
            if ((st = g->solve(*a.resid, prec(), b, x, o.tol, o.max_iters, &its, &rel))) return st;
            stats.ms_solve += tm.stop();
            stats.iterations += its;
            stats.residual = worst_residual(stats.residual, rel);
            if (!converged(rel, o.tol)) return not_converged(ctx, rel, "GMRES");
            return FP_OK;
        }
        // refinement paths: FP32 inner solver, FP64 residual and update
        double *rp = r.as<double>(), *dp = d.as<double>();
        const double bn = dnorm_host(ctx, b, n, part.as<double>(), dout.as<double>(), &st);
        if (st) return st;
        FP_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(rp, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        double rel = bn > 0 ? 1.0 : (bn == 0 ? 0.0 : NAN), rnorm = bn;   // ||r|| of the current residual (r = b at first)
        int k = 0, its_total = 0;
        for (;;) {
        while (!converged(rel, o.tol) && std::isfinite(rel) && k < o.max_refine) {
            // correction d ~ A^-1 r in the low-precision solver
            if (a.solver == FP_SOLVER_LU32_IR) {
                void *wk = rbt ? work4.p : work.p;
                if ((st = lu32_correct(rp, dp, 1, wk, fp_getrs_work_bytes(rbt ? n4 : n, 1)))) return st;
            } else {
                int its = 0;
                double irel = 0;
                const int left = std::max(1, o.max_iters - its_total);
                // the correction only has to bring the residual below tol: once
                // rel is small the inner solve may stop earlier
                const double itol = std::max(o.inner_tol, std::min(0.5, 0.3 * o.tol / std::max(rel, 1e-300)));
                if ((st = g->solve(*A32op, prec(), rp, dp, itol, left, &its, &irel, rnorm))) return st;
                its_total += its;
            }
            if ((st = fp_axpby(ctx, 1.0, dp, 1.0, x, n))) return st;
            // FP64 residual
            if ((st = a.resid->apply(x, rp))) return st;
            if ((st = fp_residual_update(ctx, b, rp, n))) return st;
            const double rn = dnorm_host(ctx, rp, n, part.as<double>(), dout.as<double>(), &st);
            if (st) return st;
            rel = rn / bn;
            rnorm = rn;
            ++k;
        }
        // the randomized no-pivot factors failed to converge: start over with the pivoted LU
        if (rbt && !converged(rel, o.tol) && std::isfinite(bn)) {
            if ((st = lu32_pivoted(stats))) return st;
            FP_HIP(ctx, hipMemsetAsync(x, 0, n * sizeof(double), ctx->stream));
            FP_HIP(ctx, hipMemcpyAsync(rp, b, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            rel = bn > 0 ? 1.0 : 0.0;
            rnorm = bn;
            continue;
        }
        break;
        }
        stats.ms_solve += tm.stop();
        stats.refinements += k;
        stats.iterations += its_total;
        stats.residual = worst_residual(stats.residual, rel);
        if (!converged(rel, o.tol)) return not_converged(ctx, rel, "iterative refinement");
        return FP_OK;
    }
    // X = A^-1 B for k right-hand sides (B, X column-major n x k).  LU paths
    // substitute all columns per sweep and refine them together (residuals by
    // a multi-vector GEMV of the stored FP64 matrix); Krylov paths solve the
    // columns one after the other with the same preconditioner.
    int solve_multi(const double *B, double *X, int64_t k, fp_solve_stats &stats)
    {
        const int64_t n = a.n;
        int st;
        diag_ready = nullptr;   // the workspaces below are new
        if (k == 1 || a.solver == FP_SOLVER_GMRES || a.solver == FP_SOLVER_GMRES_IR ||
            (a.solver == FP_SOLVER_LU32_IR && !a.R64)) {
            for (int64_t c = 0; c < k; ++c)
                if ((st = solve(B + c * n, X + c * n, stats))) return st;
            return FP_OK;
        }
        Timer tm(ctx->stream);
        DBuf wk;
        const size_t wbytes = fp_getrs_work_bytes(n, k);
        if ((st = dalloc(ctx, wk, wbytes))) return st;
        FP_HIP(ctx, hipMemcpyAsync(X, B, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        if (a.solver == FP_SOLVER_LU) {
            if ((st = fp_getrs<double>(ctx, a.A64, n, a.ld, perm.as<int>(), X, k, n, wk.p, wbytes))) return st;
            stats.ms_solve += tm.stop();
            return FP_OK;
        }
        // FP32 LU + FP64 refinement of all columns together
        DBuf R, D, nr;
        if ((st = dalloc(ctx, R, (size_t)n * k * sizeof(double))) || (st = dalloc(ctx, D, (size_t)n * k * sizeof(double))) ||
            (st = dalloc(ctx, nr, k * sizeof(double))))
            return st;
        std::vector<double> bn(k), rn(k);
        if ((st = fp_colnorm2(ctx, B, n, n, k, nr.as<double>()))) return st;
        FP_HIP(ctx, hipMemcpyAsync(bn.data(), nr.p, k * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipMemsetAsync(X, 0, (size_t)n * k * sizeof(double), ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(R.p, B, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double rel = 1.0;
        for (int64_t c = 0; c < k; ++c)
            if (!std::isfinite(bn[c])) rel = NAN;
        int it = 0;
        DBuf wk4;
        const size_t wbytes4 = rbt ? fp_getrs_work_bytes(n4, k) : 0;
        if (rbt && (st = dalloc(ctx, wk4, wbytes4))) return st;
        const double rel0 = rel;
        for (;;) {
        while (!converged(rel, o.tol) && std::isfinite(rel) && it < o.max_refine) {
            if ((st = lu32_correct(R.as<double>(), D.as<double>(), k, rbt ? wk4.p : wk.p, rbt ? wbytes4 : wbytes)))
                return st;
            if ((st = fp_axpby(ctx, 1.0, D.as<double>(), 1.0, X, n * k))) return st;
            if ((st = fp_gemv_multi<double>(ctx, a.R64, n, a.ld, X, n, k, R.as<double>(), n))) return st;
            if ((st = fp_residual_update(ctx, B, R.as<double>(), n * k))) return st;
            if ((st = fp_colnorm2(ctx, R.as<double>(), n, n, k, nr.as<double>()))) return st;
            FP_HIP(ctx, hipMemcpyAsync(rn.data(), nr.p, k * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            rel = 0;
            for (int64_t c = 0; c < k; ++c)
                rel = worst_residual(rel, bn[c] > 0 ? std::sqrt(rn[c] / bn[c]) : (rn[c] == 0 ? 0.0 : NAN));
            ++it;
        }
        // the randomized no-pivot factors failed to converge: start over with the pivoted LU
        if (rbt && !converged(rel, o.tol) && std::isfinite(rel0)) {
            if ((st = lu32_pivoted(stats))) return st;
            FP_HIP(ctx, hipMemsetAsync(X, 0, (size_t)n * k * sizeof(double), ctx->stream));
            FP_HIP(ctx, hipMemcpyAsync(R.p, B, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            rel = rel0;
            continue;
        }
        break;
        }
        stats.ms_solve += tm.stop();
        stats.refinements += it;
        stats.residual = worst_residual(stats.residual, rel);
        if (!converged(rel, o.tol)) return not_converged(ctx, rel, "iterative refinement");
        return FP_OK;
    }
};


// host AoS 3 x n -> device SoA
int upload_soa3(fp_ctx *ctx, const double *h, int64_t n, double *d)
{
    std::vector<double> tmp(3 * n);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) tmp[c * n + i] = h[3 * i + c];
    FP_HIP(ctx, hipMemcpy(d, tmp.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
    return FP_OK;
}

int build_horseshoes(fp_ctx *ctx, fp_body *b)
{
    ++b->version;
    if (b->d_hh) { (void)hipFree(b->d_hh); b->d_hh = nullptr; }
    if (b->d_hcol) { (void)hipFree(b->d_hcol); b->d_hcol = nullptr; }
    b->nh = 0;
    if (b->ns == 0) return FP_OK;
    std::vector<double> rows;  // per horseshoe 12 values
    std::vector<int> cols;
    for (int64_t e = 0; e < b->ns; ++e) {
        const int64_t *s = &b->shedding[6 * e];
        const double *da = &b->das[3 * e], *db = &b->dbs[3 * e];
        for (int side = 0; side < 2; ++side) {
            const int64_t p = side == 0 ? s[0] : s[3];
            if (p == -1) continue;
            const int64_t na = side == 0 ? s[1] : s[4], nbv = side == 0 ? s[2] : s[5];
            const int64_t ia = b->cell_idx[b->cell_ptr[p] + na], ib = b->cell_idx[b->cell_ptr[p] + nbv];
            const double *d1 = side == 0 ? da : db, *d2 = side == 0 ? db : da;
            double h[12];
            for (int c = 0; c < 3; ++c) {
                h[c] = b->nodes[3 * ia + c];
                h[3 + c] = b->nodes[3 * ib + c];
                h[6 + c] = d1[c];
                h[9 + c] = d2[c];
            }
            rows.insert(rows.end(), h, h + 12);
            cols.push_back((int)p);
        }
    }
    const int nh = (int)cols.size();
    std::vector<double> soa((size_t)12 * nh);
    for (int h = 0; h < nh; ++h)
        for (int c = 0; c < 12; ++c) soa[(size_t)c * nh + h] = rows[(size_t)h * 12 + c];
    FP_HIP(ctx, hipMalloc((void **)&b->d_hh, soa.size() * sizeof(double)));
    FP_HIP(ctx, hipMalloc((void **)&b->d_hcol, nh * sizeof(int)));
    FP_HIP(ctx, hipMemcpy(b->d_hh, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
    FP_HIP(ctx, hipMemcpy(b->d_hcol, cols.data(), nh * sizeof(int), hipMemcpyHostToDevice));
    b->nh = nh;
    return FP_OK;
}

__global__ void k_aos_add_soa(const double *__restrict__ soa, int64_t n, double *__restrict__ aos)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[3 * i] += soa[i];
    aos[3 * i + 1] += soa[n + i];
    aos[3 * i + 2] += soa[2 * n + i];
}

template <typename T>
__global__ void k_download_cm(const T *__restrict__ src, int64_t n, int64_t ld, double *__restrict__ dst)
{
    // dst column-major: dst[j*n + i] = src[i*ld + j]; one block row per column j
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < n) dst[j * n + i] = (double)src[i * ld + j];
}

// Cp with Kutta averaging (src/FLOWPanel_postprocess.jl:1059-1112)
__global__ void k_cp(const double *__restrict__ U, int64_t n, double uref, double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u = sqrt(U[3 * i] * U[3 * i] + U[3 * i + 1] * U[3 * i + 1] + U[3 * i + 2] * U[3 * i + 2]) / uref;
    out[i] += 1 - u * u;
}

__global__ void k_kutta(const int64_t *__restrict__ sh, int64_t ns, int mode, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int64_t e = 0; e < ns; ++e) {
        const int64_t pi = sh[6 * e], pj = sh[6 * e + 3];
        if (mode == FP_KUTTA_REFERENCE) {
            if (pj != -1) {
                const double ave = (out[pi] + out[pi + 1] + out[pj] + out[pj - 1]) / 4;
                out[pi] = ave; out[pi + 1] = ave; out[pj] = ave; out[pj - 1] = ave;
            } else {
                const double ave = (out[pi] + out[pi + 1]) / 2;
                out[pi] = ave; out[pi + 1] = ave;
            }
        } else if (pj != -1) {
            const double ave = (out[pi] + out[pj]) / 2;
            out[pi] = ave; out[pj] = ave;
        }
    }
}


// ----------------------------------------------------------------------------
// Batch of independent bodies (each its own system, no mutual influence).
// One persistent queue launch assembles every body's vortex rings; each body's
// wake columns, factorisation and refinement sweeps then run on one of
// BATCH_STREAMS parallel streams.  The whole sequence is captured once into a
// hipGraph and replayed; bodies are re-captured only when their device
// buffers change.
constexpr int BATCH_STREAMS = 4;

struct BatchBody {
    fp_body *b = nullptr;
    uint64_t version = 0;
    int64_t n = 0, ld = 0;
    int stream = 0;
    Targets t;
    DBuf A64, A32, rhs, x, r, d, perm, info, fwork, swork;
    std::unique_ptr<MatrixFreeOp> mf;
    // FP32 refinement: LU of the randomized matrix without pivoting (`pivoted` once it
    // met a bad pivot or did not converge: the tournament-pivoted LU of A32 from then on)
    bool pivoted = false;
    int64_t n4 = 0, ld4 = 0;
    DBuf Ar, W, perm4, work4, nwork, t4;
    // the getrs workspace (work4, or swork once pivoted) holds the diagonal-block
    // inverses of the current factors
    bool diag_ready = false;
};

struct BatchImpl {
    fp_ctx *ctx = nullptr;
    fp_solve_opts o{};
    std::vector<std::unique_ptr<BatchBody>> bodies;
    int nvp = 4, tm = 64, blocks = 0, sweeps = 3;
    double tune_ms[9] = {0};
    int tune_tm[9] = {0}, tune_blocks[9] = {0}, ntune = 0;
    int64_t total = 0;
    DBuf items, counter, norms;
    hipStream_t streams[BATCH_STREAMS] = {};
    hipEvent_t fork = nullptr, join[BATCH_STREAMS] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_nodes = 0;
    bool recapture = false;   // a body switched to the pivoted LU since the last capture
    // batched FP32 LU of the randomized bodies (FLOWPANEL_BATCH_LU=bodies: each body its own)
    bool batched_lu = true;
    DBuf lu_tab;
    std::vector<int64_t> lu_n;
    hipEvent_t fork2 = nullptr;
    bool lu32() const { return o.solver == FP_SOLVER_LU32_IR; }
    bool stored() const { return o.solver == FP_SOLVER_LU || o.residual_mode == FP_RESID_STORED; }

    ~BatchImpl()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &s : streams)
            if (s) (void)hipStreamDestroy(s);
        if (fork) (void)hipEventDestroy(fork);
        if (fork2) (void)hipEventDestroy(fork2);
        for (auto &e : join)
            if (e) (void)hipEventDestroy(e);
    }

    int alloc()
    {
        int st;
        for (int q = 0; q < BATCH_STREAMS; ++q) {
            FP_HIP(ctx, hipStreamCreateWithFlags(&streams[q], hipStreamNonBlocking));
            FP_HIP(ctx, hipEventCreateWithFlags(&join[q], hipEventDisableTiming));
        }
        FP_HIP(ctx, hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        if ((st = dalloc(ctx, counter, 16)) || (st = dalloc(ctx, norms, 2 * bodies.size() * sizeof(double))))
            return st;
        // balance the factorisations (n^3) over the streams, largest first
        std::vector<int> order(bodies.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return bodies[a]->n > bodies[b]->n; });
        double load[BATCH_STREAMS] = {0};
        for (int k : order) {
            int q = (int)(std::min_element(load, load + BATCH_STREAMS) - load);
            bodies[k]->stream = q;
            load[q] += std::pow((double)bodies[k]->n, 3);
        }
        for (auto &bb : bodies) {
            const int64_t n = bb->n, ld = bb->ld;
            if (stored() && (st = dalloc(ctx, bb->A64, (size_t)n * ld * sizeof(double)))) return st;
            if (lu32() && (st = dalloc(ctx, bb->A32, (size_t)n * ld * sizeof(float)))) return st;
            if (lu32()) {
                bb->n4 = fp_rbt_size(n);
                bb->ld4 = round_ld(bb->n4);
                if ((st = dalloc(ctx, bb->Ar, (size_t)bb->n4 * bb->ld4 * sizeof(float))) ||
                    (st = dalloc(ctx, bb->W, fp_rbt_weights_bytes(n))) || (st = dalloc(ctx, bb->perm4, bb->n4 * sizeof(int))) ||
                    (st = dalloc(ctx, bb->work4, fp_getrs_work_bytes(bb->n4, 1))) ||
                    (st = dalloc(ctx, bb->nwork, fp_getrf_nopiv_work_bytes())) ||
                    (st = dalloc(ctx, bb->t4, bb->n4 * sizeof(double))))
                    return st;
                if ((st = fp_rbt_weights(ctx, n, RBT_SEED, bb->W.as<double>()))) return st;
                std::vector<int> id(bb->n4);
                for (int64_t i = 0; i < bb->n4; ++i) id[i] = (int)i;
                FP_HIP(ctx, hipMemcpy(bb->perm4.p, id.data(), bb->n4 * sizeof(int), hipMemcpyHostToDevice));
            }
            if ((st = dalloc(ctx, bb->rhs, n * sizeof(double))) || (st = dalloc(ctx, bb->x, n * sizeof(double))) ||
                (st = dalloc(ctx, bb->r, n * sizeof(double))) || (st = dalloc(ctx, bb->d, n * sizeof(double))) ||
                (st = dalloc(ctx, bb->perm, n * sizeof(int))) || (st = dalloc(ctx, bb->info, sizeof(int))) ||
                (st = dalloc(ctx, bb->fwork, fp_getrf_work_bytes(n, lu32() ? sizeof(float) : sizeof(double)))) ||
                (st = dalloc(ctx, bb->swork, fp_getrs_work_bytes(n, 1))))
                return st;
        }
        return FP_OK;
    }

    // (re)build the device table of the assembly queue
    int build_items()
    {
        std::vector<AssemblyItem> it;
        total = 0;
        for (auto &bb : bodies) {
            AssemblyItem a;
            a.pan = panel_view(bb->b);
            a.tv = bb->t.tv;
            a.G = stored() ? bb->A64.p : bb->A32.p;
            a.ld = bb->ld;
            a.tile0 = total;
            int nbx;
            int64_t nt;
            fp_assembly_queue_tiles(bb->n, bb->n, tm, &nbx, &nt);
            a.nbx = nbx;
            total += nt;
            it.push_back(a);
        }
        int st = dalloc(ctx, items, it.size() * sizeof(AssemblyItem));
        if (st) return st;
        FP_HIP(ctx, hipMemcpy(items.p, it.data(), it.size() * sizeof(AssemblyItem), hipMemcpyHostToDevice));
        return FP_OK;
    }

    int assemble_queue()
    {
        return fp_launch_assemble_queue(ctx, items.as<AssemblyItem>(), (int)bodies.size(), total, nvp, tm,
                                        stored() ? FP_F64 : FP_F32, counter.as<int>(), blocks);
    }

    // time the queue assembly for each tile shape and workgroup count; keep the fastest
    int autotune()
    {
        const int tms[3] = {32, 64, 128}, per_cu[3] = {2, 4, 8};
        double best = 1e300;
        int btm = 64, bbl = 4 * ctx->num_cu;
        ntune = 0;
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                tm = tms[a];
                blocks = per_cu[c] * ctx->num_cu;
                int st = build_items();
                if (st) return st;
                if ((st = assemble_queue())) return st;  // warm
                Timer t(ctx->stream);
                if ((st = assemble_queue())) return st;
                const double ms = t.stop();
                tune_ms[ntune] = ms;
                tune_tm[ntune] = tm;
                tune_blocks[ntune] = blocks;
                ++ntune;
                if (ms < best) { best = ms; btm = tm; bbl = blocks; }
            }
        tm = btm;
        blocks = bbl;
        return build_items();
    }

    // the work of one body after the queue assembly, on the current ctx->stream
    // phase ISSUE_ALL: a body's whole solve; with the batched LU, ISSUE_PRE stops a
    // randomized body after U^T A V (its factorisation joins the batched one) and
    // ISSUE_POST issues its refinement (pivoted bodies do everything in ISSUE_PRE)
    enum { ISSUE_ALL, ISSUE_PRE, ISSUE_POST };
    int issue_body(BatchBody &bb, int phase = ISSUE_ALL)
    {
        fp_body *b = bb.b;
        const int64_t n = bb.n, ld = bb.ld;
        int st;
        double *x = bb.x.as<double>(), *r = bb.r.as<double>(), *rhs = bb.rhs.as<double>();
        double *nb = norms.as<double>();
        size_t k = 0;
        while (bodies[k].get() != &bb) ++k;
        const bool joint = lu32() && !bb.pivoted;   // takes part in the batched LU
        if (phase == ISSUE_POST) {
            if (!joint) return FP_OK;
            return refine_body(bb, k);
        }
        if (stored()) {
            if ((st = fp_launch_assemble_wake<double>(ctx, b, bb.t.tv, bb.A64.as<double>(), ld, 0))) return st;
        } else {
            if ((st = fp_launch_assemble_wake<float>(ctx, b, bb.t.tv, bb.A32.as<float>(), ld, 0))) return st;
        }
        if (lu32() && stored() && bb.pivoted && (st = fp_mat_narrow(ctx, bb.A64.as<double>(), bb.A32.as<float>(), n, ld)))
            return st;
        if ((st = fp_rhs_kernel(ctx, b, rhs))) return st;
        if ((st = fp_colnorm2(ctx, rhs, n, n, 1, nb + k))) return st;
        if (!lu32()) {
            // FP64 LU + one substitution; the FP64 residual is reported
            if ((st = fp_getrf<double>(ctx, bb.A64.as<double>(), n, ld, bb.perm.as<int>(), bb.info.as<int>(), bb.fwork.p,
                                       fp_getrf_work_bytes(n, sizeof(double)))))
                return st;
            FP_HIP(ctx, hipMemcpyAsync(x, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            if ((st = fp_getrs<double>(ctx, bb.A64.as<double>(), n, ld, bb.perm.as<int>(), x, 1, n, bb.swork.p,
                                       fp_getrs_work_bytes(n, 1))))
                return st;
            if ((st = bb.mf->apply(x, r))) return st;
            if ((st = fp_residual_update(ctx, rhs, r, n))) return st;
        } else {
            if (bb.pivoted) {
                if ((st = fp_getrf<float>(ctx, bb.A32.as<float>(), n, ld, bb.perm.as<int>(), bb.info.as<int>(),
                                          bb.fwork.p, fp_getrf_work_bytes(n, sizeof(float)))))
                    return st;
                bb.diag_ready = false;
            } else if (phase == ISSUE_PRE) {
                return rbt_matrix(bb);   // factored by the batched LU, refined in ISSUE_POST
            } else if ((st = rbt_factor(bb))) {
                return st;
            }
            return refine_body(bb, k);
        }
        if ((st = fp_colnorm2(ctx, r, n, n, 1, nb + bodies.size() + k))) return st;
        FP_HIP(ctx, hipMemcpyAsync(b->d_strength, x, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        return FP_OK;
    }

    // the FP32 refinement sweeps of body k from x = 0, then its residual norm and strengths
    int refine_body(BatchBody &bb, size_t k)
    {
        const int64_t n = bb.n;
        int st;
        double *x = bb.x.as<double>(), *r = bb.r.as<double>(), *rhs = bb.rhs.as<double>();
        if ((st = fp_fill(ctx, x, n, 0.0))) return st;
        FP_HIP(ctx, hipMemcpyAsync(r, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        for (int s = 0; s < sweeps; ++s)
            if ((st = sweep(bb))) return st;
        if ((st = fp_colnorm2(ctx, r, n, n, 1, norms.as<double>() + bodies.size() + k))) return st;
        FP_HIP(ctx, hipMemcpyAsync(bb.b->d_strength, x, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        return FP_OK;
    }

    // U^T A V of the FP64 (stored) or FP32 AIC
    int rbt_matrix(BatchBody &bb)
    {
        bb.diag_ready = false;
        return stored() ? fp_rbt_matrix<double>(ctx, bb.A64.as<double>(), bb.n, bb.ld, bb.Ar.as<float>(), bb.ld4,
                                                bb.W.as<double>())
                        : fp_rbt_matrix<float>(ctx, bb.A32.as<float>(), bb.n, bb.ld, bb.Ar.as<float>(), bb.ld4,
                                               bb.W.as<double>());
    }

    // U^T A V, then its LU without pivoting
    int rbt_factor(BatchBody &bb)
    {
        int st = rbt_matrix(bb);
        if (st) return st;
        return fp_getrf_nopiv_f32(ctx, bb.Ar.as<float>(), bb.n4, bb.ld4, bb.info.as<int>(), bb.nwork.p);
    }

    // the device table of the batched LU: every randomized (not pivoted) body's U^T A V
    int build_lu_table()
    {
        std::vector<fp_lu_desc> t;
        lu_n.clear();
        for (auto &bb : bodies)
            if (!bb->pivoted) {
                t.push_back(fp_lu_desc{bb->Ar.as<float>(), bb->ld4, bb->n4, (float *)bb->nwork.p, bb->info.as<int>()});
                lu_n.push_back(bb->n4);
            }
        if (t.empty()) return FP_OK;
        int st;
        if (!lu_tab.p || lu_tab.bytes < t.size() * sizeof(fp_lu_desc)) {
            lu_tab.reset();
            if ((st = dalloc(ctx, lu_tab, t.size() * sizeof(fp_lu_desc)))) return st;
        }
        FP_HIP(ctx, hipMemcpy(lu_tab.p, t.data(), t.size() * sizeof(fp_lu_desc), hipMemcpyHostToDevice));
        return FP_OK;
    }

    // one refinement sweep: d = A32^-1 r, x += d, r = b - A x (FP64)
    int sweep(BatchBody &bb)
    {
        const int64_t n = bb.n, ld = bb.ld;
        double *x = bb.x.as<double>(), *r = bb.r.as<double>(), *dd = bb.d.as<double>();
        int st;
        const bool reuse = bb.diag_ready;
        bb.diag_ready = true;
        if (!bb.pivoted) {
            double *t = bb.t4.as<double>();
            if ((st = fp_rbt_vec(ctx, r, n, t, bb.n4, n, 1, bb.W.as<double>(), 1))) return st;
            if ((st = fp_getrs<float>(ctx, bb.Ar.as<float>(), bb.n4, bb.ld4, bb.perm4.as<int>(), t, 1, bb.n4,
                                      bb.work4.p, fp_getrs_work_bytes(bb.n4, 1), reuse)))
                return st;
            if ((st = fp_rbt_vec(ctx, t, bb.n4, dd, n, n, 1, bb.W.as<double>(), 0))) return st;
        } else {
            FP_HIP(ctx, hipMemcpyAsync(dd, r, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            if ((st = fp_getrs<float>(ctx, bb.A32.as<float>(), n, ld, bb.perm.as<int>(), dd, 1, n, bb.swork.p,
                                      fp_getrs_work_bytes(n, 1), reuse)))
                return st;
        }
        if ((st = fp_axpby(ctx, 1.0, dd, 1.0, x, n))) return st;
        if (stored()) st = fp_gemv<double>(ctx, bb.A64.as<double>(), n, ld, x, r, 1.0, 0.0);
        else st = bb.mf->apply(x, r);
        if (st) return st;
        return fp_residual_update(ctx, bb.rhs.as<double>(), r, n);
    }

    // eager re-solve of one body with the pivoted FP32 LU (stays pivoted; the graph is
    // re-captured with it at the next solve)
    int fallback(BatchBody &bb)
    {
        const int64_t n = bb.n, ld = bb.ld;
        int st;
        bb.pivoted = true;
        recapture = true;
        size_t k = 0;
        while (bodies[k].get() != &bb) ++k;
        if (stored() && (st = fp_mat_narrow(ctx, bb.A64.as<double>(), bb.A32.as<float>(), n, ld))) return st;
        if (!stored()) {
            // the FP32 AIC was assembled in place; the randomized copy left it intact
        }
        if ((st = fp_getrf<float>(ctx, bb.A32.as<float>(), n, ld, bb.perm.as<int>(), bb.info.as<int>(), bb.fwork.p,
                                  fp_getrf_work_bytes(n, sizeof(float)))))
            return st;
        bb.diag_ready = false;
        if ((st = fp_fill(ctx, bb.x.as<double>(), n, 0.0))) return st;
        FP_HIP(ctx, hipMemcpyAsync(bb.r.p, bb.rhs.p, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        for (int s = 0; s < sweeps; ++s)
            if ((st = sweep(bb))) return st;
        if ((st = fp_colnorm2(ctx, bb.r.as<double>(), n, n, 1, norms.as<double>() + bodies.size() + k))) return st;
        FP_HIP(ctx, hipMemcpyAsync(bb.b->d_strength, bb.x.p, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        int info = 0;
        FP_HIP(ctx, hipMemcpyAsync(&info, bb.info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (info) {
            char buf[128];
            snprintf(buf, sizeof buf, "singular matrix: zero pivot at column %d", info);
            return fp_fail(ctx, FP_ERR_SINGULAR, buf);
        }
        return FP_OK;
    }

    int capture()
    {
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        const bool prof = ctx->prof;
        hipStream_t main = ctx->stream;
        int st = FP_OK;
        const char *bl = getenv("FLOWPANEL_BATCH_LU");
        batched_lu = !(bl && !strcmp(bl, "bodies"));
        const bool joint = lu32() && batched_lu;
        if (joint) {
            if ((st = build_lu_table())) return st;
            if (!fork2) FP_HIP(ctx, hipEventCreateWithFlags(&fork2, hipEventDisableTiming));
        }
        ctx->prof = false;  // no profiling events inside the graph
        hipError_t e = hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) { ctx->prof = prof; return fp_hip_fail(ctx, e, "hipStreamBeginCapture"); }
        st = assemble_queue();
        if (!st && (e = hipEventRecord(fork, main)) != hipSuccess) st = fp_hip_fail(ctx, e, "hipEventRecord");
        for (int q = 0; q < BATCH_STREAMS && !st; ++q)
            if ((e = hipStreamWaitEvent(streams[q], fork, 0)) != hipSuccess) st = fp_hip_fail(ctx, e, "hipStreamWaitEvent");
        for (auto &bb : bodies) {
            if (st) break;
            ctx->stream = streams[bb->stream];
            st = issue_body(*bb, joint ? ISSUE_PRE : ISSUE_ALL);
            ctx->stream = main;
        }
        if (joint && !st && !lu_n.empty()) {
            // every stream's U^T A V joins the main stream, the batched LU runs there, and the
            // refinements fork back onto the bodies' streams
            for (int q = 0; q < BATCH_STREAMS && !st; ++q) {
                if ((e = hipEventRecord(join[q], streams[q])) != hipSuccess) st = fp_hip_fail(ctx, e, "hipEventRecord");
                else if ((e = hipStreamWaitEvent(main, join[q], 0)) != hipSuccess)
                    st = fp_hip_fail(ctx, e, "hipStreamWaitEvent");
            }
            // (releasing each body's refinement right after its last outer block, beside the larger
            // bodies' steps, measured 112.8 against 108.7 ms per replay: the sweeps slow the chain)
            if (!st) st = fp_getrf_nopiv_f32_batched(ctx, lu_tab.as<fp_lu_desc>(), lu_n.data(), (int)lu_n.size());
            if (!st && (e = hipEventRecord(fork2, main)) != hipSuccess) st = fp_hip_fail(ctx, e, "hipEventRecord");
            for (int q = 0; q < BATCH_STREAMS && !st; ++q)
                if ((e = hipStreamWaitEvent(streams[q], fork2, 0)) != hipSuccess)
                    st = fp_hip_fail(ctx, e, "hipStreamWaitEvent");
            for (auto &bb : bodies) {
                if (st) break;
                ctx->stream = streams[bb->stream];
                st = issue_body(*bb, ISSUE_POST);
                ctx->stream = main;
            }
        }
        for (int q = 0; q < BATCH_STREAMS && !st; ++q) {
            if ((e = hipEventRecord(join[q], streams[q])) != hipSuccess) st = fp_hip_fail(ctx, e, "hipEventRecord");
            else if ((e = hipStreamWaitEvent(main, join[q], 0)) != hipSuccess)
                st = fp_hip_fail(ctx, e, "hipStreamWaitEvent");
        }
        ctx->stream = main;
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(main, &g);
        ctx->prof = prof;
        if (st) { if (g) (void)hipGraphDestroy(g); return st; }
        if (e != hipSuccess) return fp_hip_fail(ctx, e, "hipStreamEndCapture");
        graph = g;
        recapture = false;
        size_t nn = 0;
        if (hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess) graph_nodes = (int)nn;
        FP_HIP(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        for (auto &bb : bodies) bb->version = bb->b->version;
        return FP_OK;
    }
};

}  // namespace

// ============================================================================
void fp_pool_put(fp_ctx *ctx, void *p, size_t bytes)
{
    ctx->pool.push_back({p, bytes});
    ctx->pool_bytes += bytes;
    while (ctx->pool_bytes > FP_POOL_MAX && !ctx->pool.empty()) {
        (void)hipFree(ctx->pool.front().p);
        ctx->pool_bytes -= ctx->pool.front().bytes;
        ctx->pool.erase(ctx->pool.begin());
    }
}

void fp_pool_trim(fp_ctx *ctx)
{
    for (auto &b : ctx->pool) (void)hipFree(b.p);
    ctx->pool.clear();
    ctx->pool_bytes = 0;
}

fp_ctx::~fp_ctx() { fp_pool_trim(this); }

extern "C" {

void fp_solve_opts_default(fp_solve_opts *o, int solver)
{
    if (!o) return;
    o->solver = solver;
    o->restart = 50;
    o->max_iters = 2000;
    o->max_refine = 20;
    o->precond_block = 0;
    o->coarse_block = 0;
    o->residual_mode = FP_RESID_STORED;
    o->precond_order = FP_ORDER_INDEX;
    o->inner_tol = 1e-5;
    o->tol = 1e-12;
    if (solver == FP_SOLVER_GMRES) {
        o->tol = std::sqrt(DBL_EPSILON);  // Krylov.gmres atol = rtol = sqrt(eps) (src/FLOWPanel_linearsolver.jl:136-139)
    } else if (solver == FP_SOLVER_GMRES_IR) {
        o->precond_block = 512;
    }
}

int fp_ctx_create(int device, fp_ctx **out)
{
    if (!out) return FP_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FP_ERR_DEVICE;
    if (device < 0 || device >= ndev) return FP_ERR_ARG;
    std::unique_ptr<fp_ctx> c(new fp_ctx());
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return FP_ERR_DEVICE;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return FP_ERR_DEVICE;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) == hipSuccess) c->num_cu = p.multiProcessorCount;
    fp_noalloc_init(c.get());
    *out = c.release();
    return FP_OK;
}

int fp_ctx_trim(fp_ctx *ctx)
{
    if (!ctx) return FP_ERR_ARG;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->side) FP_HIP(ctx, hipStreamSynchronize(ctx->side));
    if (ctx->side2) FP_HIP(ctx, hipStreamSynchronize(ctx->side2));
    if (ctx->side3) FP_HIP(ctx, hipStreamSynchronize(ctx->side3));
    fp_pool_trim(ctx);
    return FP_OK;
}

void fp_ctx_destroy(fp_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    for (hipStream_t sd : {ctx->side, ctx->side2, ctx->side3})
        if (sd) {
            (void)hipStreamSynchronize(sd);
            (void)hipStreamDestroy(sd);
        }
    for (hipEvent_t e : {ctx->gj_fork, ctx->gj_join, ctx->cgj_fork, ctx->cgj_join})
        if (e) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    for (auto &r : ctx->prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->prof_pool) (void)hipEventDestroy(e);
    delete ctx;
}

const char *fp_ctx_last_error(const fp_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static void prof_fold(fp_ctx *ctx)
{
    if (ctx->prof_pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->prof_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto &acc = ctx->prof_acc[r.name];
            acc.first += 1;
            acc.second += ms;
        }
        ctx->prof_pool.push_back(r.a);
        ctx->prof_pool.push_back(r.b);
    }
    ctx->prof_pending.clear();
}

int fp_ctx_profile(fp_ctx *ctx, int enable)
{
    if (!ctx) return FP_ERR_ARG;
    prof_fold(ctx);
    if (enable) ctx->prof_acc.clear();
    ctx->prof = enable != 0;
    ctx->prof_only.clear();
    return FP_OK;
}

int fp_ctx_profile_only(fp_ctx *ctx, const char *kernel)
{
    if (!ctx) return FP_ERR_ARG;
    prof_fold(ctx);
    ctx->prof_acc.clear();
    ctx->prof = true;
    ctx->prof_only = kernel ? kernel : "";
    return FP_OK;
}

int fp_ctx_copy_calibrate(fp_ctx *ctx, int64_t bytes)
{
    if (!ctx || bytes <= 0) return FP_ERR_ARG;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    DBuf a, b;
    int st;
    if ((st = dalloc(ctx, a, bytes)) || (st = dalloc(ctx, b, bytes))) return st;
    FP_HIP(ctx, hipMemsetAsync(a.p, 0, bytes, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

int fp_ctx_profile_get(fp_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms)
{
    if (!ctx || !kernel || !launches || !total_ms) return FP_ERR_ARG;
    prof_fold(ctx);
    auto it = ctx->prof_acc.find(kernel);
    *launches = it == ctx->prof_acc.end() ? 0 : it->second.first;
    *total_ms = it == ctx->prof_acc.end() ? 0.0 : it->second.second;
    return FP_OK;
}

int fp_ctx_synchronize(fp_ctx *ctx)
{
    if (!ctx) return FP_ERR_ARG;
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

int fp_body_create(fp_ctx *ctx, const double *nodes, int64_t nnodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                   int64_t ncells, int index_base, int element, const int64_t *shedding, int64_t ns,
                   double cp_offset, int charlen_kind, double kernel_offset, double kernel_cutoff, fp_body **out)
{
    if (!ctx || !out) return FP_ERR_ARG;
    *out = nullptr;
    if (!nodes || !cell_ptr || !cell_idx || nnodes <= 0 || ncells <= 0)
        return fp_fail(ctx, FP_ERR_ARG, "empty or null geometry");
    if (index_base != 0 && index_base != 1) return fp_fail(ctx, FP_ERR_ARG, "index_base must be 0 or 1");
    if (element != FP_ELEM_SOURCE && element != FP_ELEM_VORTEXRING)
        return fp_fail(ctx, FP_ERR_ARG, "unknown element type");
    if (charlen_kind < 0 || charlen_kind > 3) return fp_fail(ctx, FP_ERR_ARG, "unknown characteristic length");
    if (ns < 0 || (ns > 0 && !shedding)) return fp_fail(ctx, FP_ERR_ARG, "null shedding matrix");
    if (element == FP_ELEM_SOURCE && ns > 0)
        return fp_fail(ctx, FP_ERR_ARG, "a constant-source body sheds no wake");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<fp_body> b(new fp_body());
    b->ctx = ctx;
    b->element = element;
    b->N = ncells;
    b->kernel_offset = kernel_offset;
    b->kernel_cutoff = kernel_cutoff;
    b->cp_offset = cp_offset;
    b->charlen = charlen_kind;
    b->nodes.assign(nodes, nodes + 3 * nnodes);
    b->cell_ptr.resize(ncells + 1);
    const int64_t p0 = cell_ptr[0];
    int maxnv = 0;
    for (int64_t c = 0; c <= ncells; ++c) b->cell_ptr[c] = cell_ptr[c] - p0;
    for (int64_t c = 0; c < ncells; ++c) {
        const int64_t nv = b->cell_ptr[c + 1] - b->cell_ptr[c];
        if (nv < 3 || nv > FP_MAX_NV) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "panels need 3..8 vertices");
        maxnv = std::max<int>(maxnv, (int)nv);
    }
    const int64_t nidx = b->cell_ptr[ncells];
    b->cell_idx.resize(nidx);
    for (int64_t k = 0; k < nidx; ++k) {
        const int64_t v = cell_idx[k] - index_base;
        if (v < 0 || v >= nnodes) return fp_fail(ctx, FP_ERR_ARG, "cell vertex index out of range");
        b->cell_idx[k] = v;
    }
    b->ns = ns;
    b->shedding.resize(6 * ns);
    for (int64_t e = 0; e < ns; ++e) {
        for (int k = 0; k < 6; ++k) {
            const int64_t v = shedding[6 * e + k];
            b->shedding[6 * e + k] = (k == 3 && v == -1) ? -1 : v - index_base;
        }
        const int64_t *s = &b->shedding[6 * e];
        for (int side = 0; side < 2; ++side) {
            const int64_t p = s[3 * side];
            if (side == 1 && p == -1) continue;
            if (p < 0 || p >= ncells) return fp_fail(ctx, FP_ERR_ARG, "shedding panel index out of range");
            const int64_t nv = b->cell_ptr[p + 1] - b->cell_ptr[p];
            for (int k = 1; k < 3; ++k)
                if (s[3 * side + k] < 0 || s[3 * side + k] >= nv)
                    return fp_fail(ctx, FP_ERR_ARG, "shedding vertex index out of range");
        }
    }
    b->nvp = maxnv <= 4 ? 4 : 8;
    const int nvp = b->nvp;
    const int64_t N = ncells;
    std::vector<double> pv((size_t)3 * nvp * N);
    std::vector<int> nvv(N);
    for (int64_t j = 0; j < N; ++j) {
        const int64_t *pc = &b->cell_idx[b->cell_ptr[j]];
        const int nv = (int)(b->cell_ptr[j + 1] - b->cell_ptr[j]);
        nvv[j] = nv;
        for (int k = 0; k < nvp; ++k) {
            const int64_t v = pc[k < nv ? k : 0];
            for (int d = 0; d < 3; ++d) pv[(size_t)(3 * k + d) * N + j] = b->nodes[3 * v + d];
        }
    }
    FP_HIP(ctx, hipMalloc((void **)&b->d_pv, pv.size() * sizeof(double)));
    FP_HIP(ctx, hipMalloc((void **)&b->d_nv, N * sizeof(int)));
    FP_HIP(ctx, hipMalloc((void **)&b->d_hs, (size_t)FP_HS_STRIDE(nvp) * N * sizeof(double)));
    FP_HIP(ctx, hipMalloc((void **)&b->d_geo, (size_t)FP_GEO_STRIDE * N * sizeof(double)));
    FP_HIP(ctx, hipMalloc((void **)&b->d_uinf, (size_t)3 * N * sizeof(double)));
    FP_HIP(ctx, hipMalloc((void **)&b->d_strength, N * sizeof(double)));
    FP_HIP(ctx, hipMemcpy(b->d_pv, pv.data(), pv.size() * sizeof(double), hipMemcpyHostToDevice));
    FP_HIP(ctx, hipMemcpy(b->d_nv, nvv.data(), N * sizeof(int), hipMemcpyHostToDevice));
    FP_HIP(ctx, hipMemset(b->d_strength, 0, N * sizeof(double)));
    int st = fp_launch_geometry(ctx, b.get());
    if (st) return st;
    if ((st = fp_body_build_edge_blocks(ctx, b.get()))) return st;
    if ((st = fp_body_build_trails(ctx, b.get()))) return st;
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = b.release();
    return FP_OK;
}

void fp_body_destroy(fp_body *b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (void *p : {(void *)b->d_pv, (void *)b->d_nv, (void *)b->d_hs, (void *)b->d_geo, (void *)b->d_uinf,
                    (void *)b->d_hh, (void *)b->d_hcol, (void *)b->d_strength, (void *)b->d_nbr, (void *)b->d_sh,
                    (void *)b->d_eb, (void *)b->d_pb, (void *)b->d_tr_tab, (void *)b->d_tr})
        if (p) (void)hipFree(p);
    delete b;
}

int64_t fp_body_ncells(const fp_body *b) { return b ? b->N : -1; }
int64_t fp_body_nsheddings(const fp_body *b) { return b ? b->ns : -1; }

int fp_body_geometry(fp_ctx *ctx, const fp_body *b, double *normals, double *cps, double *areas)
{
    if (!ctx || !b || b->ctx != ctx) return FP_ERR_ARG;
    const int64_t N = b->N;
    std::vector<double> g((size_t)FP_GEO_STRIDE * N);
    FP_HIP(ctx, hipMemcpyAsync(g.data(), b->d_geo, g.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t j = 0; j < N; ++j) {
        for (int c = 0; c < 3; ++c) {
            if (normals) normals[3 * j + c] = g[c * N + j];
            if (cps) cps[3 * j + c] = g[(3 + c) * N + j];
        }
        if (areas) areas[j] = g[6 * N + j];
    }
    return FP_OK;
}

int fp_body_set_freestream(fp_ctx *ctx, fp_body *b, const double *Uinfs, const double *Das, const double *Dbs)
{
    if (!ctx || !b || b->ctx != ctx) return FP_ERR_ARG;
    if (!Uinfs) return fp_fail(ctx, FP_ERR_SHAPE, "Invalid Uinfs; expected size (3, ncells)");
    if (b->ns > 0 && (!Das || !Dbs)) return fp_fail(ctx, FP_ERR_SHAPE, "Invalid Das/Dbs; expected size (3, nsheddings)");
    for (int64_t e = 0; e < b->ns; ++e)
        if (!check_unit(Das + 3 * e) || !check_unit(Dbs + 3 * e))
            return fp_fail(ctx, FP_ERR_NONUNIT_DIRECTION, "Found non-unitary semi-infinite direction");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    int st = upload_soa3(ctx, Uinfs, b->N, b->d_uinf);
    if (st) return st;
    b->das.assign(Das ? Das : Uinfs, Das ? Das + 3 * b->ns : Uinfs);
    b->dbs.assign(Dbs ? Dbs : Uinfs, Dbs ? Dbs + 3 * b->ns : Uinfs);
    if ((st = build_horseshoes(ctx, b))) return st;
    b->has_uinf = true;
    b->has_wake = true;
    return FP_OK;
}

int fp_body_set_strength(fp_ctx *ctx, fp_body *b, const double *s)
{
    if (!ctx || !b || b->ctx != ctx || !s) return FP_ERR_ARG;
    if (b->ns > 0 && !b->has_wake) return fp_fail(ctx, FP_ERR_ARG, "set the wake directions first");
    FP_HIP(ctx, hipMemcpy(b->d_strength, s, b->N * sizeof(double), hipMemcpyHostToDevice));
    b->solved = true;
    return FP_OK;
}

int fp_body_get_strength(fp_ctx *ctx, const fp_body *b, double *s)
{
    if (!ctx || !b || b->ctx != ctx || !s) return FP_ERR_ARG;
    FP_HIP(ctx, hipMemcpyAsync(s, b->d_strength, b->N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

static int assemble_into(fp_ctx *ctx, fp_body *const *bodies, int nb, const TargetView &tv, int precision, void *A,
                         int64_t ld, float *A32 = nullptr)
{
    int64_t col = 0;
    for (int k = 0; k < nb; ++k) {
        const fp_body *b = bodies[k];
        if (b->element == FP_ELEM_VORTEXRING && b->ns > 0 && !b->has_wake)
            return fp_fail(ctx, FP_ERR_ARG, "wake directions not set (call fp_body_set_freestream)");
        if (b->kernel_offset != bodies[0]->kernel_offset || b->kernel_cutoff != bodies[0]->kernel_cutoff) {
            // each body assembles with its own scalars; nothing to check
        }
        int st = precision == FP_F64 ? fp_launch_assemble<double>(ctx, b, tv, (double *)A, ld, col, A32)
                                     : fp_launch_assemble<float>(ctx, b, tv, (float *)A, ld, col);
        if (st) return st;
        col += b->N;
    }
    return FP_OK;
}

// The block-Jacobi inverses of GMRES-IR can start before the AIC is whole when
// their blocks are index ranges of one vortex-ring body: its diagonal blocks are
// assembled on their own first and inverted beside the assembly.  Off unless
// FLOWPANEL_BJ_EARLY=1: on the 30,000-panel loft the Gauss-Jordan products slow
// the assembly by as much as they save (39.1 ms per solve either way).
static bool early_block_jacobi(const fp_solve_opts &o, fp_body *const *bodies, int nb)
{
    const char *e = getenv("FLOWPANEL_BJ_EARLY");
    if (!(e && e[0] == '1')) return false;
    return o.solver == FP_SOLVER_GMRES_IR && o.precond_order == FP_ORDER_INDEX && nb == 1 &&
           bodies[0]->element == FP_ELEM_VORTEXRING && o.precond_block <= 1024;
}

// Assembly of the bodies' AIC and solution for k right-hand sides (device,
// column-major n x k).  The matrix storage follows the solver: FP64 for LU and
// GMRES, FP64 + its FP32 narrowing for the refinement paths with a stored
// residual, FP32 alone with a matrix-free FP64 residual.
static int solve_system(fp_ctx *ctx, fp_body *const *bodies, int nb, const fp_solve_opts &o, const double *rhs,
                        int64_t k, double *x, fp_solve_stats &s)
{
    int st;
    Targets t;
    if ((st = gather_targets(ctx, bodies, nb, t))) return st;
    const int64_t n = t.M, ld = round_ld(n);
    std::vector<fp_body *> bl(bodies, bodies + nb);
    MatrixFreeOp mf(ctx, bl, t.tv);
    if ((st = mf.init())) return st;
    SolveArgs a;
    a.solver = o.solver;
    a.n = n;
    a.ld = ld;
    a.geo = &t.tv;
    a.resid = &mf;
    Timer tm(ctx->stream);
    const bool f64 = o.solver == FP_SOLVER_LU || o.solver == FP_SOLVER_GMRES;
    // refinement with a stored FP64 residual: assemble in FP64 once, keep it,
    // and store its FP32 rounding in the same pass for the low-precision solver
    const bool dual = !f64 && o.residual_mode == FP_RESID_STORED;
    DBuf A, A32buf;
    std::unique_ptr<DenseOp<double>> op64;
    std::unique_ptr<BJPrecond> pre;
    if (early_block_jacobi(o, bodies, nb)) {
        // the block inverses run on their own stream beside the assembly
        pre.reset(new BJPrecond(ctx, n, (int)std::min<int64_t>(o.precond_block > 0 ? o.precond_block : 512, n)));
        if ((st = pre->assemble_early(bodies[0], t.tv))) return st;
    }
    if ((st = dalloc(ctx, A, (size_t)n * ld * ((f64 || dual) ? sizeof(double) : sizeof(float))))) return st;
    if (dual && (st = dalloc(ctx, A32buf, (size_t)n * ld * sizeof(float)))) return st;
    if ((st = assemble_into(ctx, bodies, nb, t.tv, (f64 || dual) ? FP_F64 : FP_F32, A.p, ld,
                            dual ? A32buf.as<float>() : nullptr)))
        return st;
    if (pre && (st = pre->invert_early())) return st;
    if (dual) {
        a.A32 = A32buf.as<float>();
        op64.reset(new DenseOp<double>(ctx, A.as<double>(), n, ld));
        a.resid = op64.get();
        a.R64 = A.as<double>();
    } else if (f64) {
        a.A64 = A.as<double>();
    } else {
        a.A32 = A.as<float>();
    }
    tm.mark();   // read after the solve: the set-up is queued while the assembly runs
    if (o.solver == FP_SOLVER_GMRES) {
        op64.reset(new DenseOp<double>(ctx, a.A64, n, ld));
        a.resid = op64.get();
    }
    a.nrhs = k;
    Solver solver(ctx, a, o);
    if (pre) solver.adopt(std::move(pre));
    if ((st = solver.setup(s))) return st;
    st = solver.solve_multi(rhs, x, k, s);
    s.ms_assemble = tm.elapsed();
    return st;
}

static int check_solve_inputs(fp_ctx *ctx, fp_body *const *bodies, int nb, const fp_solve_opts *opts,
                              fp_solve_opts &o)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    for (int k = 0; k < nb; ++k)
        if (!bodies[k]->has_uinf)
            return fp_fail(ctx, FP_ERR_SHAPE, "Invalid Uinfs; call fp_body_set_freestream before solve");
    if (opts) o = *opts;
    else fp_solve_opts_default(&o, FP_SOLVER_LU);
    FP_HIP(ctx, hipSetDevice(ctx->device));
    return FP_OK;
}

int fp_solve_bodies(fp_ctx *ctx, fp_body *const *bodies, int nb, const fp_solve_opts *opts, fp_solve_stats *stats)
{
    fp_solve_opts o;
    int st = check_solve_inputs(ctx, bodies, nb, opts, o);
    if (st) return st;
    fp_solve_stats s{};
    int64_t n = 0;
    for (int k = 0; k < nb; ++k) n += bodies[k]->N;
    DBuf rhs, x;
    if ((st = dalloc(ctx, rhs, n * sizeof(double)))) return st;
    if ((st = dalloc(ctx, x, n * sizeof(double)))) return st;
    int64_t col = 0;
    for (int k = 0; k < nb; ++k) {
        if ((st = fp_rhs_kernel(ctx, bodies[k], rhs.as<double>() + col))) return st;
        col += bodies[k]->N;
    }
    st = solve_system(ctx, bodies, nb, o, rhs.as<double>(), 1, x.as<double>(), s);
    if (stats) *stats = s;
    if (st) return st;
    col = 0;
    for (int k = 0; k < nb; ++k) {
        FP_HIP(ctx, hipMemcpyAsync(bodies[k]->d_strength, x.as<double>() + col, bodies[k]->N * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
        bodies[k]->solved = true;
        col += bodies[k]->N;
    }
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

// solve(RigidWakeBody{VortexRing, 2}, Uinfs, Das, Dbs; elprescribe): the
// watertight least-squares formulation (src/FLOWPanel_liftingbody.jl:236-286,
// 357-447).  The prescribed strengths move to the right-hand side, and the
// normal equations G_red^T G_red x = G_red^T b are formed on MFMA as the full
// G^T G (one FP64 GEMM of the transposed AIC with the AIC) with the prescribed
// rows and columns replaced by the identity, which gives the same x and the
// prescribed values in place (set_solution, :290-316).
// The same least-squares solution without the normal equations.  With k
// prescribed strengths the reduced matrix G_red (n x (n-k), full column rank)
// has a range of codimension k, spanned in its complement by Z = null(G_red^T).
// B = G with the prescribed columns replaced by G[:, p] + w_p (w_p: the panel
// areas of p's body, scaled to |G_pp|, for the first prescribed strength of each
// body -- a vector outside the range of G also for a closed body, whose uniform
// strengths G annihilates) is square and nonsingular, and
//   B^T Z = E_P      (E_P the unit vectors of the prescribed strengths)
// gives Z exactly: its rows j not in P say G[:, j]^T Z = 0.  The least-squares
// residual is the projection r* = Z (Z^T Z)^-1 Z^T b_red of b_red onto that
// complement, and  B y = b_red - r*  is consistent with y_P = 0 and y = the
// least-squares solution elsewhere.  Two square solves with the chosen solver
// (one transposed, k right-hand sides): condition kappa(G), not kappa(G)^2, and
// (4/3) n^3 flop for LU against 2 n^3 + (2/3) n^3 through G^T G.
static int lsq_bordered(fp_ctx *ctx, fp_body *const *bodies, int nb, const std::vector<int64_t> &pidx,
                        const double *pres_val, const fp_solve_opts &o, fp_solve_stats &s, double *x_out)
{
    const int np = (int)pidx.size();
    int st;
    Timer tm(ctx->stream);
    Targets t;
    if ((st = gather_targets(ctx, bodies, nb, t))) return st;
    const int64_t n = t.M, ld = round_ld(n);
    DBuf G, B, Bt, rhs, dp, dv, w, Z, E, part, dots, x;
    if ((st = dalloc(ctx, G, (size_t)n * ld * sizeof(double))) || (st = dalloc(ctx, B, (size_t)n * ld * sizeof(double))) ||
        (st = dalloc(ctx, Bt, (size_t)n * ld * sizeof(double))) || (st = dalloc(ctx, rhs, n * sizeof(double))) ||
        (st = dalloc(ctx, dp, std::max(1, np) * sizeof(int64_t))) || (st = dalloc(ctx, dv, std::max(1, np) * sizeof(double))) ||
        (st = dalloc(ctx, w, (size_t)std::max(1, np) * n * sizeof(double))) ||
        (st = dalloc(ctx, Z, (size_t)std::max(1, np) * n * sizeof(double))) ||
        (st = dalloc(ctx, E, (size_t)std::max(1, np) * n * sizeof(double))) ||
        (st = dalloc(ctx, part, (size_t)512 * (np + 2) * sizeof(double))) ||
        (st = dalloc(ctx, dots, (size_t)(np + 1) * (np + 1) * sizeof(double))) || (st = dalloc(ctx, x, n * sizeof(double))))
        return st;
    if (np) {
        FP_HIP(ctx, hipMemcpyAsync(dp.p, pidx.data(), np * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(dv.p, pres_val, np * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    int64_t col = 0;
    for (int b = 0; b < nb; ++b) {
        if ((st = fp_rhs_kernel(ctx, bodies[b], rhs.as<double>() + col))) return st;
        col += bodies[b]->N;
    }
    if ((st = assemble_into(ctx, bodies, nb, t.tv, FP_F64, G.p, ld))) return st;
    if ((st = fp_lsq_rhs(ctx, G.as<double>(), n, ld, dp.as<int64_t>(), dv.as<double>(), np, rhs.as<double>())))
        return st;
    // border vectors: the areas of the body holding p, scaled to |G_pp|, for the first p of every body
    std::vector<double> areas(n), gpp(np);
    col = 0;
    for (int b = 0; b < nb; ++b) {
        FP_HIP(ctx, hipMemcpyAsync(areas.data() + col, bodies[b]->d_geo + 6 * bodies[b]->N, bodies[b]->N * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
        col += bodies[b]->N;
    }
    for (int q = 0; q < np; ++q)
        FP_HIP(ctx, hipMemcpyAsync(&gpp[q], G.as<double>() + pidx[q] * ld + pidx[q], sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> wh((size_t)std::max(1, np) * n, 0.0), eh((size_t)std::max(1, np) * n, 0.0);
    std::vector<char> seen(nb, 0);
    for (int q = 0; q < np; ++q) {
        int64_t lo = 0;
        int b = 0;
        while (b < nb && pidx[q] >= lo + bodies[b]->N) { lo += bodies[b]->N; ++b; }
        eh[(size_t)q * n + pidx[q]] = 1.0;
        if (seen[b]) continue;
        seen[b] = 1;
        double amean = 0;
        for (int64_t i = lo; i < lo + bodies[b]->N; ++i) amean += areas[i];
        amean /= (double)bodies[b]->N;
        const double sc = (std::fabs(gpp[q]) > 0 ? std::fabs(gpp[q]) : 1.0) / (amean > 0 ? amean : 1.0);
        for (int64_t i = lo; i < lo + bodies[b]->N; ++i) wh[(size_t)q * n + i] = sc * areas[i];
    }
    if (np) {
        FP_HIP(ctx, hipMemcpyAsync(w.p, wh.data(), (size_t)np * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(E.p, eh.data(), (size_t)np * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    FP_HIP(ctx, hipMemcpyAsync(B.p, G.p, (size_t)n * ld * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if ((st = fp_lsq_border(ctx, B.as<double>(), n, ld, dp.as<int64_t>(), np, w.as<double>()))) return st;
    if ((st = fp_mat_transpose(ctx, B.as<double>(), n, ld, Bt.p, FP_F64))) return st;
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));   // the host vectors go out of scope
    s.ms_assemble = tm.stop();
    // solve B^T Z = E_P, then B y = b_red - Z (Z^T Z)^-1 Z^T b_red, with the chosen solver
    auto square_solve = [&](DBuf &M, const double *rhs_d, int64_t k, double *out) -> int {
        SolveArgs a;
        a.solver = o.solver;
        a.n = n;
        a.ld = ld;
        a.geo = &t.tv;
        DBuf W, W32;
        std::unique_ptr<DenseOp<double>> op(new DenseOp<double>(ctx, M.as<double>(), n, ld));
        a.resid = op.get();
        a.R64 = M.as<double>();
        int s2;
        if (o.solver == FP_SOLVER_LU) {
            if ((s2 = dalloc(ctx, W, (size_t)n * ld * sizeof(double)))) return s2;
            FP_HIP(ctx, hipMemcpyAsync(W.p, M.p, (size_t)n * ld * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            a.A64 = W.as<double>();
        } else if (o.solver == FP_SOLVER_GMRES) {
            a.A64 = M.as<double>();
        } else {
            if ((s2 = dalloc(ctx, W32, (size_t)n * ld * sizeof(float)))) return s2;
            if ((s2 = fp_mat_narrow(ctx, M.as<double>(), W32.as<float>(), n, ld))) return s2;
            a.A32 = W32.as<float>();
        }
        a.nrhs = k;
        Solver solver(ctx, a, o);
        if ((s2 = solver.setup(s))) return s2;
        return solver.solve_multi(rhs_d, out, k, s);
    };
    if (np) {
        if ((st = square_solve(Bt, E.as<double>(), np, Z.as<double>()))) return st;
        // c = (Z^T Z)^-1 Z^T b_red on the host (np x np), then rhs -= Z c
        std::vector<double> ztz((size_t)np * np), ztb(np);
        double *dd = dots.as<double>();
        for (int q = 0; q < np; ++q)
            if ((st = fp_mdot(ctx, Z.as<double>(), n, np, Z.as<double>() + (size_t)q * n, n, part.as<double>(),
                              dd + (size_t)q * np)))
                return st;
        if ((st = fp_mdot(ctx, Z.as<double>(), n, np, rhs.as<double>(), n, part.as<double>(), dd + (size_t)np * np)))
            return st;
        FP_HIP(ctx, hipMemcpyAsync(ztz.data(), dd, (size_t)np * np * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(ztb.data(), dd + (size_t)np * np, np * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        // Gaussian elimination with partial pivoting (np is small)
        std::vector<double> c = ztb;
        for (int k = 0; k < np; ++k) {
            int pr = k;
            for (int r = k + 1; r < np; ++r)
                if (std::fabs(ztz[(size_t)r * np + k]) > std::fabs(ztz[(size_t)pr * np + k])) pr = r;
            if (!(std::fabs(ztz[(size_t)pr * np + k]) > 0))
                return fp_fail(ctx, FP_ERR_SINGULAR, "least squares: the complement of the reduced range is degenerate");
            for (int cc = 0; cc < np; ++cc) std::swap(ztz[(size_t)k * np + cc], ztz[(size_t)pr * np + cc]);
            std::swap(c[k], c[pr]);
            for (int r = k + 1; r < np; ++r) {
                const double f = ztz[(size_t)r * np + k] / ztz[(size_t)k * np + k];
                for (int cc = k; cc < np; ++cc) ztz[(size_t)r * np + cc] -= f * ztz[(size_t)k * np + cc];
                c[r] -= f * c[k];
            }
        }
        for (int k = np - 1; k >= 0; --k) {
            for (int cc = k + 1; cc < np; ++cc) c[k] -= ztz[(size_t)k * np + cc] * c[cc];
            c[k] /= ztz[(size_t)k * np + k];
        }
        for (int q = 0; q < np; ++q)
            if ((st = fp_axpby(ctx, -c[q], Z.as<double>() + (size_t)q * n, 1.0, rhs.as<double>(), n))) return st;
    }
    Bt.reset();
    if ((st = square_solve(B, rhs.as<double>(), 1, x.as<double>()))) return st;
    // the prescribed values in place (y_P is the border's coefficient, zero up to rounding)
    std::vector<double> xp(n);
    FP_HIP(ctx, hipMemcpyAsync(xp.data(), x.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < np; ++q) xp[pidx[q]] = pres_val[q];
    std::memcpy(x_out, xp.data(), n * sizeof(double));
    return FP_OK;
}

int fp_solve_bodies_lsq(fp_ctx *ctx, fp_body *const *bodies, int nb, const int64_t *pres_idx, const double *pres_val,
                        int64_t npres, int index_base, const fp_solve_opts *opts, fp_solve_stats *stats)
{
    fp_solve_opts o;
    int st = check_solve_inputs(ctx, bodies, nb, opts, o);
    if (st) return st;
    int64_t n = 0;
    for (int b = 0; b < nb; ++b) {
        if (bodies[b]->element != FP_ELEM_VORTEXRING)
            return fp_fail(ctx, FP_ERR_UNSUPPORTED, "the least-squares formulation takes vortex-ring bodies");
        n += bodies[b]->N;
    }
    if (npres < 0 || (npres > 0 && (!pres_idx || !pres_val))) return fp_fail(ctx, FP_ERR_ARG, "null elprescribe");
    if (npres >= n) return fp_fail(ctx, FP_ERR_ARG, "every strength is prescribed");
    std::vector<int64_t> pidx(npres);
    for (int64_t q = 0; q < npres; ++q) {
        pidx[q] = pres_idx[q] - index_base;
        if (pidx[q] < 0 || pidx[q] >= n) return fp_fail(ctx, FP_ERR_ARG, "prescribed element index out of range");
        for (int64_t r = 0; r < q; ++r)
            if (pidx[r] == pidx[q]) return fp_fail(ctx, FP_ERR_ARG, "element prescribed twice");
    }
    fp_solve_stats s{};
    const char *ne = getenv("FLOWPANEL_LSQ_NORMAL");
    if (!(ne && ne[0] == '1')) {
        std::vector<double> xh(n);
        st = lsq_bordered(ctx, bodies, nb, pidx, pres_val, o, s, xh.data());
        if (stats) *stats = s;
        if (st) return st;
        int64_t off = 0;
        for (int b = 0; b < nb; ++b) {
            FP_HIP(ctx, hipMemcpyAsync(bodies[b]->d_strength, xh.data() + off, bodies[b]->N * sizeof(double),
                                       hipMemcpyHostToDevice, ctx->stream));
            bodies[b]->solved = true;
            off += bodies[b]->N;
        }
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return FP_OK;
    }
    Timer tm(ctx->stream);
    Targets t;
    if ((st = gather_targets(ctx, bodies, nb, t))) return st;
    const int64_t ld = round_ld(n);
    DBuf G, GT, C, rhs, bls, x, dp, dv;
    if ((st = dalloc(ctx, G, (size_t)n * ld * sizeof(double))) || (st = dalloc(ctx, GT, (size_t)n * ld * sizeof(double))) ||
        (st = dalloc(ctx, C, (size_t)n * ld * sizeof(double))) || (st = dalloc(ctx, rhs, n * sizeof(double))) ||
        (st = dalloc(ctx, bls, n * sizeof(double))) || (st = dalloc(ctx, x, n * sizeof(double))) ||
        (st = dalloc(ctx, dp, std::max<int64_t>(1, npres) * sizeof(int64_t))) ||
        (st = dalloc(ctx, dv, std::max<int64_t>(1, npres) * sizeof(double))))
        return st;
    if (npres) {
        FP_HIP(ctx, hipMemcpyAsync(dp.p, pidx.data(), npres * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(dv.p, pres_val, npres * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    int64_t col = 0;
    for (int b = 0; b < nb; ++b) {
        if ((st = fp_rhs_kernel(ctx, bodies[b], rhs.as<double>() + col))) return st;
        col += bodies[b]->N;
    }
    if ((st = assemble_into(ctx, bodies, nb, t.tv, FP_F64, G.p, ld))) return st;
    if ((st = fp_lsq_rhs(ctx, G.as<double>(), n, ld, dp.as<int64_t>(), dv.as<double>(), (int)npres, rhs.as<double>())))
        return st;
    if ((st = fp_mat_transpose(ctx, G.as<double>(), n, ld, GT.p, FP_F64))) return st;
    if ((st = fp_fill(ctx, C.as<double>(), n * ld, 0.0))) return st;
    if ((st = fp_gemm_sub<double>(ctx, n, n, n, GT.as<double>(), ld, G.as<double>(), ld, C.as<double>(), ld))) return st;
    if ((st = fp_gemv<double>(ctx, GT.as<double>(), n, ld, rhs.as<double>(), bls.as<double>(), 1.0, 0.0))) return st;
    if ((st = fp_lsq_fix(ctx, C.as<double>(), n, ld, dp.as<int64_t>(), (int)npres, bls.as<double>(), dv.as<double>())))
        return st;
    s.ms_assemble = tm.stop();
    G.reset();
    GT.reset();
    // normal equations with the chosen solver (FP64 copy kept for the residual)
    SolveArgs a;
    a.solver = o.solver;
    a.n = n;
    a.ld = ld;
    a.geo = &t.tv;
    DBuf W, W32;
    std::unique_ptr<DenseOp<double>> op;
    op.reset(new DenseOp<double>(ctx, C.as<double>(), n, ld));
    a.resid = op.get();
    a.R64 = C.as<double>();
    if (o.solver == FP_SOLVER_LU) {
        if ((st = dalloc(ctx, W, (size_t)n * ld * sizeof(double)))) return st;
        FP_HIP(ctx, hipMemcpyAsync(W.p, C.p, (size_t)n * ld * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        a.A64 = W.as<double>();
    } else if (o.solver == FP_SOLVER_GMRES) {
        a.A64 = C.as<double>();
    } else {
        if ((st = dalloc(ctx, W32, (size_t)n * ld * sizeof(float)))) return st;
        if ((st = fp_mat_narrow(ctx, C.as<double>(), W32.as<float>(), n, ld))) return st;
        a.A32 = W32.as<float>();
    }
    Solver solver(ctx, a, o);
    st = solver.setup(s);
    if (!st) st = solver.solve(bls.as<double>(), x.as<double>(), s);
    if (stats) *stats = s;
    if (st) return st;
    col = 0;
    for (int b = 0; b < nb; ++b) {
        FP_HIP(ctx, hipMemcpyAsync(bodies[b]->d_strength, x.as<double>() + col, bodies[b]->N * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
        bodies[b]->solved = true;
        col += bodies[b]->N;
    }
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

int fp_solve_bodies_multi(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *Uinfs, int64_t k,
                          const fp_solve_opts *opts, double *strengths, fp_solve_stats *stats)
{
    fp_solve_opts o;
    int st = check_solve_inputs(ctx, bodies, nb, opts, o);
    if (st) return st;
    if (k <= 0 || !Uinfs || !strengths) return fp_fail(ctx, FP_ERR_ARG, "need k >= 1 freestreams and outputs");
    fp_solve_stats s{};
    int64_t n = 0;
    for (int b = 0; b < nb; ++b) n += bodies[b]->N;
    for (int64_t q = 0; q < 3 * n * k; ++q)
        if (!std::isfinite(Uinfs[q])) return fp_fail(ctx, FP_ERR_ARG, "non-finite freestream");
    DBuf rhs, x, U;
    if ((st = dalloc(ctx, rhs, (size_t)n * k * sizeof(double))) || (st = dalloc(ctx, x, (size_t)n * k * sizeof(double))) ||
        (st = dalloc(ctx, U, (size_t)3 * n * k * sizeof(double))))
        return st;
    // freestream c of body b as SoA 3 x N_b at U + 3*(c*n + col_b)
    std::vector<double> soa((size_t)3 * n * k);
    for (int64_t c = 0; c < k; ++c) {
        int64_t col = 0;
        for (int b = 0; b < nb; ++b) {
            const int64_t N = bodies[b]->N;
            double *dst = &soa[(size_t)3 * (c * n + col)];
            for (int64_t i = 0; i < N; ++i)
                for (int d = 0; d < 3; ++d) dst[d * N + i] = Uinfs[3 * (c * n + col + i) + d];
            col += N;
        }
    }
    FP_HIP(ctx, hipMemcpyAsync(U.p, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    for (int64_t c = 0; c < k; ++c) {
        int64_t col = 0;
        for (int b = 0; b < nb; ++b) {
            if ((st = fp_rhs_from(ctx, bodies[b], U.as<double>() + 3 * (c * n + col), rhs.as<double>() + c * n + col)))
                return st;
            col += bodies[b]->N;
        }
    }
    st = solve_system(ctx, bodies, nb, o, rhs.as<double>(), k, x.as<double>(), s);
    if (stats) *stats = s;
    if (st) return st;
    FP_HIP(ctx, hipMemcpyAsync(strengths, x.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

int fp_assemble(fp_ctx *ctx, fp_body *const *bodies, int nb, int precision, fp_mat **out)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    if (!out) return fp_fail(ctx, FP_ERR_ARG, "null output");
    if (precision != FP_F64 && precision != FP_F32) return fp_fail(ctx, FP_ERR_ARG, "unknown precision");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    Targets t;
    if ((st = gather_targets(ctx, bodies, nb, t))) return st;
    std::unique_ptr<fp_mat> m(new fp_mat());
    m->ctx = ctx;
    m->n = t.M;
    m->ld = round_ld(t.M);
    m->precision = precision;
    const size_t es = precision == FP_F64 ? sizeof(double) : sizeof(float);
    FP_HIP(ctx, hipMalloc(&m->A, (size_t)m->n * m->ld * es));
    if ((st = assemble_into(ctx, bodies, nb, t.tv, precision, m->A, m->ld))) { (void)hipFree(m->A); return st; }
    m->bodies.assign(bodies, bodies + nb);
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = m.release();
    return FP_OK;
}

int fp_mat_upload(fp_ctx *ctx, int64_t n, const double *Ah, fp_mat **out)
{
    if (!ctx || !out || !Ah || n <= 0) return FP_ERR_ARG;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<fp_mat> m(new fp_mat());
    m->ctx = ctx;
    m->n = n;
    m->ld = round_ld(n);
    m->precision = FP_F64;
    DBuf tmp;
    int st = dalloc(ctx, tmp, (size_t)n * n * sizeof(double));
    if (st) return st;
    FP_HIP(ctx, hipMemcpy(tmp.p, Ah, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice));
    FP_HIP(ctx, hipMalloc(&m->A, (size_t)n * m->ld * sizeof(double)));
    if ((st = fp_mat_convert(ctx, tmp.as<double>(), n, n, m->A, m->ld, FP_F64, 1))) { (void)hipFree(m->A); return st; }
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = m.release();
    return FP_OK;
}

int64_t fp_mat_size(const fp_mat *m) { return m ? m->n : -1; }

int fp_mat_block_inverses(fp_ctx *ctx, const fp_mat *m, int block, float *out)
{
    if (!ctx || !m || !out || m->ctx != ctx || block <= 0) return FP_ERR_ARG;
    if (m->factored != -1) return fp_fail(ctx, FP_ERR_ARG, "block inverses need a matrix that is not factorised");
    if (block > 1024) return fp_fail(ctx, FP_ERR_ARG, "block-Jacobi block larger than 1024");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t n = m->n;
    const int blk = (int)std::min<int64_t>(block, n);
    const int64_t nb = (n + blk - 1) / blk;
    DBuf inv, scr, info;
    int st;
    if ((st = dalloc(ctx, inv, (size_t)nb * blk * blk * sizeof(float)))) return st;
    const size_t sf = fp_bj_gj_scratch_floats(n, blk);
    if (sf && (st = dalloc(ctx, scr, sf * sizeof(float)))) return st;
    if ((st = dalloc(ctx, info, sizeof(int)))) return st;
    st = m->precision == FP_F64
             ? fp_bj_gj_inverse<double>(ctx, (const double *)m->A, n, m->ld, blk, inv.as<float>(), info.as<int>(),
                                        sf ? scr.as<float>() : nullptr)
             : fp_bj_gj_inverse<float>(ctx, (const float *)m->A, n, m->ld, blk, inv.as<float>(), info.as<int>(),
                                       sf ? scr.as<float>() : nullptr);
    if (st) return st;
    if (ctx->side2) FP_HIP(ctx, hipStreamSynchronize(ctx->side2));
    int hinfo = 0;
    FP_HIP(ctx, hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(out, inv.p, (size_t)nb * blk * blk * sizeof(float), hipMemcpyDeviceToHost,
                               ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hinfo) {
        char buf[128];
        snprintf(buf, sizeof buf, "block inverses: zero pivot in diagonal block %lld (global row %lld)",
                 (long long)((hinfo - 1) / blk), (long long)(hinfo - 1));
        return fp_fail(ctx, FP_ERR_SINGULAR, buf);
    }
    return FP_OK;
}

int fp_mat_download(fp_ctx *ctx, const fp_mat *m, double *Ah)
{
    if (!ctx || !m || !Ah || m->ctx != ctx) return FP_ERR_ARG;
    DBuf tmp;
    int st = dalloc(ctx, tmp, (size_t)m->n * m->n * sizeof(double));
    if (st) return st;
    dim3 grid((unsigned)((m->n + 255) / 256), (unsigned)m->n);
    if (m->precision == FP_F64)
        hipLaunchKernelGGL(k_download_cm<double>, grid, dim3(256), 0, ctx->stream, (const double *)m->A, m->n, m->ld,
                           tmp.as<double>());
    else
        hipLaunchKernelGGL(k_download_cm<float>, grid, dim3(256), 0, ctx->stream, (const float *)m->A, m->n, m->ld,
                           tmp.as<double>());
    FP_HIP(ctx, hipGetLastError());
    FP_HIP(ctx, hipMemcpyAsync(Ah, tmp.p, (size_t)m->n * m->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

void fp_mat_destroy(fp_mat *m)
{
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    if (m->A) (void)hipFree(m->A);
    delete m;
}

int fp_mat_getrf(fp_ctx *ctx, fp_mat *m, int precision, double *LU, int64_t *ipiv)
{
    if (!ctx || !m || m->ctx != ctx || !LU || !ipiv) return FP_ERR_ARG;
    if (precision != FP_F64 && precision != FP_F32) return fp_fail(ctx, FP_ERR_ARG, "unknown precision");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t n = m->n, ld = m->ld;
    DBuf F, p, pm, src64, tmp;
    int st;
    const size_t es = precision == FP_F64 ? sizeof(double) : sizeof(float);
    if ((st = dalloc(ctx, F, (size_t)n * ld * es))) return st;
    if ((st = dalloc(ctx, p, n * sizeof(int)))) return st;
    if ((st = dalloc(ctx, pm, n * sizeof(int)))) return st;
    // copy/convert the matrix into F
    if (m->precision == FP_F64 && precision == FP_F64) {
        FP_HIP(ctx, hipMemcpyAsync(F.p, m->A, (size_t)n * ld * es, hipMemcpyDeviceToDevice, ctx->stream));
    } else if (m->precision == FP_F64) {
        if ((st = fp_mat_narrow(ctx, (const double *)m->A, F.as<float>(), n, ld))) return st;
    } else if (precision == FP_F32) {
        FP_HIP(ctx, hipMemcpyAsync(F.p, m->A, (size_t)n * ld * es, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        return fp_fail(ctx, FP_ERR_UNSUPPORTED, "FP64 factors of an FP32 matrix");
    }
    int info = 0;
    st = precision == FP_F64 ? factor<double>(ctx, F.as<double>(), n, ld, pm.as<int>(), &info)
                             : factor<float>(ctx, F.as<float>(), n, ld, pm.as<int>(), &info);
    if (st && st != FP_ERR_SINGULAR) return st;
    if ((st == FP_OK || st == FP_ERR_SINGULAR)) {
        int st2 = dalloc(ctx, tmp, (size_t)n * n * sizeof(double));
        if (st2) return st2;
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
        if (precision == FP_F64)
            hipLaunchKernelGGL(k_download_cm<double>, grid, dim3(256), 0, ctx->stream, F.as<double>(), n, ld,
                               tmp.as<double>());
        else
            hipLaunchKernelGGL(k_download_cm<float>, grid, dim3(256), 0, ctx->stream, F.as<float>(), n, ld,
                               tmp.as<double>());
        FP_HIP(ctx, hipMemcpyAsync(LU, tmp.p, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        std::vector<int> pr(n);
        FP_HIP(ctx, hipMemcpyAsync(pr.data(), pm.p, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const std::vector<int> ip = perm_to_ipiv(pr);
        for (int64_t i = 0; i < n; ++i) ipiv[i] = ip[i] + 1;
    }
    return st;
}

int fp_solve_mat(fp_ctx *ctx, fp_mat *m, const double *bh, int64_t k, double *xh, const fp_solve_opts *opts,
                 fp_solve_stats *stats)
{
    if (!ctx || !m || m->ctx != ctx || !bh || !xh || k <= 0) return FP_ERR_ARG;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    fp_solve_opts o;
    if (opts) o = *opts;
    else fp_solve_opts_default(&o, FP_SOLVER_LU);
    const int64_t n = m->n, ld = m->ld;
    int st;
    DBuf b, x, W;
    if ((st = dalloc(ctx, b, (size_t)n * k * sizeof(double)))) return st;
    if ((st = dalloc(ctx, x, (size_t)n * k * sizeof(double)))) return st;
    FP_HIP(ctx, hipMemcpy(b.p, bh, (size_t)n * k * sizeof(double), hipMemcpyHostToDevice));
    // FP64 residual operator: stored FP64 matrix, else matrix-free from the bodies
    std::unique_ptr<Operator> resid;
    std::unique_ptr<MatrixFreeOp> mf;
    Targets t;
    if (m->precision == FP_F64) {
        resid.reset(new DenseOp<double>(ctx, (const double *)m->A, n, ld));
    } else {
        if (m->bodies.empty()) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "FP32 matrix without bodies");
        if ((st = gather_targets(ctx, m->bodies.data(), (int)m->bodies.size(), t))) return st;
        mf.reset(new MatrixFreeOp(ctx, m->bodies, t.tv));
        if ((st = mf->init())) return st;
    }
    Operator *R = resid ? resid.get() : (Operator *)mf.get();
    if (o.precond_order == FP_ORDER_SPATIAL && !m->bodies.empty() && t.M == 0 &&
        (st = gather_targets(ctx, m->bodies.data(), (int)m->bodies.size(), t)))
        return st;
    // one factorisation / preconditioner for all k right-hand sides
    fp_solve_stats s{};
    SolveArgs a;
    a.solver = o.solver;
    a.n = n;
    a.ld = ld;
    a.geo = m->bodies.empty() ? nullptr : &t.tv;
    a.resid = R;
    if (o.solver == FP_SOLVER_LU) {
        if (m->precision != FP_F64) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "FP64 LU of an FP32 matrix");
        if ((st = dalloc(ctx, W, (size_t)n * ld * sizeof(double)))) return st;
        FP_HIP(ctx, hipMemcpyAsync(W.p, m->A, (size_t)n * ld * sizeof(double), hipMemcpyDeviceToDevice,
                                   ctx->stream));
        a.A64 = W.as<double>();
    } else if (o.solver == FP_SOLVER_GMRES) {
        if (m->precision == FP_F64) a.A64 = (double *)m->A;
        else a.A32 = (float *)m->A;
    } else {
        if ((st = dalloc(ctx, W, (size_t)n * ld * sizeof(float)))) return st;
        if (m->precision == FP_F64) {
            if ((st = fp_mat_narrow(ctx, (const double *)m->A, W.as<float>(), n, ld))) return st;
        } else {
            FP_HIP(ctx, hipMemcpyAsync(W.p, m->A, (size_t)n * ld * sizeof(float), hipMemcpyDeviceToDevice,
                                       ctx->stream));
        }
        a.A32 = W.as<float>();
    }
    if (m->precision == FP_F64) a.R64 = (const double *)m->A;
    a.nrhs = k;
    Solver solver(ctx, a, o);
    st = solver.setup(s);
    if (!st) st = solver.solve_multi(b.as<double>(), x.as<double>(), k, s);
    if (st) { if (stats) *stats = s; return st; }
    FP_HIP(ctx, hipMemcpy(xh, x.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToHost));
    if (stats) *stats = s;
    return FP_OK;
}

static int field_eval(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *targets, int64_t nt, double *out,
                      int mode)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    if (nt < 0 || (nt > 0 && (!targets || !out))) return fp_fail(ctx, FP_ERR_ARG, "null targets/out");
    for (int k = 0; k < nb; ++k)
        if (!bodies[k]->solved)
            return fp_fail(ctx, FP_ERR_UNSOLVED, "Source body hasn't been solved yet. Please call `solve(...)` function first.");
    if (nt == 0) return FP_OK;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    const int comps = mode == 0 ? 3 : 1;
    DBuf tg, o, part;
    if ((st = dalloc(ctx, tg, 3 * nt * sizeof(double)))) return st;
    if ((st = dalloc(ctx, o, comps * nt * sizeof(double)))) return st;
    if ((st = upload_soa3(ctx, targets, nt, tg.as<double>()))) return st;
    FP_HIP(ctx, hipMemsetAsync(o.p, 0, comps * nt * sizeof(double), ctx->stream));
    TargetView tv{};
    tv.x = tg.as<double>(); tv.y = tv.x + nt; tv.z = tv.x + 2 * nt; tv.M = nt;
    size_t pb = 0;
    for (int k = 0; k < nb; ++k) pb = std::max(pb, fp_field_partial_bytes(nt, bodies[k]->N));
    if (pb && (st = dalloc(ctx, part, pb))) return st;
    for (int k = 0; k < nb; ++k)
        if ((st = fp_launch_field(ctx, bodies[k], bodies[k]->d_strength, tv, mode, o.as<double>(), part.as<double>(), pb)))
            return st;
    std::vector<double> h(comps * nt);
    FP_HIP(ctx, hipMemcpyAsync(h.data(), o.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < h.size(); ++i) out[i] += h[i];
    return FP_OK;
}

int fp_uind(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *targets, int64_t nt, double *out)
{
    return field_eval(ctx, bodies, nb, targets, nt, out, 0);
}

int fp_phi(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *targets, int64_t nt, double *out)
{
    return field_eval(ctx, bodies, nb, targets, nt, out, 2);
}

int fp_calcfield_U(fp_ctx *ctx, fp_body *const *bodies, int nb, double *out)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    if (!out) return fp_fail(ctx, FP_ERR_ARG, "null out");
    for (int k = 0; k < nb; ++k) {
        if (!bodies[k]->solved)
            return fp_fail(ctx, FP_ERR_UNSOLVED, "Source body hasn't been solved yet. Please call `solve(...)` function first.");
        if (!bodies[k]->has_uinf) return fp_fail(ctx, FP_ERR_ARG, "Target body doesn't have freestream field `Uinf`");
    }
    FP_HIP(ctx, hipSetDevice(ctx->device));
    Targets t;
    if ((st = gather_targets(ctx, bodies, nb, t))) return st;
    const int64_t M = t.M;
    DBuf o, part;
    if ((st = dalloc(ctx, o, 3 * M * sizeof(double)))) return st;
    FP_HIP(ctx, hipMemsetAsync(o.p, 0, 3 * M * sizeof(double), ctx->stream));
    int64_t off = 0;
    for (int k = 0; k < nb; ++k) {
        hipLaunchKernelGGL(k_aos_add_soa, dim3((unsigned)((bodies[k]->N + 255) / 256)), dim3(256), 0, ctx->stream,
                           bodies[k]->d_uinf, bodies[k]->N, o.as<double>() + 3 * off);
        off += bodies[k]->N;
    }
    size_t pb = 0;
    for (int k = 0; k < nb; ++k) pb = std::max(pb, fp_field_partial_bytes(M, bodies[k]->N));
    if (pb && (st = dalloc(ctx, part, pb))) return st;
    for (int k = 0; k < nb; ++k)
        if ((st = fp_launch_field(ctx, bodies[k], bodies[k]->d_strength, t.tv, 0, o.as<double>(), part.as<double>(), pb)))
            return st;
    std::vector<double> h(3 * M);
    FP_HIP(ctx, hipMemcpyAsync(h.data(), o.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < h.size(); ++i) out[i] += h[i];
    return FP_OK;
}

int fp_calcfield_cp(fp_ctx *ctx, const double *Us, int64_t n, double uref, const int64_t *shedding, int64_t ns,
                    int index_base, int kutta_mode, double *out)
{
    if (!ctx || !Us || !out || n < 0 || ns < 0) return FP_ERR_ARG;
    if (kutta_mode < FP_KUTTA_NONE || kutta_mode > FP_KUTTA_PAIR) return fp_fail(ctx, FP_ERR_ARG, "bad kutta mode");
    if (ns > 0 && !shedding && kutta_mode != FP_KUTTA_NONE) return fp_fail(ctx, FP_ERR_ARG, "null shedding");
    if (n == 0) return FP_OK;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<int64_t> sh;
    if (kutta_mode != FP_KUTTA_NONE) {
        sh.resize(6 * ns);
        for (int64_t e = 0; e < ns; ++e)
            for (int k = 0; k < 6; ++k) {
                const int64_t v = shedding[6 * e + k];
                sh[6 * e + k] = (k == 3 && v == -1) ? -1 : v - index_base;
            }
        for (int64_t e = 0; e < ns; ++e) {
            const int64_t pi = sh[6 * e], pj = sh[6 * e + 3];
            const bool ref = kutta_mode == FP_KUTTA_REFERENCE;
            if (pi < 0 || pi + (ref ? 1 : 0) >= n || (pj != -1 && (pj - (ref ? 1 : 0) < 0 || pj >= n)))
                return fp_fail(ctx, FP_ERR_ARG, "Kutta panel index out of range");
        }
    }
    DBuf U, o, S;
    int st;
    if ((st = dalloc(ctx, U, 3 * n * sizeof(double)))) return st;
    if ((st = dalloc(ctx, o, n * sizeof(double)))) return st;
    FP_HIP(ctx, hipMemcpyAsync(U.p, Us, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(o.p, out, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_cp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, U.as<double>(), n, uref,
                       o.as<double>());
    if (kutta_mode != FP_KUTTA_NONE && ns > 0) {
        if ((st = dalloc(ctx, S, sh.size() * sizeof(int64_t)))) return st;
        FP_HIP(ctx, hipMemcpyAsync(S.p, sh.data(), sh.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_kutta, dim3(1), dim3(64), 0, ctx->stream, S.as<int64_t>(), ns, kutta_mode, o.as<double>());
    }
    FP_HIP(ctx, hipGetLastError());
    FP_HIP(ctx, hipMemcpyAsync(out, o.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

static int ensure_shedding_on_device(fp_ctx *ctx, fp_body *b)
{
    if (b->ns > 0 && !b->d_sh) {
        FP_HIP(ctx, hipMalloc((void **)&b->d_sh, 6 * b->ns * sizeof(int64_t)));
        FP_HIP(ctx, hipMemcpy(b->d_sh, b->shedding.data(), 6 * b->ns * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    return FP_OK;
}

// calcfield_F! with calcfield_Ftot! and calcfield_Mtot! (src/FLOWPanel_postprocess.jl:1199-1307, 1543-1784)
int fp_calcfield_F(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *Cps, double Uinf, double rho,
                   int kutta_mode, const double *Xac, const double *X, double *F, double *Ftot, double *Mtot)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    if (!Cps) return fp_fail(ctx, FP_ERR_ARG, "null Cps");
    if (kutta_mode < FP_KUTTA_NONE || kutta_mode > FP_KUTTA_PAIR) return fp_fail(ctx, FP_ERR_ARG, "unknown Kutta mode");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    int64_t M = 0;
    size_t pb = 0;
    for (int k = 0; k < nb; ++k) {
        M += bodies[k]->N;
        pb = std::max(pb, fp_loads_partial_bytes(bodies[k]->N));
    }
    DBuf cp, Fd, part, o6, Xd;
    if ((st = dalloc(ctx, cp, M * sizeof(double))) || (st = dalloc(ctx, Fd, 3 * M * sizeof(double))) ||
        (st = dalloc(ctx, part, pb)) || (st = dalloc(ctx, o6, 6 * nb * sizeof(double))))
        return st;
    FP_HIP(ctx, hipMemcpyAsync(cp.p, Cps, M * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (X) {
        if ((st = dalloc(ctx, Xd, 3 * M * sizeof(double)))) return st;
        int64_t off = 0;
        for (int k = 0; k < nb; ++k) {
            if ((st = upload_soa3(ctx, X + 3 * off, bodies[k]->N, Xd.as<double>() + 3 * off))) return st;
            off += bodies[k]->N;
        }
    }
    const double zero[3] = {0, 0, 0};
    const double q = 0.5 * rho * Uinf * Uinf;
    std::vector<char> te(M, 0);
    int64_t off = 0;
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        const bool kutta = kutta_mode != FP_KUTTA_NONE && b->element == FP_ELEM_VORTEXRING && b->ns > 0;
        if (kutta) {
            // the Kutta override of the force: the averaged Cp of the trailing-edge cells replaces theirs
            if ((st = ensure_shedding_on_device(ctx, b))) return st;
            hipLaunchKernelGGL(k_kutta, dim3(1), dim3(64), 0, ctx->stream, (const int64_t *)b->d_sh, b->ns, kutta_mode,
                               cp.as<double>() + off);
            for (int64_t e = 0; e < b->ns; ++e) {
                const int64_t pi = b->shedding[6 * e], pj = b->shedding[6 * e + 3];
                if (kutta_mode == FP_KUTTA_REFERENCE) {
                    te[off + pi] = te[off + pi + 1] = 1;
                    if (pj != -1) te[off + pj] = te[off + pj - 1] = 1;
                } else if (pj != -1) {
                    te[off + pi] = te[off + pj] = 1;
                }
            }
        }
        if ((st = fp_launch_loads(ctx, b, cp.as<double>() + off, q, X ? Xd.as<double>() + 3 * off : nullptr,
                                  Xac ? Xac : zero, Fd.as<double>() + 3 * off, part.as<double>(),
                                  o6.as<double>() + 6 * k)))
            return st;
        off += b->N;
    }
    std::vector<double> h6(6 * nb), hF(F ? 3 * M : 0);
    FP_HIP(ctx, hipMemcpyAsync(h6.data(), o6.p, h6.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (F) FP_HIP(ctx, hipMemcpyAsync(hF.data(), Fd.p, hF.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (F)
        for (int64_t i = 0; i < M; ++i)
            for (int d = 0; d < 3; ++d) F[3 * i + d] = te[i] ? hF[3 * i + d] : F[3 * i + d] + hF[3 * i + d];
    for (int k = 0; k < nb; ++k)
        for (int d = 0; d < 3; ++d) {
            if (Ftot) Ftot[d] += h6[6 * k + d];
            if (Mtot) Mtot[d] += h6[6 * k + 3 + d];
        }
    return FP_OK;
}

// calcfield_sectionalforce! (src/FLOWPanel_postprocess.jl:1411-1542)
int fp_calcfield_sectionalforce(fp_ctx *ctx, const fp_body *b, const double *F, const double *X, int64_t nchord,
                                int64_t nspan, const double *spandir, double *outF, double *outpos)
{
    if (!ctx) return FP_ERR_ARG;
    if (!b || b->ctx != ctx) return fp_fail(ctx, FP_ERR_ARG, "invalid body handle");
    if (!F || !spandir || !outF || !outpos) return fp_fail(ctx, FP_ERR_ARG, "null argument");
    if (nchord <= 0 || nspan < 2 || nchord * nspan != b->N)
        return fp_fail(ctx, FP_ERR_SHAPE, "sectional loads need a structured grid of nchord x nspan >= 2 cells");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = b->N;
    DBuf Fd, Xd, of, op;
    int st;
    if ((st = dalloc(ctx, Fd, 3 * N * sizeof(double))) || (st = dalloc(ctx, of, 3 * nspan * sizeof(double))) ||
        (st = dalloc(ctx, op, nspan * sizeof(double))))
        return st;
    FP_HIP(ctx, hipMemcpyAsync(Fd.p, F, 3 * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const double *Xs = b->d_geo + 3 * N;   // the control points (SoA)
    if (X) {
        if ((st = dalloc(ctx, Xd, 3 * N * sizeof(double))) || (st = upload_soa3(ctx, X, N, Xd.as<double>()))) return st;
        Xs = Xd.as<double>();
    }
    if ((st = fp_launch_sectional(ctx, Fd.as<double>(), Xs, N, nchord, nspan, spandir, of.as<double>(),
                                  op.as<double>())))
        return st;
    std::vector<double> hf(3 * nspan), hp(nspan);
    FP_HIP(ctx, hipMemcpyAsync(hf.data(), of.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(hp.data(), op.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // per unit span: central differences of the station positions, one-sided at the ends
    for (int64_t j = 0; j < nspan; ++j) {
        const double dp = j == 0 ? hp[1] - hp[0] : (j == nspan - 1 ? hp[j] - hp[j - 1] : (hp[j + 1] - hp[j - 1]) / 2);
        for (int d = 0; d < 3; ++d) outF[3 * j + d] = hf[3 * j + d] / std::fabs(dp);
        outpos[j] = hp[j];
    }
    return FP_OK;
}

int64_t fp_body_nnodes(const fp_body *b) { return b ? (int64_t)(b->nodes.size() / 3) : -1; }

int fp_body_set_nodes(fp_ctx *ctx, fp_body *b, const double *nodes)
{
    if (!ctx || !b || b->ctx != ctx || !nodes) return FP_ERR_ARG;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t nn = (int64_t)b->nodes.size() / 3;
    for (int64_t k = 0; k < 3 * nn; ++k)
        if (!std::isfinite(nodes[k])) return fp_fail(ctx, FP_ERR_ARG, "non-finite node coordinate");
    b->nodes.assign(nodes, nodes + 3 * nn);
    ++b->version;
    const int nvp = b->nvp;
    const int64_t N = b->N;
    std::vector<double> pv((size_t)3 * nvp * N);
    for (int64_t j = 0; j < N; ++j) {
        const int64_t *pc = &b->cell_idx[b->cell_ptr[j]];
        const int nv = (int)(b->cell_ptr[j + 1] - b->cell_ptr[j]);
        for (int k = 0; k < nvp; ++k) {
            const int64_t v = pc[k < nv ? k : 0];
            for (int d = 0; d < 3; ++d) pv[(size_t)(3 * k + d) * N + j] = b->nodes[3 * v + d];
        }
    }
    FP_HIP(ctx, hipMemcpyAsync(b->d_pv, pv.data(), pv.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int st = fp_launch_geometry(ctx, b);
    if (st) return st;
    if (b->has_wake && (st = build_horseshoes(ctx, b))) return st;
    b->solved = false;
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

void fp_liftdrag_opts_default(fp_liftdrag_opts *o, int solver)
{
    if (!o) return;
    o->probe_offset = 0.02;
    o->probe_length = 1.0;
    o->probe_charlen = FP_CHARLEN_UNITARY;
    o->kutta_mode = FP_KUTTA_PAIR;
    o->span_dir[0] = 0; o->span_dir[1] = 1; o->span_dir[2] = 0;
    o->ref_area = 1.0;
    fp_solve_opts_default(&o->solve, solver);
}

// Forward (and optionally reverse) pass of the lift/drag objective.
// The two adjoint systems A^T lambda_m = g_m (m = CL, CDi) with a Krylov solver:
// system 0 on the context stream, system 1 on a second stream (default
// priority) owned by a child context and driven from a host thread.  Each has
// its own solver workspace and preconditioner; the matrices are shared
// read-only.  The child stream first waits for everything already queued on
// the context stream.
static int transpose_solves_concurrent(fp_ctx *ctx, const SolveArgs &at, const fp_solve_opts &so,
                                       const double *Ares, const double *g, double *lam, int64_t M,
                                       fp_solve_stats &ts)
{
    fp_ctx child;
    child.device = ctx->device;
    child.num_cu = ctx->num_cu;
    FP_HIP(ctx, hipStreamCreateWithFlags(&child.stream, hipStreamNonBlocking));
    hipEvent_t ready = nullptr;
    hipError_t e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ready, ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(child.stream, ready, 0);
    if (e != hipSuccess) {
        if (ready) (void)hipEventDestroy(ready);
        (void)hipStreamDestroy(child.stream);
        FP_HIP(ctx, e);
    }
    int st1 = FP_OK;
    fp_solve_stats ts1{};
    std::thread worker([&]() {
        if (hipSetDevice(child.device) != hipSuccess) {
            st1 = fp_fail(&child, FP_ERR_DEVICE, "hipSetDevice failed in the adjoint worker thread");
            return;
        }
        SolveArgs a1 = at;
        DenseOp<double> op1(&child, Ares, at.n, at.ld);
        a1.resid = &op1;
        {
            Solver s1(&child, a1, so);
            st1 = s1.setup(ts1);
            if (!st1) st1 = s1.solve(g + M, lam + M, ts1);
        }
        if (hipStreamSynchronize(child.stream) != hipSuccess && !st1)
            st1 = fp_fail(&child, FP_ERR_DEVICE, "adjoint worker stream synchronisation failed");
    });
    int st0;
    {
        Solver s0(ctx, at, so);
        st0 = s0.setup(ts);
        if (!st0) st0 = s0.solve(g, lam, ts);
    }
    worker.join();
    (void)hipEventDestroy(ready);
    for (hipStream_t sd : {child.side, child.side2, child.side3})
        if (sd) (void)hipStreamDestroy(sd);
    for (hipEvent_t ev : {child.gj_fork, child.gj_join, child.cgj_fork, child.cgj_join})
        if (ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(child.stream);
    child.stream = nullptr;
    if (st0) return st0;
    if (st1) return fp_fail(ctx, st1, child.err.c_str());
    ts.iterations += ts1.iterations;
    ts.refinements += ts1.refinements;
    ts.residual = worst_residual(ts.residual, ts1.residual);
    return FP_OK;
}

static int liftdrag_run(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *U, const fp_liftdrag_opts *lo,
                        bool adjoint, double *CLo, double *CDo, double *dCL_dn, double *dCD_dn, double *dCL_dU,
                        double *dCD_dU, fp_liftdrag_stats *stats_out)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    if (!U) return fp_fail(ctx, FP_ERR_ARG, "null Uinf");
    fp_liftdrag_opts o;
    if (lo) o = *lo;
    else fp_liftdrag_opts_default(&o, FP_SOLVER_LU);
    const double un = std::sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
    if (!(un > 0) || !std::isfinite(un)) return fp_fail(ctx, FP_ERR_ARG, "Uinf must be finite and non-zero");
    if (!(o.ref_area > 0)) return fp_fail(ctx, FP_ERR_ARG, "ref_area must be positive");
    if (!std::isfinite(o.probe_offset) || !std::isfinite(o.probe_length))
        return fp_fail(ctx, FP_ERR_ARG, "probe offset/length must be finite");
    if (o.probe_charlen < 0 || o.probe_charlen > 3) return fp_fail(ctx, FP_ERR_ARG, "unknown probe_charlen");
    if (o.kutta_mode < FP_KUTTA_NONE || o.kutta_mode > FP_KUTTA_PAIR) return fp_fail(ctx, FP_ERR_ARG, "bad kutta mode");
    const double dh[3] = {U[0] / un, U[1] / un, U[2] / un};
    const double *sd = o.span_dir;
    const double wv[3] = {dh[1] * sd[2] - dh[2] * sd[1], dh[2] * sd[0] - dh[0] * sd[2], dh[0] * sd[1] - dh[1] * sd[0]};
    const double wn = std::sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
    if (!(wn > 1e-12)) return fp_fail(ctx, FP_ERR_ARG, "span_dir is parallel to Uinf");
    const double Lh[3] = {wv[0] / wn, wv[1] / wn, wv[2] / wn};
    int64_t M = 0;
    for (int k = 0; k < nb; ++k) {
        const fp_body *b = bodies[k];
        if (b->element != FP_ELEM_VORTEXRING)
            return fp_fail(ctx, FP_ERR_UNSUPPORTED, "lift/drag objective needs vortex-ring bodies");
        for (int64_t e = 0; e < b->ns && o.kutta_mode != FP_KUTTA_NONE; ++e) {
            const int64_t pi = b->shedding[6 * e], pj = b->shedding[6 * e + 3];
            const bool ref = o.kutta_mode == FP_KUTTA_REFERENCE;
            if (pi + (ref ? 1 : 0) >= b->N || (pj != -1 && pj - (ref ? 1 : 0) < 0))
                return fp_fail(ctx, FP_ERR_ARG, "Kutta panel index out of range");
        }
        M += b->N;
    }
    FP_HIP(ctx, hipSetDevice(ctx->device));
    fp_liftdrag_stats S{};
    Timer tm(ctx->stream);
    // 1) freestream at every control point, wake along Uinf
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        std::vector<double> ui(3 * b->N);
        for (int64_t i = 0; i < b->N; ++i)
            for (int c = 0; c < 3; ++c) ui[3 * i + c] = U[c];
        if ((st = upload_soa3(ctx, ui.data(), b->N, b->d_uinf))) return st;
        b->das.assign(3 * b->ns, 0.0);
        for (int64_t e = 0; e < b->ns; ++e)
            for (int c = 0; c < 3; ++c) b->das[3 * e + c] = dh[c];
        b->dbs = b->das;
        if ((st = build_horseshoes(ctx, b))) return st;
        b->has_uinf = b->has_wake = true;
    }
    // 2) assembly and solve (FP64 stored matrix; transposed copy for the adjoint)
    fp_solve_opts so = o.solve;
    const bool f64 = so.solver == FP_SOLVER_LU || so.solver == FP_SOLVER_GMRES;
    if (adjoint) so.residual_mode = FP_RESID_STORED;
    const bool stored = f64 || so.residual_mode == FP_RESID_STORED;
    Targets t;
    if ((st = gather_targets(ctx, bodies, nb, t))) return st;
    const int64_t n = t.M, ld = round_ld(n);
    DBuf A, A32, AT, AT32, rhs, x;
    if ((st = dalloc(ctx, rhs, n * sizeof(double))) || (st = dalloc(ctx, x, n * sizeof(double)))) return st;
    int64_t col = 0;
    for (int k = 0; k < nb; ++k) {
        if ((st = fp_rhs_kernel(ctx, bodies[k], rhs.as<double>() + col))) return st;
        col += bodies[k]->N;
    }
    std::vector<fp_body *> bl(bodies, bodies + nb);
    MatrixFreeOp mf(ctx, bl, t.tv);
    if ((st = mf.init())) return st;
    std::unique_ptr<DenseOp<double>> op64, opT;
    SolveArgs a;
    a.solver = so.solver;
    a.n = n;
    a.ld = ld;
    a.geo = &t.tv;
    a.resid = &mf;
    if ((st = dalloc(ctx, A, (size_t)n * ld * (stored ? sizeof(double) : sizeof(float))))) return st;
    if (!f64 && stored && (st = dalloc(ctx, A32, (size_t)n * ld * sizeof(float)))) return st;
    if ((st = assemble_into(ctx, bodies, nb, t.tv, stored ? FP_F64 : FP_F32, A.p, ld,
                            (!f64 && stored) ? A32.as<float>() : nullptr)))
        return st;
    if (adjoint) {  // before any in-place factorisation of A
        if ((st = dalloc(ctx, AT, (size_t)n * ld * sizeof(double)))) return st;
        if ((st = fp_mat_transpose(ctx, A.as<double>(), n, ld, AT.p, FP_F64))) return st;
    }
    if (f64) {
        a.A64 = A.as<double>();
        if (so.solver == FP_SOLVER_GMRES) {
            op64.reset(new DenseOp<double>(ctx, a.A64, n, ld));
            a.resid = op64.get();
        }
    } else if (stored) {
        a.A32 = A32.as<float>();
        op64.reset(new DenseOp<double>(ctx, A.as<double>(), n, ld));
        a.resid = op64.get();
    } else {
        a.A32 = A.as<float>();
    }
    S.forward.ms_assemble = tm.stop();
    {
        Solver fwd(ctx, a, so);
        st = fwd.setup(S.forward);
        if (!st) st = fwd.solve(rhs.as<double>(), x.as<double>(), S.forward);
        if (st) { if (stats_out) *stats_out = S; return st; }
    }
    col = 0;
    for (int k = 0; k < nb; ++k) {
        FP_HIP(ctx, hipMemcpyAsync(bodies[k]->d_strength, x.as<double>() + col, bodies[k]->N * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
        bodies[k]->solved = true;
        col += bodies[k]->N;
    }
    // 3) probes, velocities, Cp, forces
    DBuf geo, V, Cp, T, ones, part, fpart, out8, shs;
    if ((st = dalloc(ctx, geo, 10 * M * sizeof(double)))) return st;
    double *nrm = geo.as<double>(), *area = nrm + 3 * M, *cpt = area + M, *prb = cpt + 3 * M;
    if ((st = dalloc(ctx, V, 3 * M * sizeof(double))) || (st = dalloc(ctx, Cp, M * sizeof(double))) ||
        (st = dalloc(ctx, T, 8 * M * sizeof(double))) || (st = dalloc(ctx, ones, M * sizeof(double))) ||
        (st = dalloc(ctx, part, 256 * 8 * sizeof(double))) || (st = dalloc(ctx, out8, 8 * sizeof(double))))
        return st;
    {
        std::vector<double> one(M, 1.0);
        FP_HIP(ctx, hipMemcpyAsync(ones.p, one.data(), M * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    col = 0;
    for (int k = 0; k < nb; ++k) {
        if ((st = fp_adj_geometry(ctx, bodies[k], o.probe_offset, o.probe_charlen, o.probe_length, M, col, nrm, area,
                                  cpt, prb)))
            return st;
        col += bodies[k]->N;
    }
    TargetView tp;
    tp.x = prb; tp.y = prb + M; tp.z = prb + 2 * M;
    tp.nx = nrm; tp.ny = nrm + M; tp.nz = nrm + 2 * M;
    tp.M = M;
    size_t fpb = 0;
    for (int k = 0; k < nb; ++k) fpb = std::max(fpb, fp_field_partial_bytes(M, bodies[k]->N));
    if (fpb && (st = dalloc(ctx, fpart, fpb))) return st;
    FP_HIP(ctx, hipMemsetAsync(V.p, 0, 3 * M * sizeof(double), ctx->stream));
    for (int k = 0; k < nb; ++k)
        if ((st = fp_launch_field(ctx, bodies[k], bodies[k]->d_strength, tp, 0, V.as<double>(), fpart.as<double>(), fpb)))
            return st;
    if ((st = fp_adj_cp(ctx, V.as<double>(), M, U, Cp.as<double>()))) return st;
    // per-body shedding tables on the device (Kutta)
    std::vector<DBuf> dsh(nb);
    col = 0;
    for (int k = 0; k < nb; ++k) {
        const fp_body *b = bodies[k];
        if (b->ns > 0) {
            if ((st = dalloc(ctx, dsh[k], 6 * b->ns * sizeof(int64_t)))) return st;
            FP_HIP(ctx, hipMemcpyAsync(dsh[k].p, b->shedding.data(), 6 * b->ns * sizeof(int64_t),
                                       hipMemcpyHostToDevice, ctx->stream));
            if ((st = fp_adj_kutta(ctx, dsh[k].as<int64_t>(), b->ns, o.kutta_mode, Cp.as<double>() + col, nullptr, 0)))
                return st;
        }
        col += b->N;
    }
    if ((st = fp_adj_force_terms(ctx, Cp.as<double>(), area, nrm, M, T.as<double>()))) return st;
    if ((st = fp_mdot(ctx, T.as<double>(), M, 3, ones.as<double>(), M, part.as<double>(), out8.as<double>()))) return st;
    double fs[8];
    FP_HIP(ctx, hipMemcpyAsync(fs, out8.p, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double Fn[3] = {-fs[0] / o.ref_area, -fs[1] / o.ref_area, -fs[2] / o.ref_area};
    S.CL = Fn[0] * Lh[0] + Fn[1] * Lh[1] + Fn[2] * Lh[2];
    S.CDi = Fn[0] * dh[0] + Fn[1] * dh[1] + Fn[2] * dh[2];
    if (CLo) *CLo = S.CL;
    if (CDo) *CDo = S.CDi;
    S.ms_forward = tm.stop();
    if (!adjoint) {
        if (stats_out) *stats_out = S;
        return FP_OK;
    }
    // 4) reverse pass
    tm.restart();
    DBuf Cpb, Ab, nbar, W1, W2, Xbp, Xbc, g, lam, Uc, work;
    if ((st = dalloc(ctx, Cpb, 2 * M * sizeof(double))) || (st = dalloc(ctx, Ab, 2 * M * sizeof(double))) ||
        (st = dalloc(ctx, nbar, 6 * M * sizeof(double))) || (st = dalloc(ctx, W1, 8 * M * sizeof(double))) ||
        (st = dalloc(ctx, W2, 6 * M * sizeof(double))) || (st = dalloc(ctx, Xbp, 6 * M * sizeof(double))) ||
        (st = dalloc(ctx, Xbc, 6 * M * sizeof(double))) || (st = dalloc(ctx, g, 2 * M * sizeof(double))) ||
        (st = dalloc(ctx, lam, 2 * M * sizeof(double))) || (st = dalloc(ctx, Uc, 3 * M * sizeof(double))))
        return st;
    FP_HIP(ctx, hipMemsetAsync(Xbp.p, 0, 6 * M * sizeof(double), ctx->stream));
    FP_HIP(ctx, hipMemsetAsync(Xbc.p, 0, 6 * M * sizeof(double), ctx->stream));
    FP_HIP(ctx, hipMemsetAsync(g.p, 0, 2 * M * sizeof(double), ctx->stream));
    if ((st = fp_adj_objective(ctx, Cp.as<double>(), area, nrm, M, Lh, dh, 1.0 / o.ref_area, Cpb.as<double>(),
                               Ab.as<double>(), nbar.as<double>())))
        return st;
    col = 0;
    for (int k = 0; k < nb; ++k) {
        const fp_body *b = bodies[k];
        if (b->ns > 0 && (st = fp_adj_kutta(ctx, dsh[k].as<int64_t>(), b->ns, o.kutta_mode, Cpb.as<double>() + col,
                                            Cpb.as<double>() + M + col, 1)))
            return st;
        col += b->N;
    }
    if ((st = fp_adj_vbar(ctx, Cpb.as<double>(), V.as<double>(), M, 1.0 / (un * un), W1.as<double>()))) return st;
    double s1[8];
    if ((st = fp_mdot(ctx, W1.as<double>(), M, 8, ones.as<double>(), M, part.as<double>(), out8.as<double>()))) return st;
    FP_HIP(ctx, hipMemcpyAsync(s1, out8.p, 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    // probe pass: strength adjoints, probe-target and vertex adjoints
    size_t wb = 0;
    for (int k = 0; k < nb; ++k) wb = std::max(wb, fp_adj_pairs_work_bytes(ctx, bodies[k]->N, bodies[k]->nvp, M));
    if ((st = dalloc(ctx, work, wb))) return st;
    std::vector<DBuf> vbar(nb), hpart(nb);
    col = 0;
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        const size_t vb = (size_t)6 * b->nvp * b->N * sizeof(double);
        if ((st = dalloc(ctx, vbar[k], vb)) || (st = dalloc(ctx, hpart[k], (size_t)2 * 26 * std::max(1, b->nh) * sizeof(double))))
            return st;
        FP_HIP(ctx, hipMemsetAsync(vbar[k].p, 0, vb, ctx->stream));
        if ((st = fp_adj_pairs(ctx, b, prb, M, W1.as<double>(), 1, Xbp.as<double>(), vbar[k].as<double>(),
                               g.as<double>() + col, M, work.p, wb)))
            return st;
        if ((st = fp_adj_wake(ctx, b, prb, M, W1.as<double>(), 1, Xbp.as<double>(), hpart[k].as<double>(),
                              g.as<double>(), col)))
            return st;
        col += b->N;
    }
    // transpose solve A^T lambda_m = g_m
    {
        Timer tt(ctx->stream);
        fp_solve_stats ts{};
        SolveArgs at;
        at.solver = so.solver;
        at.n = n;
        at.ld = ld;
        at.geo = &t.tv;
        at.A64 = f64 ? AT.as<double>() : nullptr;
        if (!f64) {
            if ((st = dalloc(ctx, AT32, (size_t)n * ld * sizeof(float)))) return st;
            if ((st = fp_mat_narrow(ctx, AT.as<double>(), AT32.as<float>(), n, ld))) return st;
            at.A32 = AT32.as<float>();
        }
        // the FP64 residual of the transposed system needs an unfactorised copy
        DBuf ATr;
        const double *Ares = AT.as<double>();
        if (so.solver == FP_SOLVER_LU) {
            if ((st = dalloc(ctx, ATr, (size_t)n * ld * sizeof(double)))) return st;
            FP_HIP(ctx, hipMemcpyAsync(ATr.p, AT.p, (size_t)n * ld * sizeof(double), hipMemcpyDeviceToDevice,
                                       ctx->stream));
            Ares = ATr.as<double>();
        }
        opT.reset(new DenseOp<double>(ctx, Ares, n, ld));
        at.resid = opT.get();
        if (so.solver == FP_SOLVER_GMRES || so.solver == FP_SOLVER_GMRES_IR) {
            // Krylov solves touch the matrices read-only: the two coefficients' systems run at the
            // same time, the second on a stream of its own driven by a host thread (each is bound by
            // kernel latency at these sizes, so the two fill each other's gaps)
            st = transpose_solves_concurrent(ctx, at, so, Ares, g.as<double>(), lam.as<double>(), M, ts);
        } else {
            Solver ts_solver(ctx, at, so);
            st = ts_solver.setup(ts);
            for (int m = 0; m < 2 && !st; ++m)
                st = ts_solver.solve(g.as<double>() + m * M, lam.as<double>() + m * M, ts);
        }
        if (st) { if (stats_out) *stats_out = S; return st; }
        S.iterations_adjoint = ts.iterations;
        S.ms_transpose_solve = tt.stop();
    }
    AT.reset();
    AT32.reset();
    A.reset();
    A32.reset();
    // control-point pass: induced velocity at the control points, weights, pairs
    TargetView tc;
    tc.x = cpt; tc.y = cpt + M; tc.z = cpt + 2 * M;
    tc.nx = nrm; tc.ny = nrm + M; tc.nz = nrm + 2 * M;
    tc.M = M;
    FP_HIP(ctx, hipMemsetAsync(Uc.p, 0, 3 * M * sizeof(double), ctx->stream));
    for (int k = 0; k < nb; ++k)
        if ((st = fp_launch_field(ctx, bodies[k], bodies[k]->d_strength, tc, 0, Uc.as<double>(), fpart.as<double>(), fpb)))
            return st;
    if ((st = fp_adj_cpw(ctx, lam.as<double>(), nrm, Uc.as<double>(), M, U, W2.as<double>(), nbar.as<double>())))
        return st;
    double s2[8];
    if ((st = fp_mdot(ctx, W2.as<double>(), M, 6, ones.as<double>(), M, part.as<double>(), out8.as<double>()))) return st;
    FP_HIP(ctx, hipMemcpyAsync(s2, out8.p, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    col = 0;
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        if ((st = fp_adj_pairs(ctx, b, cpt, M, W2.as<double>(), 0, Xbc.as<double>(), vbar[k].as<double>(), nullptr,
                               0, work.p, wb)))
            return st;
        if ((st = fp_adj_wake(ctx, b, cpt, M, W2.as<double>(), 0, Xbc.as<double>(),
                              hpart[k].as<double>() + (size_t)26 * std::max(1, b->nh), nullptr, col)))
            return st;
        col += b->N;
    }
    // geometry reverse and node gather; horseshoe end points and directions on the host
    double dbar[2][3] = {{0, 0, 0}, {0, 0, 0}};
    col = 0;
    int64_t noff = 0;
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        const int64_t nn = (int64_t)b->nodes.size() / 3, N = b->N;
        if ((st = fp_adj_geom_rev(ctx, b, o.probe_offset, o.probe_charlen, o.probe_length, M, col, nbar.as<double>(),
                                  Xbp.as<double>(), Xbc.as<double>(), Ab.as<double>(), vbar[k].as<double>())))
            return st;
        // node -> panel-vertex slots (k * N + j)
        std::vector<int64_t> ptr(nn + 1, 0), ent;
        for (int64_t j = 0; j < N; ++j)
            for (int64_t q = b->cell_ptr[j]; q < b->cell_ptr[j + 1]; ++q) ptr[b->cell_idx[q] + 1]++;
        for (int64_t v = 0; v < nn; ++v) ptr[v + 1] += ptr[v];
        ent.resize(ptr[nn]);
        std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
        for (int64_t j = 0; j < N; ++j)
            for (int64_t q = b->cell_ptr[j]; q < b->cell_ptr[j + 1]; ++q)
                ent[fill[b->cell_idx[q]]++] = (q - b->cell_ptr[j]) * N + j;
        DBuf dptr, dent, dout;
        if ((st = dalloc(ctx, dptr, (nn + 1) * sizeof(int64_t))) ||
            (st = dalloc(ctx, dent, std::max<size_t>(1, ent.size()) * sizeof(int64_t))) ||
            (st = dalloc(ctx, dout, 6 * nn * sizeof(double))))
            return st;
        FP_HIP(ctx, hipMemcpyAsync(dptr.p, ptr.data(), ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(dent.p, ent.data(), ent.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        if ((st = fp_adj_nodes(ctx, b, dptr.as<int64_t>(), dent.as<int64_t>(), nn, vbar[k].as<double>(),
                               dout.as<double>())))
            return st;
        std::vector<double> nodeb(6 * nn), hp((size_t)2 * 26 * std::max(1, b->nh)), gam(N);
        FP_HIP(ctx, hipMemcpyAsync(nodeb.data(), dout.p, nodeb.size() * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(hp.data(), hpart[k].p, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(gam.data(), b->d_strength, N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        // horseshoes in the order build_horseshoes made them
        int h = 0;
        const size_t second = (size_t)26 * std::max(1, b->nh);
        for (int64_t e = 0; e < b->ns; ++e) {
            const int64_t *s = &b->shedding[6 * e];
            for (int side = 0; side < 2; ++side) {
                const int64_t pnl = side == 0 ? s[0] : s[3];
                if (pnl == -1) continue;
                const int64_t ia = b->cell_idx[b->cell_ptr[pnl] + s[3 * side + 1]];
                const int64_t ib = b->cell_idx[b->cell_ptr[pnl] + s[3 * side + 2]];
                const double G = gam[pnl];
                for (int m = 0; m < 2; ++m)
                    for (int d = 0; d < 3; ++d) {
                        const double *h0 = &hp[(size_t)h * 26 + 12 * m], *h1 = &hp[second + (size_t)h * 26 + 12 * m];
                        nodeb[(3 * m + d) * nn + ia] += G * (h0[d] + h1[d]);
                        nodeb[(3 * m + d) * nn + ib] += G * (h0[3 + d] + h1[3 + d]);
                        dbar[m][d] += G * (h0[6 + d] + h1[6 + d] + h0[9 + d] + h1[9 + d]);
                    }
                ++h;
            }
        }
        for (int64_t v = 0; v < nn; ++v)
            for (int d = 0; d < 3; ++d) {
                if (dCL_dn) dCL_dn[3 * (noff + v) + d] = nodeb[d * nn + v];
                if (dCD_dn) dCD_dn[3 * (noff + v) + d] = nodeb[(3 + d) * nn + v];
            }
        noff += nn;
        col += N;
    }
    // freestream adjoint: probe velocities, Cp normalisation, no-flow-through
    // RHS, and the direction Dhat = Uinf/|Uinf| (wake, Lhat, Dhat)
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double u4 = un * un * un * un;
    // objective directions: CL = Fn.Lhat, Lhat = w/|w|, w = Dhat x span; CDi = Fn.Dhat
    {
        const double lf = Lh[0] * Fn[0] + Lh[1] * Fn[1] + Lh[2] * Fn[2];
        const double wb3[3] = {(Fn[0] - Lh[0] * lf) / wn, (Fn[1] - Lh[1] * lf) / wn, (Fn[2] - Lh[2] * lf) / wn};
        // dbar += span x wbar
        dbar[0][0] += sd[1] * wb3[2] - sd[2] * wb3[1];
        dbar[0][1] += sd[2] * wb3[0] - sd[0] * wb3[2];
        dbar[0][2] += sd[0] * wb3[1] - sd[1] * wb3[0];
        for (int d = 0; d < 3; ++d) dbar[1][d] += Fn[d];
    }
    double Ub[2][3];
    for (int m = 0; m < 2; ++m) {
        const double dd = dh[0] * dbar[m][0] + dh[1] * dbar[m][1] + dh[2] * dbar[m][2];
        for (int d = 0; d < 3; ++d)
            Ub[m][d] = s1[3 * m + d] + 2 * s1[6 + m] * U[d] / u4 + s2[3 * m + d] + (dbar[m][d] - dh[d] * dd) / un;
    }
    if (dCL_dU)
        for (int d = 0; d < 3; ++d) dCL_dU[d] = Ub[0][d];
    if (dCD_dU)
        for (int d = 0; d < 3; ++d) dCD_dU[d] = Ub[1][d];
    S.ms_adjoint = tm.stop();
    if (stats_out) *stats_out = S;
    return FP_OK;
}

int fp_lift_drag(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *Uinf, const fp_liftdrag_opts *opts,
                 double *CL, double *CDi, fp_liftdrag_stats *stats)
{
    return liftdrag_run(ctx, bodies, nb, Uinf, opts, false, CL, CDi, nullptr, nullptr, nullptr, nullptr, stats);
}

int fp_lift_drag_adjoint(fp_ctx *ctx, fp_body *const *bodies, int nb, const double *Uinf, const fp_liftdrag_opts *opts,
                         double *dCL_dnodes, double *dCDi_dnodes, double *dCL_dUinf, double *dCDi_dUinf,
                         fp_liftdrag_stats *stats)
{
    return liftdrag_run(ctx, bodies, nb, Uinf, opts, true, nullptr, nullptr, dCL_dnodes, dCDi_dnodes, dCL_dUinf,
                        dCDi_dUinf, stats);
}


struct fp_batch {
    fp_ctx *ctx = nullptr;
    BatchImpl *impl = nullptr;
};

int fp_batch_create(fp_ctx *ctx, fp_body *const *bodies, int nb, const fp_solve_opts *opts, fp_batch **out)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    if (!out) return fp_fail(ctx, FP_ERR_ARG, "null output");
    *out = nullptr;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<BatchImpl> B(new BatchImpl());
    B->ctx = ctx;
    if (opts) B->o = *opts;
    else fp_solve_opts_default(&B->o, FP_SOLVER_LU32_IR);
    if (B->o.solver != FP_SOLVER_LU && B->o.solver != FP_SOLVER_LU32_IR)
        return fp_fail(ctx, FP_ERR_UNSUPPORTED, "a body batch runs FP_SOLVER_LU or FP_SOLVER_LU32_IR");
    B->sweeps = std::max(1, std::min(B->o.max_refine, 3));
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        if (b->element != FP_ELEM_VORTEXRING) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "a body batch takes vortex-ring bodies");
        if (!b->has_uinf) return fp_fail(ctx, FP_ERR_SHAPE, "Invalid Uinfs; call fp_body_set_freestream before solve");
        if (b->nvp != bodies[0]->nvp) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "bodies of a batch share the panel size class");
        for (int q = 0; q < k; ++q)
            if (bodies[q] == b) return fp_fail(ctx, FP_ERR_ARG, "a body appears twice in the batch");
        std::unique_ptr<BatchBody> bb(new BatchBody());
        bb->b = b;
        bb->n = b->N;
        bb->ld = round_ld(b->N);
        if ((st = gather_targets(ctx, &bodies[k], 1, bb->t))) return st;
        std::vector<fp_body *> one(1, b);
        bb->mf.reset(new MatrixFreeOp(ctx, one, bb->t.tv));
        if ((st = bb->mf->init())) return st;
        B->bodies.push_back(std::move(bb));
    }
    B->nvp = bodies[0]->nvp;
    if ((st = B->alloc())) return st;
    if ((st = B->autotune())) return st;
    if ((st = B->capture())) return st;
    fp_batch *h = new fp_batch();
    h->ctx = ctx;
    h->impl = B.release();
    *out = h;
    return FP_OK;
}

void fp_batch_destroy(fp_batch *h)
{
    if (!h) return;
    if (h->ctx) {
        (void)hipSetDevice(h->ctx->device);
        (void)hipStreamSynchronize(h->ctx->stream);
    }
    delete h->impl;
    delete h;
}

int fp_batch_solve(fp_ctx *ctx, fp_batch *h, fp_batch_stats *stats)
{
    if (!ctx || !h || h->ctx != ctx) return FP_ERR_ARG;
    BatchImpl &B = *h->impl;
    FP_HIP(ctx, hipSetDevice(ctx->device));
    fp_batch_stats S{};
    int st;
    bool stale = B.recapture;
    for (auto &bb : B.bodies) stale = stale || bb->version != bb->b->version || !bb->b->has_uinf;
    if (stale) {
        for (auto &bb : B.bodies) {
            if (!bb->b->has_uinf) return fp_fail(ctx, FP_ERR_SHAPE, "Invalid Uinfs; call fp_body_set_freestream before solve");
            if ((st = gather_targets(ctx, &bb->b, 1, bb->t))) return st;
            std::vector<fp_body *> one(1, bb->b);
            bb->mf.reset(new MatrixFreeOp(ctx, one, bb->t.tv));
            if ((st = bb->mf->init())) return st;
        }
        if ((st = B.build_items())) return st;
        Timer tc(ctx->stream);
        if ((st = B.capture())) return st;
        S.ms_capture = tc.stop();
        S.recaptured = 1;
    }
    Timer t(ctx->stream);
    FP_HIP(ctx, hipGraphLaunch(B.exec, ctx->stream));
    S.ms_graph = t.stop();
    const size_t nb = B.bodies.size();
    std::vector<double> nr(2 * nb);
    FP_HIP(ctx, hipMemcpyAsync(nr.data(), B.norms.p, nr.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    bool refall = false;
    for (auto &bb : B.bodies) {
        int info = 0;
        FP_HIP(ctx, hipMemcpy(&info, bb->info.p, sizeof(int), hipMemcpyDeviceToHost));
        if (info && B.lu32() && !bb->pivoted) {
            // the randomized no-pivot LU met a bad pivot: this body switches to the pivoted LU
            if ((st = B.fallback(*bb))) return st;
            refall = true;
            continue;
        }
        if (info) {
            char buf[128];
            snprintf(buf, sizeof buf, "singular matrix: zero pivot at column %d", info);
            return fp_fail(ctx, FP_ERR_SINGULAR, buf);
        }
    }
    if (refall) {
        FP_HIP(ctx, hipMemcpyAsync(nr.data(), B.norms.p, nr.size() * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    // bodies that need more refinement continue eagerly on the context stream
    Timer te(ctx->stream);
    int extra = 0;
    double worst = 0;
    for (size_t k = 0; k < nb; ++k) {
        BatchBody &bb = *B.bodies[k];
        double rel = nr[k] > 0 ? std::sqrt(nr[nb + k] / nr[k]) : (nr[nb + k] == 0 ? 0.0 : NAN);
        int steps = B.lu32() ? B.sweeps : 0;
        for (;;) {
        while (B.lu32() && !converged(rel, B.o.tol) && std::isfinite(rel) && steps < B.o.max_refine) {
            if ((st = B.sweep(bb))) return st;
            if ((st = fp_colnorm2(ctx, bb.r.as<double>(), bb.n, bb.n, 1, B.norms.as<double>() + nb + k))) return st;
            double v = 0;
            FP_HIP(ctx, hipMemcpyAsync(&v, B.norms.as<double>() + nb + k, sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
            FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            rel = nr[k] > 0 ? std::sqrt(v / nr[k]) : (v == 0 ? 0.0 : NAN);
            ++steps;
            ++extra;
            FP_HIP(ctx, hipMemcpyAsync(bb.b->d_strength, bb.x.p, bb.n * sizeof(double), hipMemcpyDeviceToDevice,
                                       ctx->stream));
        }
        if (B.lu32() && !bb.pivoted && !converged(rel, B.o.tol) && std::isfinite(nr[k])) {
            // the randomized factors did not bring the residual down: pivot this body and refine again
            if ((st = B.fallback(bb))) return st;
            double v = 0;
            FP_HIP(ctx, hipMemcpyAsync(&v, B.norms.as<double>() + nb + k, sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
            FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
            rel = nr[k] > 0 ? std::sqrt(v / nr[k]) : (v == 0 ? 0.0 : NAN);
            steps = B.sweeps;
            continue;
        }
        break;
        }
        worst = worst_residual(worst, rel);
        S.panels += bb.n;
    }
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    S.ms_extra = te.stop();
    S.extra_sweeps = extra;
    S.sweeps_in_graph = B.lu32() ? B.sweeps : 0;
    S.residual = worst;
    S.graph_nodes = B.graph_nodes;
    S.tile_targets = B.tm;
    S.queue_workgroups = B.blocks;
    S.nbodies = (int)nb;
    for (auto &bb : B.bodies) bb->b->solved = true;
    if (stats) *stats = S;
    if (!converged(worst, B.o.tol)) return not_converged(ctx, worst, "batch refinement");
    return FP_OK;
}

int fp_batch_autotune_get(const fp_batch *h, int k, int *tile_targets, int *workgroups, double *ms)
{
    if (!h || k < 0 || k >= h->impl->ntune) return FP_ERR_ARG;
    if (tile_targets) *tile_targets = h->impl->tune_tm[k];
    if (workgroups) *workgroups = h->impl->tune_blocks[k];
    if (ms) *ms = h->impl->tune_ms[k];
    return FP_OK;
}

// neighbour across edge k (vertices k, k+1) of every cell, the cell sharing
// both end points (gt.neighbor with preserveEdge), uploaded once per body
static int body_neighbors(fp_ctx *ctx, fp_body *b)
{
    if (b->d_nbr) return FP_OK;
    const int64_t N = b->N, nvp = b->nvp;
    std::map<std::pair<int64_t, int64_t>, std::vector<int64_t>> owners;
    for (int64_t j = 0; j < N; ++j) {
        const int64_t p0 = b->cell_ptr[j], nv = b->cell_ptr[j + 1] - p0;
        for (int64_t k = 0; k < nv; ++k) {
            int64_t a = b->cell_idx[p0 + k], c = b->cell_idx[p0 + (k + 1) % nv];
            if (a > c) std::swap(a, c);
            owners[{a, c}].push_back(j);
        }
    }
    std::vector<int> nb((size_t)nvp * N, -1);
    for (int64_t j = 0; j < N; ++j) {
        const int64_t p0 = b->cell_ptr[j], nv = b->cell_ptr[j + 1] - p0;
        for (int64_t k = 0; k < nv; ++k) {
            int64_t a = b->cell_idx[p0 + k], c = b->cell_idx[p0 + (k + 1) % nv];
            if (a > c) std::swap(a, c);
            for (int64_t o : owners[{a, c}])
                if (o != j) { nb[(size_t)k * N + j] = (int)o; break; }
        }
    }
    FP_HIP(ctx, hipMalloc((void **)&b->d_nbr, nb.size() * sizeof(int)));
    FP_HIP(ctx, hipMemcpy(b->d_nbr, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice));
    return FP_OK;
}

int fp_calcfield_ugradmu(fp_ctx *ctx, fp_body *const *bodies, int nb, int te_correction, double *out)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    if (!out) return fp_fail(ctx, FP_ERR_ARG, "null out");
    int64_t M = 0;
    for (int k = 0; k < nb; ++k) {
        if (!bodies[k]->solved)
            return fp_fail(ctx, FP_ERR_UNSOLVED, "Source body hasn't been solved yet. Please call `solve(...)` function first.");
        if (bodies[k]->element != FP_ELEM_VORTEXRING)
            return fp_fail(ctx, FP_ERR_UNSUPPORTED, "the doublet-gradient velocity needs vortex-ring bodies");
        M += bodies[k]->N;
    }
    FP_HIP(ctx, hipSetDevice(ctx->device));
    DBuf o;
    if ((st = dalloc(ctx, o, 3 * M * sizeof(double)))) return st;
    FP_HIP(ctx, hipMemcpyAsync(o.p, out, 3 * M * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int64_t off = 0;
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        if ((st = body_neighbors(ctx, b))) return st;
        if (b->ns > 0 && !b->d_sh) {
            FP_HIP(ctx, hipMalloc((void **)&b->d_sh, 6 * b->ns * sizeof(int64_t)));
            FP_HIP(ctx, hipMemcpy(b->d_sh, b->shedding.data(), 6 * b->ns * sizeof(int64_t), hipMemcpyHostToDevice));
        }
        if ((st = fp_launch_ugradmu(ctx, b, b->d_nbr, b->d_sh, te_correction, o.as<double>() + 3 * off))) return st;
        off += b->N;
    }
    FP_HIP(ctx, hipMemcpyAsync(out, o.p, 3 * M * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

}  // extern "C"

namespace {

int ugradmu_inputs(fp_ctx *ctx, fp_body *const *bodies, int nb, int64_t *M)
{
    int st = check_bodies(ctx, bodies, nb);
    if (st) return st;
    *M = 0;
    for (int k = 0; k < nb; ++k) {
        if (!bodies[k]->solved)
            return fp_fail(ctx, FP_ERR_UNSOLVED, "Source body hasn't been solved yet. Please call `solve(...)` function first.");
        if (bodies[k]->element != FP_ELEM_VORTEXRING)
            return fp_fail(ctx, FP_ERR_UNSUPPORTED, "the doublet-gradient velocity needs vortex-ring bodies");
        *M += bodies[k]->N;
    }
    return FP_OK;
}

// node -> cells lists (CSR over every vertex of every cell) and the first three
// nodes of each cell, uploaded for one call
struct NodeData {
    DBuf nptr, ncell, cn, nodal;
    int64_t nn = 0;
    int build(fp_ctx *ctx, const fp_body *b)
    {
        int st;
        nn = (int64_t)b->nodes.size() / 3;
        const int64_t N = b->N;
        std::vector<int64_t> ptr(nn + 1, 0);
        for (int64_t c = 0; c < N; ++c)
            for (int64_t q = b->cell_ptr[c]; q < b->cell_ptr[c + 1]; ++q) ++ptr[b->cell_idx[q] + 1];
        for (int64_t v = 0; v < nn; ++v) ptr[v + 1] += ptr[v];
        std::vector<int> cells(ptr[nn]), cn3(3 * N);
        std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
        for (int64_t c = 0; c < N; ++c) {
            for (int64_t q = b->cell_ptr[c]; q < b->cell_ptr[c + 1]; ++q) cells[fill[b->cell_idx[q]]++] = (int)c;
            for (int k = 0; k < 3; ++k) cn3[3 * c + k] = (int)b->cell_idx[b->cell_ptr[c] + k];
        }
        if ((st = dalloc(ctx, nptr, (nn + 1) * sizeof(int64_t)))) return st;
        if ((st = dalloc(ctx, ncell, std::max<size_t>(1, cells.size()) * sizeof(int)))) return st;
        if ((st = dalloc(ctx, cn, 3 * N * sizeof(int)))) return st;
        if ((st = dalloc(ctx, nodal, std::max<int64_t>(1, nn) * sizeof(double)))) return st;
        FP_HIP(ctx, hipMemcpyAsync(nptr.p, ptr.data(), (nn + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        if (!cells.empty())
            FP_HIP(ctx, hipMemcpyAsync(ncell.p, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice,
                                       ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(cn.p, cn3.data(), 3 * N * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        // the host vectors go out of scope: wait for the copies
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return FP_OK;
    }
};

// Sharp trailing edge: side 1 = the shedding panels pi (with pi + 1 when `pairs`,
// the two triangles of a structured triangulated grid's last quad), side 2 = the
// partners pj (with pj - 1); the TE nodes are the shedding edges' nodes.
struct TeTables {
    DBuf side, te, node_te, vals;
    TeSides view;
    int build(fp_ctx *ctx, const fp_body *b, int64_t nn, bool pairs)
    {
        const int64_t N = b->N;
        std::vector<int> side(N, 0), node_te(std::max<int64_t>(1, nn), -1), te;
        auto mark = [&](int64_t c, int s) { if (c >= 0 && c < N) side[c] = s; };
        for (int s = 1; s <= 2; ++s)   // side 2 last: it wins where the two overlap
            for (int64_t e = 0; e < b->ns; ++e) {
                const int64_t p = b->shedding[6 * e + 3 * (s - 1)];
                if (p < 0) continue;
                mark(p, s);
                if (pairs) mark(s == 1 ? p + 1 : p - 1, s);
            }
        for (int64_t e = 0; e < b->ns; ++e)
            for (int sd = 0; sd < 2; ++sd) {
                const int64_t p = b->shedding[6 * e + 3 * sd];
                if (p < 0) continue;
                for (int q = 1; q <= 2; ++q) {
                    const int64_t v = b->cell_idx[b->cell_ptr[p] + b->shedding[6 * e + 3 * sd + q]];
                    if (node_te[v] < 0) { node_te[v] = (int)te.size(); te.push_back((int)v); }
                }
            }
        int st;
        const int nte = (int)te.size();
        if ((st = dalloc(ctx, this->side, N * sizeof(int))) || (st = dalloc(ctx, this->te, std::max(1, nte) * sizeof(int))) ||
            (st = dalloc(ctx, this->node_te, node_te.size() * sizeof(int))) ||
            (st = dalloc(ctx, vals, 2 * std::max(1, nte) * sizeof(double))))
            return st;
        FP_HIP(ctx, hipMemcpyAsync(this->side.p, side.data(), N * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        if (nte)
            FP_HIP(ctx, hipMemcpyAsync(this->te.p, te.data(), nte * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipMemcpyAsync(this->node_te.p, node_te.data(), node_te.size() * sizeof(int), hipMemcpyHostToDevice,
                                   ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));   // the host vectors go out of scope
        view.cell_side = this->side.as<int>();
        view.te_nodes = this->te.as<int>();
        view.node_te = this->node_te.as<int>();
        view.sidev = vals.as<double>();
        view.nte = nte;
        return FP_OK;
    }
};

}  // namespace

extern "C" {

int fp_calcfield_ugradmu_node(fp_ctx *ctx, fp_body *const *bodies, int nb, int sharp_te, double *out,
                              double *nodal_out)
{
    int64_t M = 0;
    int st = ugradmu_inputs(ctx, bodies, nb, &M);
    if (st) return st;
    if (!out) return fp_fail(ctx, FP_ERR_ARG, "null out");
    if (sharp_te < FP_SHARPTE_OFF || sharp_te > FP_SHARPTE_CELLS) return fp_fail(ctx, FP_ERR_ARG, "unknown sharp_te mode");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    DBuf o;
    if ((st = dalloc(ctx, o, 3 * M * sizeof(double)))) return st;
    FP_HIP(ctx, hipMemcpyAsync(o.p, out, 3 * M * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int64_t off = 0, noff = 0;
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        NodeData nd;
        if ((st = nd.build(ctx, b))) return st;
        TeTables tt;
        const bool sharp = sharp_te != FP_SHARPTE_OFF && b->ns > 0;
        if (sharp && (st = tt.build(ctx, b, nd.nn, sharp_te == FP_SHARPTE_PAIRS))) return st;
        if ((st = fp_launch_ugradmu_node(ctx, b, nd.nptr.as<int64_t>(), nd.ncell.as<int>(), nd.nn, nd.cn.as<int>(),
                                         nd.nodal.as<double>(), o.as<double>() + 3 * off, sharp ? &tt.view : nullptr)))
            return st;
        if (nodal_out && nd.nn > 0)
            FP_HIP(ctx, hipMemcpyAsync(nodal_out + noff, nd.nodal.p, nd.nn * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        off += b->N;
        noff += nd.nn;
    }
    FP_HIP(ctx, hipMemcpyAsync(out, o.p, 3 * M * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

int fp_calcfield_ugradmu_combined(fp_ctx *ctx, fp_body *const *bodies, int nb, int force_cell_te,
                                  double anglecrit_deg, int sharp_te, double *out)
{
    int64_t M = 0;
    int st = ugradmu_inputs(ctx, bodies, nb, &M);
    if (st) return st;
    if (!out) return fp_fail(ctx, FP_ERR_ARG, "null out");
    if (sharp_te < FP_SHARPTE_OFF || sharp_te > FP_SHARPTE_CELLS) return fp_fail(ctx, FP_ERR_ARG, "unknown sharp_te mode");
    FP_HIP(ctx, hipSetDevice(ctx->device));
    DBuf o, cellv, nodev;
    if ((st = dalloc(ctx, o, 3 * M * sizeof(double)))) return st;
    FP_HIP(ctx, hipMemcpyAsync(o.p, out, 3 * M * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const double cosc = std::cos(anglecrit_deg * 3.14159265358979323846 / 180.0);
    int64_t off = 0;
    for (int k = 0; k < nb; ++k) {
        fp_body *b = bodies[k];
        const int64_t N = b->N;
        if ((st = body_neighbors(ctx, b))) return st;
        if (b->ns > 0 && !b->d_sh) {
            FP_HIP(ctx, hipMalloc((void **)&b->d_sh, 6 * b->ns * sizeof(int64_t)));
            FP_HIP(ctx, hipMemcpy(b->d_sh, b->shedding.data(), 6 * b->ns * sizeof(int64_t), hipMemcpyHostToDevice));
        }
        if ((st = dalloc(ctx, cellv, 3 * std::max<int64_t>(1, N) * sizeof(double)))) return st;
        if ((st = dalloc(ctx, nodev, 3 * std::max<int64_t>(1, N) * sizeof(double)))) return st;
        FP_HIP(ctx, hipMemsetAsync(cellv.p, 0, 3 * N * sizeof(double), ctx->stream));
        FP_HIP(ctx, hipMemsetAsync(nodev.p, 0, 3 * N * sizeof(double), ctx->stream));
        NodeData nd;
        if ((st = nd.build(ctx, b))) return st;
        TeTables tt;
        const bool sharp = sharp_te != FP_SHARPTE_OFF && b->ns > 0;
        if (sharp && (st = tt.build(ctx, b, nd.nn, sharp_te == FP_SHARPTE_PAIRS))) return st;
        // cell-centred values with the trailing-edge treatment, node-based values, then the choice per cell
        if ((st = fp_launch_ugradmu(ctx, b, b->d_nbr, b->d_sh, 1, cellv.as<double>()))) return st;
        if ((st = fp_launch_ugradmu_node(ctx, b, nd.nptr.as<int64_t>(), nd.ncell.as<int>(), nd.nn, nd.cn.as<int>(),
                                         nd.nodal.as<double>(), nodev.as<double>(), sharp ? &tt.view : nullptr)))
            return st;
        if ((st = fp_launch_ugradmu_combine(ctx, b, b->d_nbr, b->d_sh, force_cell_te, cosc, cellv.as<double>(),
                                            nodev.as<double>(), o.as<double>() + 3 * off)))
            return st;
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        off += N;
    }
    FP_HIP(ctx, hipMemcpyAsync(out, o.p, 3 * M * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

}  // extern "C"
