// fp_influence.hip -- panel geometry, AIC assembly and field kernels (gfx950).
//
// Assembly (src/FLOWPanel_liftingbody.jl:455-572, src/FLOWPanel_nonliftingbody.jl:141-186):
//   one lane per panel (its vertices and edge vectors in registers), a tile of
//   TM targets (control point + normal) staged in LDS and read by broadcast;
//   every G(i, j) is written once, a wavefront storing 64 consecutive panels of
//   one row.  Wake horseshoes are added afterwards by one lane per target that
//   walks the shedding list in order (no two lanes touch the same entry).
// Field sums (src/FLOWPanel_liftingbody.jl:576-742, src/FLOWPanel_nonliftingbody.jl:188-249):
//   one lane per target (accumulating in FP64 registers), panels staged in LDS
//   tiles shared by the 256 lanes; the panel range is split over grid.y when
//   there are few targets and the splits are summed in a fixed order.
#include "fp_internal.h"
#include "fp_elements.h"
#include "fp_geometry.h"

#include <algorithm>

static_assert(FPG_CHARLEN_UNITARY == FP_CHARLEN_UNITARY && FPG_CHARLEN_BBOX == FP_CHARLEN_BBOX &&
                  FPG_CHARLEN_MAXDIST == FP_CHARLEN_MAXDIST && FPG_CHARLEN_SQRTAREA == FP_CHARLEN_SQRTAREA,
              "characteristic-length kinds");

// ----------------------------------------------------------------------------
// geometry: one thread per panel, FP64 (src/FLOWPanel_abstractbody.jl:635-927)
template <int NVP>
__global__ __launch_bounds__(256) void k_geometry(const double *__restrict__ pv, const int *__restrict__ nvv,
                                                  int64_t N, double cpoff, int charlen, double *__restrict__ geo,
                                                  double *__restrict__ hs)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double p[NVP][3];
#pragma unroll
    for (int k = 0; k < NVP; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) p[k][d] = pv[(3 * k + d) * N + j];
    const int nv = nvv[j];
    double t[3], n[3], o[3], m[3], area, l;
    panel_geom<double, NVP>(p, nv, charlen, t, n, o, m, area, l);
    for (int d = 0; d < 3; ++d) m[d] += cpoff * l * n[d];
    geo[0 * N + j] = n[0]; geo[1 * N + j] = n[1]; geo[2 * N + j] = n[2];
    geo[3 * N + j] = m[0]; geo[4 * N + j] = m[1]; geo[5 * N + j] = m[2];
    geo[6 * N + j] = area;
    // Hess-Smith frame: x = o, y = t, z = -n, origin = first vertex
    const double ex[3] = {o[0], o[1], o[2]}, ey[3] = {t[0], t[1], t[2]}, ez[3] = {-n[0], -n[1], -n[2]};
    double *h = hs + j;
    for (int d = 0; d < 3; ++d) {
        h[(0 + d) * N] = ex[d]; h[(3 + d) * N] = ey[d]; h[(6 + d) * N] = ez[d]; h[(9 + d) * N] = p[0][d];
    }
    double xl[NVP], yl[NVP], zl[NVP];
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        const double dx = p[k][0] - p[0][0], dy = p[k][1] - p[0][1], dz = p[k][2] - p[0][2];
        xl[k] = ex[0] * dx + ex[1] * dy + ex[2] * dz;
        yl[k] = ey[0] * dx + ey[1] * dy + ey[2] * dz;
        zl[k] = ez[0] * dx + ez[1] * dy + ez[2] * dz;
        h[(12 + k) * N] = xl[k]; h[(12 + NVP + k) * N] = yl[k]; h[(12 + 2 * NVP + k) * N] = zl[k];
    }
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        const int q = (k + 1) % NVP;
        const double dd = sqrt((xl[q] - xl[k]) * (xl[q] - xl[k]) + (yl[q] - yl[k]) * (yl[q] - yl[k]) +
                               (zl[q] - zl[k]) * (zl[q] - zl[k]));
        h[(12 + 3 * NVP + k) * N] = dd;
        h[(12 + 4 * NVP + k) * N] = dd > 0 ? (yl[q] - yl[k]) / dd : 0.0;
        h[(12 + 5 * NVP + k) * N] = dd > 0 ? (xl[q] - xl[k]) / dd : 0.0;
    }
}

int fp_launch_geometry(fp_ctx *ctx, fp_body *b)
{
    const int64_t N = b->N;
    if (N == 0) return FP_OK;
    dim3 grid((unsigned)((N + 255) / 256));
    if (b->nvp == 4)
        hipLaunchKernelGGL(k_geometry<4>, grid, dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, N, b->cp_offset,
                           b->charlen, b->d_geo, b->d_hs);
    else
        hipLaunchKernelGGL(k_geometry<8>, grid, dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, N, b->cp_offset,
                           b->charlen, b->d_geo, b->d_hs);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// AIC assembly, vortex ring: one lane per panel, TM targets per block in LDS.
// One tile = 256 panels (block bx) x TM targets (tile by); the caller owns the
// LDS tile s and the barriers around a tile.
template <typename T, typename TO, int NVP, int TM>
__device__ __forceinline__ void vr_load_targets(const TargetView &tv, int64_t i0, T (&s)[6][TM])
{
    for (int t = threadIdx.x; t < TM; t += blockDim.x) {
        const int64_t i = i0 + t;
        const bool ok = i < tv.M;
        s[0][t] = ok ? T(tv.x[i]) : T(0);
        s[1][t] = ok ? T(tv.y[i]) : T(0);
        s[2][t] = ok ? T(tv.z[i]) : T(0);
        s[3][t] = ok ? T(tv.nx[i]) : T(0);
        s[4][t] = ok ? T(tv.ny[i]) : T(0);
        s[5][t] = ok ? T(tv.nz[i]) : T(0);
    }
}

template <typename T, typename TO, int NVP, int TM>
__device__ __forceinline__ void vr_tile(const PanelView &pan, const TargetView &tv, TO *__restrict__ G, int64_t ld,
                                        int64_t col0, int64_t bx, int64_t i0, const T (&s)[6][TM],
                                        float *__restrict__ G2 = nullptr)
{
    const int64_t j = bx * blockDim.x + threadIdx.x;
    if (j >= pan.N) return;
    const int64_t N = pan.N;
    T px[NVP], py[NVP], pz[NVP], ex[NVP], ey[NVP], ez[NVP];
    double q[NVP][3];
#pragma unroll
    for (int k = 0; k < NVP; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) q[k][d] = pan.pv[(3 * k + d) * N + j];
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        const int kk = (k + 1) % NVP;
        px[k] = T(q[k][0]); py[k] = T(q[k][1]); pz[k] = T(q[k][2]);
        ex[k] = T(q[kk][0] - q[k][0]); ey[k] = T(q[kk][1] - q[k][1]); ez[k] = T(q[kk][2] - q[k][2]);
    }
    const T off = T(pan.off), off2 = T(pan.off * pan.off), cut2 = T(pan.cut * pan.cut);
    const int iend = (int)min((int64_t)TM, tv.M - i0);
    TO *g = G + i0 * ld + col0 + j;
    float *g2 = G2 ? G2 + i0 * ld + col0 + j : nullptr;
    // two targets per iteration for quads (more independent work per wave; octagons
    // would drop to two waves per SIMD)
    constexpr int TU = NVP == 4 ? 2 : 1;
#pragma unroll TU
    for (int t = 0; t < iend; ++t) {
        T ux = 0, uy = 0, uz = 0;
        ring_u<T, NVP>(px, py, pz, ex, ey, ez, s[0][t], s[1][t], s[2][t], off, off2, cut2, ux, uy, uz);
        const T v = ux * s[3][t] + uy * s[4][t] + uz * s[5][t];
        g[(int64_t)t * ld] = TO(v);
        if (g2) g2[(int64_t)t * ld] = float(v);
    }
}

template <typename T, typename TO, int NVP, int TM>
__global__ __launch_bounds__(256) void k_assemble_vr(PanelView pan, TargetView tv, TO *__restrict__ G, int64_t ld,
                                                     int64_t col0, float *__restrict__ G2)
{
    __shared__ T s[6][TM];
    const int64_t i0 = (int64_t)blockIdx.y * TM;
    vr_load_targets<T, TO, NVP, TM>(tv, i0, s);
    __syncthreads();
    vr_tile<T, TO, NVP, TM>(pan, tv, G, ld, col0, blockIdx.x, i0, s, G2);
}

// The diagonal blocks of one vortex-ring body's AIC for the block-Jacobi
// inverses, rounded to FP32 into a compact identity-padded array (block b at
// X + b block^2, row-major, leading dimension block): grid (column tiles of 256,
// target tiles of TM, blocks); one lane per panel as in k_assemble_vr.
template <int NVP, int TM>
__global__ __launch_bounds__(256) void k_assemble_diag_vr(PanelView pan, TargetView tv, int block,
                                                          float *__restrict__ X)
{
    __shared__ double s[6][TM];
    const int64_t b0 = (int64_t)blockIdx.z * block;
    const int bs = (int)min((int64_t)block, pan.N - b0);
    const int r0 = blockIdx.y * TM;
    for (int t = threadIdx.x; t < TM; t += blockDim.x) {
        const bool ok = r0 + t < bs;
        const int64_t i = b0 + r0 + t;
        s[0][t] = ok ? tv.x[i] : 0.0;
        s[1][t] = ok ? tv.y[i] : 0.0;
        s[2][t] = ok ? tv.z[i] : 0.0;
        s[3][t] = ok ? tv.nx[i] : 0.0;
        s[4][t] = ok ? tv.ny[i] : 0.0;
        s[5][t] = ok ? tv.nz[i] : 0.0;
    }
    __syncthreads();
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= block) return;
    float *xb = X + (int64_t)blockIdx.z * block * block;
    const int rend = min(TM, block - r0);
    if (c >= bs) {
        for (int t = 0; t < rend; ++t) xb[(int64_t)(r0 + t) * block + c] = r0 + t == c ? 1.f : 0.f;
        return;
    }
    const int64_t N = pan.N, j = b0 + c;
    double px[NVP], py[NVP], pz[NVP], ex[NVP], ey[NVP], ez[NVP];
    double q[NVP][3];
#pragma unroll
    for (int k = 0; k < NVP; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) q[k][d] = pan.pv[(3 * k + d) * N + j];
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        const int kk = (k + 1) % NVP;
        px[k] = q[k][0]; py[k] = q[k][1]; pz[k] = q[k][2];
        ex[k] = q[kk][0] - q[k][0]; ey[k] = q[kk][1] - q[k][1]; ez[k] = q[kk][2] - q[k][2];
    }
    const double off = pan.off, off2 = pan.off * pan.off, cut2 = pan.cut * pan.cut;
    for (int t = 0; t < rend; ++t) {
        const int r = r0 + t;
        float v = r == c ? 1.f : 0.f;
        if (r < bs) {
            double ux = 0, uy = 0, uz = 0;
            ring_u<double, NVP>(px, py, pz, ex, ey, ez, s[0][t], s[1][t], s[2][t], off, off2, cut2, ux, uy, uz);
            v = float(ux * s[3][t] + uy * s[4][t] + uz * s[5][t]);
        }
        xb[(int64_t)r * block + c] = v;
    }
}

// Batched assembly of independent bodies (each its own matrix): persistent
// workgroups pull (body, panel block, target tile) items from a device queue
// (one atomic counter) until the queue is empty.  Every body's items are the
// same size, so uneven bodies share the chip without per-body launches.
template <typename T, typename TO, int NVP, int TM>
__global__ __launch_bounds__(256) void k_assemble_vr_queue(const AssemblyItem *__restrict__ items, int nitems,
                                                           int64_t total, int *__restrict__ counter)
{
    __shared__ T s[6][TM];
    __shared__ int64_t cur;
    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd((unsigned long long *)counter, 1ull);
        __syncthreads();
        const int64_t it = cur;
        if (it >= total) break;
        int b = 0;
        while (b + 1 < nitems && items[b + 1].tile0 <= it) ++b;
        const AssemblyItem &a = items[b];
        const int64_t local = it - a.tile0;
        const int64_t by = local / a.nbx, bx = local % a.nbx;
        vr_load_targets<T, TO, NVP, TM>(a.tv, by * TM, s);
        __syncthreads();
        vr_tile<T, TO, NVP, TM>(a.pan, a.tv, (TO *)a.G, a.ld, 0, bx, by * TM, s);
        __syncthreads();
    }
}

__global__ void k_zero_counter(int *counter) { *(unsigned long long *)counter = 0ull; }

// AIC assembly, constant source (Hess-Smith), one lane per panel.
template <typename T, typename TO, int NVP, int TM>
__global__ __launch_bounds__(256) void k_assemble_source(PanelView pan, TargetView tv, TO *__restrict__ G,
                                                         int64_t ld, int64_t col0, float *__restrict__ G2)
{
    __shared__ T s[6][TM];
    const int64_t i0 = (int64_t)blockIdx.y * TM;
    for (int t = threadIdx.x; t < TM; t += blockDim.x) {
        const int64_t i = i0 + t;
        const bool ok = i < tv.M;
        s[0][t] = ok ? T(tv.x[i]) : T(0);
        s[1][t] = ok ? T(tv.y[i]) : T(0);
        s[2][t] = ok ? T(tv.z[i]) : T(0);
        s[3][t] = ok ? T(tv.nx[i]) : T(0);
        s[4][t] = ok ? T(tv.ny[i]) : T(0);
        s[5][t] = ok ? T(tv.nz[i]) : T(0);
    }
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= pan.N) return;
    const int64_t N = pan.N;
    constexpr int S = 12 + 6 * NVP;
    T h[S];
#pragma unroll
    for (int c = 0; c < S; ++c) h[c] = T(pan.hs[c * N + j]);
    const int nv = pan.nv[j];
    const T off = T(pan.off);
    const T k4 = T(1.0 / (4 * FP_PI));
    const int iend = (int)min((int64_t)TM, tv.M - i0);
    TO *g = G + i0 * ld + col0 + j;
    float *g2 = G2 ? G2 + i0 * ld + col0 + j : nullptr;
    for (int t = 0; t < iend; ++t) {
        const T dx = s[0][t] - h[9], dy = s[1][t] - h[10], dz = s[2][t] - h[11];
        const T x = h[0] * dx + h[1] * dy + h[2] * dz;
        const T y = h[3] * dx + h[4] * dy + h[5] * dz;
        const T z = h[6] * dx + h[7] * dy + h[8] * dz;
        T V1, V2, V3, ps;
        hs_sums<T, NVP, true>(h + 12, h + 12 + NVP, h + 12 + 2 * NVP, h + 12 + 3 * NVP, h + 12 + 4 * NVP,
                              h + 12 + 5 * NVP, nv, x, y, z, off, V1, V2, V3, ps);
        V1 *= k4; V2 *= k4; V3 *= k4;
        const T ux = V1 * h[0] + V2 * h[3] + V3 * h[6];
        const T uy = V1 * h[1] + V2 * h[4] + V3 * h[7];
        const T uz = V1 * h[2] + V2 * h[5] + V3 * h[8];
        const T v = ux * s[3][t] + uy * s[4][t] + uz * s[5][t];
        g[(int64_t)t * ld] = TO(v);
        if (g2) g2[(int64_t)t * ld] = float(v);
    }
}

// Wake horseshoes into their columns: one lane per target, shedding list in order.
// dblk > 0: G is the compact diagonal-block array of block size dblk (block b at
// G + b dblk^2, row-major with leading dimension dblk) and only the columns in
// the target's own block are updated.
template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_assemble_wake(HorseView hv, double off_, double cut_, TargetView tv,
                                                       TO *__restrict__ G, int64_t ld, int64_t col0,
                                                       float *__restrict__ G2, int dblk)
{
    __shared__ T sh[64][12];
    __shared__ int sc[64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < tv.M;
    const T X = ok ? T(tv.x[i]) : T(0), Y = ok ? T(tv.y[i]) : T(0), Z = ok ? T(tv.z[i]) : T(0);
    const T nx = ok ? T(tv.nx[i]) : T(0), ny = ok ? T(tv.ny[i]) : T(0), nz = ok ? T(tv.nz[i]) : T(0);
    const T off = T(off_), off2 = T(off_ * off_), cut2 = T(cut_ * cut_);
    for (int h0 = 0; h0 < hv.nh; h0 += 64) {
        const int cnt = min(64, hv.nh - h0);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt * 12; t += blockDim.x) {
            const int hh = t / 12, c = t % 12;
            sh[hh][c] = T(hv.hh[(int64_t)c * hv.nh + h0 + hh]);
        }
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) sc[t] = hv.col[h0 + t];
        __syncthreads();
        if (ok) {
            for (int hh = 0; hh < cnt; ++hh) {
                T ux = 0, uy = 0, uz = 0;
                TO *g;
                if (dblk > 0) {
                    const int64_t bb = i / dblk, c = sc[hh] - bb * dblk;
                    if (c < 0 || c >= dblk) continue;
                    g = G + bb * dblk * dblk + (i - bb * dblk) * dblk + c;
                } else {
                    g = G + i * ld + col0 + sc[hh];
                }
                horseshoe_u<T>(sh[hh], X, Y, Z, T(1), off, off2, cut2, ux, uy, uz);
                const TO v = TO(T(*g) + (ux * nx + uy * ny + uz * nz));
                *g = v;
                if (G2 && dblk == 0) G2[i * ld + col0 + sc[hh]] = float(v);
            }
        }
    }
}

int fp_launch_assemble_diag(fp_ctx *ctx, const fp_body *b, const TargetView &tv, int block, float *X)
{
    constexpr int TM = 64;
    if (b->element != FP_ELEM_VORTEXRING || tv.M != b->N || block <= 0)
        return fp_fail(ctx, FP_ERR_ARG, "internal: diagonal-block assembly of one vortex-ring body");
    const int64_t nb = (b->N + block - 1) / block;
    dim3 grid((unsigned)((block + 255) / 256), (unsigned)((block + TM - 1) / TM), (unsigned)nb);
    if (b->nvp == 4)
        hipLaunchKernelGGL((k_assemble_diag_vr<4, TM>), grid, dim3(256), 0, ctx->stream, panel_view(b), tv, block, X);
    else
        hipLaunchKernelGGL((k_assemble_diag_vr<8, TM>), grid, dim3(256), 0, ctx->stream, panel_view(b), tv, block, X);
    FP_HIP(ctx, hipGetLastError());
    if (b->nh > 0) {
        dim3 g2((unsigned)((tv.M + 255) / 256));
        hipLaunchKernelGGL((k_assemble_wake<double, float>), g2, dim3(256), 0, ctx->stream, horse_view(b),
                           b->kernel_offset, b->kernel_cutoff, tv, X, (int64_t)0, (int64_t)0, (float *)nullptr, block);
        FP_HIP(ctx, hipGetLastError());
    }
    return FP_OK;
}

// The entries are always computed in FP64 arithmetic and rounded once to the
// storage type T: an FP32 kernel loses about log10(R/h) digits in the far
// field (r_i/|r_i| - r_j/|r_j| and r_i x r_j cancel), which would limit the
// refinement of the mixed-precision solve.
template <typename T>
int fp_launch_assemble(fp_ctx *ctx, const fp_body *b, const TargetView &tv, T *G, int64_t ld, int64_t col0,
                       float *G32)
{
    constexpr int TM = 64;
    const PanelView pan = panel_view(b);
    if (tv.M == 0 || b->N == 0) return FP_OK;
    dim3 grid((unsigned)((b->N + 255) / 256), (unsigned)((tv.M + TM - 1) / TM));
    if (b->element == FP_ELEM_VORTEXRING) {
        int shared = 0;
        int st = fp_launch_assemble_edges<T>(ctx, b, tv, G, ld, col0, G32, &shared);
        if (st) return st;
        ProfScope ps(ctx, shared ? nullptr : sizeof(T) == 8 ? "assemble_vr_f64" : "assemble_vr_f32");
        if (shared) {
        } else if (b->nvp == 4)
            hipLaunchKernelGGL((k_assemble_vr<double, T, 4, TM>), grid, dim3(256), 0, ctx->stream, pan, tv, G, ld, col0,
                               G32);
        else
            hipLaunchKernelGGL((k_assemble_vr<double, T, 8, TM>), grid, dim3(256), 0, ctx->stream, pan, tv, G, ld, col0,
                               G32);
        FP_HIP(ctx, hipGetLastError());
        if (b->nh > 0) {
            ProfScope pw(ctx, "assemble_wake");
            dim3 g2((unsigned)((tv.M + 255) / 256));
            hipLaunchKernelGGL((k_assemble_wake<double, T>), g2, dim3(256), 0, ctx->stream, horse_view(b),
                               b->kernel_offset, b->kernel_cutoff, tv, G, ld, col0, G32, 0);
            FP_HIP(ctx, hipGetLastError());
        }
    } else {
        ProfScope ps(ctx, sizeof(T) == 8 ? "assemble_source_f64" : "assemble_source_f32");
        if (b->nvp == 4)
            hipLaunchKernelGGL((k_assemble_source<double, T, 4, TM>), grid, dim3(256), 0, ctx->stream, pan, tv, G, ld,
                               col0, G32);
        else
            hipLaunchKernelGGL((k_assemble_source<double, T, 8, TM>), grid, dim3(256), 0, ctx->stream, pan, tv, G, ld,
                               col0, G32);
        FP_HIP(ctx, hipGetLastError());
    }
    return FP_OK;
}
template int fp_launch_assemble<double>(fp_ctx *, const fp_body *, const TargetView &, double *, int64_t, int64_t,
                                        float *);

template <typename T>
int fp_launch_assemble_wake(fp_ctx *ctx, const fp_body *b, const TargetView &tv, T *G, int64_t ld, int64_t col0)
{
    if (b->element != FP_ELEM_VORTEXRING || b->nh == 0 || tv.M == 0) return FP_OK;
    ProfScope pw(ctx, "assemble_wake");
    dim3 g2((unsigned)((tv.M + 255) / 256));
    hipLaunchKernelGGL((k_assemble_wake<double, T>), g2, dim3(256), 0, ctx->stream, horse_view(b), b->kernel_offset,
                       b->kernel_cutoff, tv, G, ld, col0, (float *)nullptr, 0);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
template int fp_launch_assemble_wake<double>(fp_ctx *, const fp_body *, const TargetView &, double *, int64_t, int64_t);
template int fp_launch_assemble_wake<float>(fp_ctx *, const fp_body *, const TargetView &, float *, int64_t, int64_t);

int fp_assembly_queue_tiles(int64_t N, int64_t M, int tm, int *nbx, int64_t *ntiles)
{
    *nbx = (int)((N + 255) / 256);
    *ntiles = (int64_t)(*nbx) * ((M + tm - 1) / tm);
    return FP_OK;
}

template <int NVP, int TM, typename TO>
static void launch_queue(fp_ctx *ctx, const AssemblyItem *items, int nitems, int64_t total, int *counter, int blocks)
{
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, ctx->stream, counter);
    hipLaunchKernelGGL((k_assemble_vr_queue<double, TO, NVP, TM>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       items, nitems, total, counter);
}

int fp_launch_assemble_queue(fp_ctx *ctx, const AssemblyItem *d_items, int nitems, int64_t total, int nvp, int tm,
                             int prec, int *counter, int blocks)
{
    if (total <= 0) return FP_OK;
    ProfScope ps(ctx, prec == FP_F64 ? "assemble_vr_queue_f64" : "assemble_vr_queue_f32");
#define FP_Q(NV, TMV)                                                                             \
    if (nvp == NV && tm == TMV) {                                                                 \
        if (prec == FP_F64) launch_queue<NV, TMV, double>(ctx, d_items, nitems, total, counter, blocks); \
        else launch_queue<NV, TMV, float>(ctx, d_items, nitems, total, counter, blocks);          \
        FP_HIP(ctx, hipGetLastError());                                                           \
        return FP_OK;                                                                             \
    }
    FP_Q(4, 32) FP_Q(4, 64) FP_Q(4, 128) FP_Q(8, 32) FP_Q(8, 64) FP_Q(8, 128)
#undef FP_Q
    return fp_fail(ctx, FP_ERR_ARG, "internal: unsupported assembly tile");
}
template int fp_launch_assemble<float>(fp_ctx *, const fp_body *, const TargetView &, float *, int64_t, int64_t,
                                       float *);

// ----------------------------------------------------------------------------
// Field sums: one lane per target, panels in LDS.
//   MODE 0: velocity (3 components), MODE 1: n . velocity, MODE 2: potential.
// ELEM vortex ring: LDS holds vertices + edges (6*NVP) + strength; source or
// doublet potential: the Hess-Smith table (12 + 6*NVP) + nv + strength.
constexpr int FIELD_TILE = 64;

template <int NVP, int MODE, int ELEM>
struct FieldLds {
    static constexpr bool HS = !(ELEM == FP_ELEM_VORTEXRING && MODE != 2);
    static constexpr int STRIDE = HS ? (12 + 6 * NVP + 2) : (6 * NVP + 1);
};

template <int NVP, int MODE, int ELEM>
__global__ __launch_bounds__(256) void k_field(PanelView pan, const double *__restrict__ strength, TargetView tv,
                                               HorseView hv, int64_t pchunk, double *__restrict__ out,
                                               double *__restrict__ partial)
{
    using L = FieldLds<NVP, MODE, ELEM>;
    __shared__ double sp[FIELD_TILE][L::STRIDE];
    const int64_t N = pan.N;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < tv.M;
    const double X = ok ? tv.x[i] : 0.0, Y = ok ? tv.y[i] : 0.0, Z = ok ? tv.z[i] : 0.0;
    const double off = pan.off, off2 = pan.off * pan.off, cut2 = pan.cut * pan.cut;
    const int64_t p0 = (int64_t)blockIdx.y * pchunk;
    const int64_t p1 = min(N, p0 + pchunk);
    double ux = 0, uy = 0, uz = 0, ph = 0;
    for (int64_t t0 = p0; t0 < p1; t0 += FIELD_TILE) {
        const int cnt = (int)min((int64_t)FIELD_TILE, p1 - t0);
        __syncthreads();
        if (!L::HS) {
            for (int t = threadIdx.x; t < cnt * (3 * NVP); t += blockDim.x) {
                const int pp = t % cnt, c = t / cnt;  // c = 3k + d
                const int k = c / 3, d = c % 3, kk = (k + 1) % NVP;
                const double a = pan.pv[(int64_t)c * N + t0 + pp];
                const double bnext = pan.pv[(int64_t)(3 * kk + d) * N + t0 + pp];
                sp[pp][c] = a;
                sp[pp][3 * NVP + c] = bnext - a;
            }
            for (int t = threadIdx.x; t < cnt; t += blockDim.x) sp[t][6 * NVP] = strength[t0 + t];
        } else {
            constexpr int S = 12 + 6 * NVP;
            for (int t = threadIdx.x; t < cnt * S; t += blockDim.x) {
                const int pp = t % cnt, c = t / cnt;
                sp[pp][c] = pan.hs[(int64_t)c * N + t0 + pp];
            }
            for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
                sp[t][S] = (double)pan.nv[t0 + t];
                sp[t][S + 1] = strength[t0 + t];
            }
        }
        __syncthreads();
        if (ok) {
            for (int pp = 0; pp < cnt; ++pp) {
                const double *q = sp[pp];
                if (!L::HS) {
                    double vx[NVP], vy[NVP], vz[NVP], wx[NVP], wy[NVP], wz[NVP];
#pragma unroll
                    for (int k = 0; k < NVP; ++k) {
                        vx[k] = q[3 * k]; vy[k] = q[3 * k + 1]; vz[k] = q[3 * k + 2];
                        wx[k] = q[3 * NVP + 3 * k]; wy[k] = q[3 * NVP + 3 * k + 1]; wz[k] = q[3 * NVP + 3 * k + 2];
                    }
                    double ax = 0, ay = 0, az = 0;
                    ring_u<double, NVP>(vx, vy, vz, wx, wy, wz, X, Y, Z, off, off2, cut2, ax, ay, az);
                    const double g = q[6 * NVP];
                    ux += g * ax; uy += g * ay; uz += g * az;
                } else {
                    constexpr int S = 12 + 6 * NVP;
                    const double dx = X - q[9], dy = Y - q[10], dz = Z - q[11];
                    const double x = q[0] * dx + q[1] * dy + q[2] * dz;
                    const double y = q[3] * dx + q[4] * dy + q[5] * dz;
                    const double z = q[6] * dx + q[7] * dy + q[8] * dz;
                    const int nv = (int)q[S];
                    const double g = q[S + 1];
                    double V1, V2, V3, ps;
                    if (ELEM == FP_ELEM_SOURCE) {
                        hs_sums<double, NVP, true>(q + 12, q + 12 + NVP, q + 12 + 2 * NVP, q + 12 + 3 * NVP,
                                                   q + 12 + 4 * NVP, q + 12 + 5 * NVP, nv, x, y, z, off, V1, V2,
                                                   V3, ps);
                    } else {
                        hs_sums<double, NVP, false>(q + 12, q + 12 + NVP, q + 12 + 2 * NVP, q + 12 + 3 * NVP,
                                                    q + 12 + 4 * NVP, q + 12 + 5 * NVP, nv, x, y, z, off, V1,
                                                    V2, V3, ps);
                    }
                    const double k4 = 1.0 / (4 * FP_PI);
                    if (MODE == 2) {
                        ph += (ELEM == FP_ELEM_SOURCE) ? g * k4 * ps : g * V3 * k4;
                    } else {
                        V1 *= k4; V2 *= k4; V3 *= k4;
                        ux += g * (V1 * q[0] + V2 * q[3] + V3 * q[6]);
                        uy += g * (V1 * q[1] + V2 * q[4] + V3 * q[7]);
                        uz += g * (V1 * q[2] + V2 * q[5] + V3 * q[8]);
                    }
                }
            }
        }
    }
    // wake horseshoes (split 0 only), strengths of their panels
    if (ELEM == FP_ELEM_VORTEXRING && blockIdx.y == 0 && hv.nh > 0) {
        __shared__ double sh[32][13];
        for (int h0 = 0; h0 < hv.nh; h0 += 32) {
            const int cnt = min(32, hv.nh - h0);
            __syncthreads();
            for (int t = threadIdx.x; t < cnt * 12; t += blockDim.x) {
                const int hh = t / 12, c = t % 12;
                sh[hh][c] = hv.hh[(int64_t)c * hv.nh + h0 + hh];
            }
            for (int t = threadIdx.x; t < cnt; t += blockDim.x) sh[t][12] = strength[hv.col[h0 + t]];
            __syncthreads();
            if (ok) {
                for (int hh = 0; hh < cnt; ++hh) {
                    if (MODE == 2) {
                        ph += sh[hh][12] * semiinf_doublet_phi<double>(sh[hh], X, Y, Z);
                    } else {
                        horseshoe_u<double>(sh[hh], X, Y, Z, sh[hh][12], off, off2, cut2, ux, uy, uz);
                    }
                }
            }
        }
    }
    if (!ok) return;
    const int64_t M = tv.M;
    if (partial) {
        double *pp = partial + (int64_t)blockIdx.y * (MODE == 0 ? 3 : 1) * M;
        if (MODE == 0) { pp[i] = ux; pp[M + i] = uy; pp[2 * M + i] = uz; }
        else if (MODE == 1) pp[i] = ux * tv.nx[i] + uy * tv.ny[i] + uz * tv.nz[i];
        else pp[i] = ph;
    } else {
        if (MODE == 0) { out[3 * i] += ux; out[3 * i + 1] += uy; out[3 * i + 2] += uz; }
        else if (MODE == 1) out[i] += ux * tv.nx[i] + uy * tv.ny[i] + uz * tv.nz[i];
        else out[i] += ph;
    }
}

__global__ void k_field_reduce(const double *__restrict__ partial, int splits, int64_t M, int comps,
                               double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    for (int c = 0; c < comps; ++c) {
        double s = 0;
        for (int k = 0; k < splits; ++k) s += partial[((int64_t)k * comps + c) * M + i];
        if (comps == 3) out[3 * i + c] += s;
        else out[i] += s;
    }
}

static int field_splits(int64_t M, int64_t N)
{
    const int64_t bx = (M + 255) / 256;
    int64_t s = (2048 + bx - 1) / bx;
    s = std::min<int64_t>(s, (N + FIELD_TILE - 1) / FIELD_TILE);
    return (int)std::max<int64_t>(1, std::min<int64_t>(s, 256));
}

size_t fp_field_partial_bytes(int64_t M, int64_t N)
{
    const int s = field_splits(M, N);
    return s > 1 ? (size_t)s * 3 * M * sizeof(double) : 0;
}

template <int NVP, int MODE, int ELEM>
static void launch_field_t(fp_ctx *ctx, const fp_body *b, const double *strength, const TargetView &tv,
                           double *out, double *partial, int splits, int64_t pchunk)
{
    dim3 grid((unsigned)((tv.M + 255) / 256), (unsigned)splits);
    hipLaunchKernelGGL((k_field<NVP, MODE, ELEM>), grid, dim3(256), 0, ctx->stream, panel_view(b), strength, tv,
                       horse_view(b), pchunk, out, partial);
}

template <int NVP>
static void launch_field_nvp(fp_ctx *ctx, const fp_body *b, const double *strength, const TargetView &tv, int mode,
                             double *out, double *partial, int splits, int64_t pchunk)
{
    const bool vr = b->element == FP_ELEM_VORTEXRING;
    if (mode == 0) {
        if (vr) launch_field_t<NVP, 0, FP_ELEM_VORTEXRING>(ctx, b, strength, tv, out, partial, splits, pchunk);
        else launch_field_t<NVP, 0, FP_ELEM_SOURCE>(ctx, b, strength, tv, out, partial, splits, pchunk);
    } else if (mode == 1) {
        if (vr) launch_field_t<NVP, 1, FP_ELEM_VORTEXRING>(ctx, b, strength, tv, out, partial, splits, pchunk);
        else launch_field_t<NVP, 1, FP_ELEM_SOURCE>(ctx, b, strength, tv, out, partial, splits, pchunk);
    } else {
        if (vr) launch_field_t<NVP, 2, FP_ELEM_VORTEXRING>(ctx, b, strength, tv, out, partial, splits, pchunk);
        else launch_field_t<NVP, 2, FP_ELEM_SOURCE>(ctx, b, strength, tv, out, partial, splits, pchunk);
    }
}

int fp_launch_field(fp_ctx *ctx, const fp_body *b, const double *strength, const TargetView &tv, int mode,
                    double *out, double *partial, size_t partial_bytes)
{
    if (tv.M == 0 || b->N == 0) return FP_OK;
    int splits = field_splits(tv.M, b->N);
    const int comps = mode == 0 ? 3 : 1;
    if (splits > 1 && (partial == nullptr || partial_bytes < (size_t)splits * comps * tv.M * sizeof(double)))
        splits = 1;
    int64_t pchunk = (b->N + splits - 1) / splits;
    pchunk = (pchunk + FIELD_TILE - 1) / FIELD_TILE * FIELD_TILE;
    splits = (int)((b->N + pchunk - 1) / pchunk);
    double *part = splits > 1 ? partial : nullptr;
    if (b->element == FP_ELEM_VORTEXRING) {
        int used = 0;
        int st = fp_launch_field_trails(ctx, b, strength, tv, mode, out, part, splits, &used);
        if (st || used) return st;
    }
    ProfScope ps(ctx, mode == 1 ? "field_matvec_f64" : (mode == 0 ? "field_velocity_f64" : "field_potential_f64"));
    if (b->nvp == 4) launch_field_nvp<4>(ctx, b, strength, tv, mode, out, part, splits, pchunk);
    else launch_field_nvp<8>(ctx, b, strength, tv, mode, out, part, splits, pchunk);
    FP_HIP(ctx, hipGetLastError());
    if (part) {
        hipLaunchKernelGGL(k_field_reduce, dim3((unsigned)((tv.M + 255) / 256)), dim3(256), 0, ctx->stream, part,
                           splits, tv.M, comps, out);
        FP_HIP(ctx, hipGetLastError());
    }
    return FP_OK;
}

// ----------------------------------------------------------------------------
// RHS of the no-flow-through condition (src/FLOWPanel_linearsolver.jl:297-328)
__global__ void k_rhs(const double *__restrict__ uinf, const double *__restrict__ geo, int64_t N,
                      double *__restrict__ rhs)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    rhs[j] = -(uinf[j] * geo[j] + uinf[N + j] * geo[N + j] + uinf[2 * N + j] * geo[2 * N + j]);
}

int fp_rhs_from(fp_ctx *ctx, const fp_body *b, const double *uinf_soa, double *rhs)
{
    if (b->N == 0) return FP_OK;
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((b->N + 255) / 256)), dim3(256), 0, ctx->stream, uinf_soa, b->d_geo,
                       b->N, rhs);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

int fp_rhs_kernel(fp_ctx *ctx, const fp_body *b, double *rhs)
{
    if (b->N == 0) return FP_OK;
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)((b->N + 255) / 256)), dim3(256), 0, ctx->stream, b->d_uinf, b->d_geo,
                       b->N, rhs);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
