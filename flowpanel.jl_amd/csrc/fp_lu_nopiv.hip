// fp_lu_nopiv.hip -- FP32 LU of a randomized matrix without pivoting, for the
// mixed-precision refinement path (FP_SOLVER_LU32_IR).
//
// The refinement needs only an FP32 approximate inverse; the FP64 residual
// r = b - A x decides the result.  Partial pivoting is what makes the blocked
// LU latency-bound at the sizes of a wing (4k-12k unknowns): every 32-column
// sub-panel runs a tournament of pivot searches.  A random butterfly
// transformation (RBT, Parker 1995; Baboulin, Dongarra, Herrmann, Tomov 2013)
// A_r = U^T A V with two depth-2 recursive butterflies U, V makes pivoting
// unnecessary with probability close to one, and the LU of A_r is then a
// plain blocked right-looking factorisation:
//
//   per 64-column step:  k_nopiv_diag   the 64 x 64 diagonal block factored in
//                                       registers (one row per lane, row k
//                                       broadcast by v_readlane), then the
//                                       explicit inverses of L11 and U11 (two
//                                       waves, one each);
//                        k_nopiv_panel  L21 = A21 U11^-1 and U12 = L11^-1 A12 as
//                                       64 x 64 x 64 products on MFMA, in place
//                                       (U12 within the 256-column outer block);
//                        fp_gemm_sub    the update of the outer block's columns;
//   per 256-column outer block: the U rows right of it (blocked forward
//   substitution: MFMA products and k_nopiv_panel), then the rank-256 MFMA
//   trailing update.
//
// A x = b becomes  A_r y = U^T b,  x = V y.  A depth-2 butterfly couples only
// the four indices {p, p + m, p + 2m, p + 3m} (m = n4 / 4, n4 = n rounded up
// to a multiple of 4, the pad being the identity), so U^T A V is one pass of
// 4 x 4 . 4 x 4 . 4 x 4 products over the (p, q) orbit blocks, and the vector
// transforms are 4 x 4 products per orbit.  A zero or non-finite pivot is
// reported through `info` and the caller falls back to the pivoted LU.
#include "fp_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace {

constexpr int NPB = 64;   // block of the no-pivot LU

typedef float nf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Compile-time loop: f(cint<I>{}) for I = B, B + S, ... while I < E (S > 0)
// or I > E (S < 0), so every register-array index below is a constant.
template <int V> struct cint {
    static constexpr int v = V;
};
template <int B, int E, int S, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (S > 0 ? B < E : B > E) {
        f(cint<B>{});
        static_for<B + S, E, S>(f);
    }
}

// v[q] -= f * (v[q] of lane j) for q in [LO, HI): the broadcasts go out in
// groups of four before their FMAs, so that no v_readlane's result is
// consumed by the very next instruction (a readlane followed at once by its
// use costs two wait states, which doubled the length of these loops).
template <int LO, int HI, int G = 4>
__device__ __forceinline__ void bcast_axpy(float (&v)[NPB], float f, int j)
{
#pragma unroll
    for (int q0 = LO; q0 < HI; q0 += G) {
        float t[G];
#pragma unroll
        for (int u = 0; u < G; ++u)
            if (q0 + u < HI) t[u] = bcast(v[q0 + u], j);
#pragma unroll
        for (int u = 0; u < G; ++u)
            if (q0 + u < HI) v[q0 + u] = fmaf(-f, t[u], v[q0 + u]);
    }
}

// Columns [Q0, Q1) of X = U11^-1 for row `lane`, from row `lane` of U in u[]:
// rows bottom-up, Y_i -= (u_ic / u_cc) Y_c for i < c, then X_i = Y_i / u_ii.
// Column q takes the steps c <= q, so a column range is independent of the rest.
template <int Q0, int Q1>
__device__ __forceinline__ void nopiv_uinv_cols(const float (&u)[NPB], const float (*lu)[NPB + 1], int lane, int bs,
                                                int64_t k0, int64_t ld, float *__restrict__ Uinv,
                                                float *__restrict__ YU)
{
    float x[NPB];
#pragma unroll
    for (int q = Q0; q < Q1; ++q) x[q] = q == lane ? 1.f : 0.f;
    float dii = 1.f;
    static_for<NPB - 1, -1, -1>([&](auto C) {
        constexpr int c = decltype(C)::v;
        dii = lane == c ? u[c] : dii;
        if constexpr (c < Q1) {   // else no column of the range at or right of c
            const float coef = lane < c ? u[c] * (1.f / lu[c][c]) : 0.f;
            bcast_axpy<(c > Q0 ? c : Q0), Q1>(x, coef, c);
        }
    });
    const float rd = 1.f / dii;
#pragma unroll
    for (int c = Q0; c < Q1; c += 4)
        *reinterpret_cast<float4 *>(Uinv + lane * NPB + c) =
            make_float4(x[c] * rd, x[c + 1] * rd, x[c + 2] * rd, x[c + 3] * rd);
    if (YU && lane < bs) {   // -U11^-1 into the diagonal block of the full inverse (ld as A)
#pragma unroll
        for (int c = Q0; c < Q1; ++c)
            if (c < bs) YU[(k0 + lane) * ld + k0 + c] = -(x[c] * rd);
    }
}

// Columns [Q0, Q1) of X = L11^-1 (unit lower) for row `lane`, the row's
// multipliers l_lane,c in l[] (zero at and right of the diagonal): rows
// top-down, X_i -= l_ic X_c for i > c; column q takes the steps c >= q.
template <int Q0, int Q1>
__device__ __forceinline__ void nopiv_linv_cols(const float (&l)[NPB], int lane, int bs, int64_t k0, int64_t ld,
                                                float *__restrict__ Linv, float *__restrict__ YL)
{
    float x[NPB];
#pragma unroll
    for (int q = Q0; q < Q1; ++q) x[q] = q == lane ? 1.f : 0.f;
    static_for<Q0, NPB, 1>([&](auto C) {
        constexpr int c = decltype(C)::v;
        bcast_axpy<Q0, (c + 1 < Q1 ? c + 1 : Q1)>(x, l[c], c);
    });
#pragma unroll
    for (int c = Q0; c < Q1; c += 4)
        *reinterpret_cast<float4 *>(Linv + lane * NPB + c) = make_float4(x[c], x[c + 1], x[c + 2], x[c + 3]);
    if (YL && lane < bs) {   // -L11^-1 likewise
#pragma unroll
        for (int c = Q0; c < Q1; ++c)
            if (c < bs) YL[(k0 + lane) * ld + k0 + c] = -x[c];
    }
}

// Diagonal block at (k0, k0): LU without pivoting in place (rows beyond n are
// the identity), then Linv = L11^-1 (unit lower) and Uinv = U11^-1, 64 x 64
// row-major; 512 threads.  Wave 0 holds row `lane` of the block in registers
// and factors it; the broadcast of row j is one v_readlane per entry (j is a
// compile-time index, every loop is unrolled), and every row takes the update,
// rows at and above the pivot with a zero multiplier, so no update hides
// behind a branch (the branchy form cost 37 us per block, a broadcast through
// LDS 46 us).  Then the eight waves invert: waves 0-3 four column ranges of
// U^-1, waves 4-7 four of L^-1, the ranges cut so that each wave takes about
// an eighth of the 2 x 2,080 broadcast-and-FMA steps (U column q costs q + 1,
// L column q costs 64 - q).
__device__ __forceinline__ void nopiv_diag_dev(float *__restrict__ A, int64_t ld, int64_t n, int64_t k0,
                                               float *__restrict__ Linv, float *__restrict__ Uinv,
                                               int *__restrict__ info, float *__restrict__ YL, float *__restrict__ YU)
{
    __shared__ float lu[NPB][NPB + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int bs = (int)min((int64_t)NPB, n - k0);
    float a[NPB];
    if (wave == 0) {
        float *row = A + (k0 + lane) * ld + k0;
        if (bs == NPB) {
#pragma unroll
            for (int c = 0; c < NPB; c += 4) {
                const float4 v = *reinterpret_cast<const float4 *>(row + c);
                a[c] = v.x; a[c + 1] = v.y; a[c + 2] = v.z; a[c + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int c = 0; c < NPB; ++c) a[c] = (lane < bs && c < bs) ? row[c] : (lane == c ? 1.f : 0.f);
        }
        bool bad = false;
        static_for<0, NPB, 1>([&](auto J) {
            constexpr int j = decltype(J)::v;
            const float piv = bcast(a[j], j);
            bad = bad || !(fabsf(piv) > 0.f) || !(fabsf(piv) <= 3.4e38f);
            // rows at and above j take a zero multiplier (unconditional updates, no branches)
            const float l = lane > j ? a[j] / piv : 0.f;
            a[j] = lane > j ? l : a[j];
            bcast_axpy<j + 1, NPB>(a, l, j);
        });
        if (bad && lane == 0) atomicCAS(info, 0, (int)(k0 + 1));
        if (lane < bs) {
            if (bs == NPB) {
#pragma unroll
                for (int c = 0; c < NPB; c += 4)
                    *reinterpret_cast<float4 *>(row + c) = make_float4(a[c], a[c + 1], a[c + 2], a[c + 3]);
            } else {
#pragma unroll
                for (int c = 0; c < NPB; ++c)
                    if (c < bs) row[c] = a[c];
            }
        }
#pragma unroll
        for (int c = 0; c < NPB; ++c) lu[lane][c] = a[c];
    }
    __syncthreads();
    if (wave == 0) {
        nopiv_uinv_cols<0, 32>(a, lu, lane, bs, k0, ld, Uinv, YU);
        return;
    }
    if (wave < 4) {   // U part of row `lane` (the strictly lower entries are never read)
#pragma unroll
        for (int c = 0; c < NPB; ++c) a[c] = lu[lane][c];
        if (wave == 1) nopiv_uinv_cols<32, 44>(a, lu, lane, bs, k0, ld, Uinv, YU);
        else if (wave == 2) nopiv_uinv_cols<44, 56>(a, lu, lane, bs, k0, ld, Uinv, YU);
        else nopiv_uinv_cols<56, 64>(a, lu, lane, bs, k0, ld, Uinv, YU);
        return;
    }
#pragma unroll
    for (int c = 0; c < NPB; ++c) a[c] = c < lane ? lu[lane][c] : 0.f;
    if (wave == 4) nopiv_linv_cols<0, 8>(a, lane, bs, k0, ld, Linv, YL);
    else if (wave == 5) nopiv_linv_cols<8, 20>(a, lane, bs, k0, ld, Linv, YL);
    else if (wave == 6) nopiv_linv_cols<20, 32>(a, lane, bs, k0, ld, Linv, YL);
    else nopiv_linv_cols<32, 64>(a, lane, bs, k0, ld, Linv, YL);
}

__global__ __launch_bounds__(512) void k_nopiv_diag(float *__restrict__ A, int64_t ld, int64_t n, int64_t k0,
                                                   float *__restrict__ Linv, float *__restrict__ Uinv,
                                                   int *__restrict__ info, float *__restrict__ YL,
                                                   float *__restrict__ YU)
{
    nopiv_diag_dev(A, ld, n, k0, Linv, Uinv, info, YL, YU);
}

__global__ void k_fill_info(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// One 64 x 64 tile of the block row / column of step k0, in place:
// blocks [0, nl): L21 tile rows off = k0 + 64 (1 + b) .. off+63 := A21 U11^-1;
// blocks [nl, nl + nu): U12 tile columns off = u_off + 64 (b - nl) .. off+63 := L11^-1 A12.
// Both operands in LDS; wave w computes output rows 16w..16w+15 with
// v_mfma_f32_16x16x4 (lane l supplies P(l%16, k + l/16) and Q(k + l/16, l%16);
// accumulator element r holds C(4 (l/16) + r, l%16)).
__device__ __forceinline__ void nopiv_panel_dev(float *__restrict__ A, int64_t ld, int64_t n, int64_t k0,
                                                const float *__restrict__ Linv, const float *__restrict__ Uinv, int nl,
                                                int64_t u_off, int b)
{
    __shared__ float P[NPB][NPB + 4];
    __shared__ float Q[NPB][NPB + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool lpanel = b < nl;
    const int64_t off = lpanel ? k0 + NPB + (int64_t)NPB * b : u_off + (int64_t)NPB * (b - nl);
    float *tile = lpanel ? A + off * ld + k0 : A + k0 * ld + off;
    const int tr = (int)min((int64_t)NPB, lpanel ? n - off : n - k0);
    const int tc = (int)min((int64_t)NPB, lpanel ? n - k0 : n - off);
    // the tile is the left operand of L21 (times U11^-1), the right one of U12 (L11^-1 times it)
    const float *inv = lpanel ? Uinv : Linv;
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    const bool full = tr == NPB && tc == NPB;
#pragma unroll
    for (int e = 0; e < 16; e += 4) {
        float4 v;
        if (full) {
            v = *reinterpret_cast<const float4 *>(tile + r * ld + c0 + e);
        } else {
            v.x = (r < tr && c0 + e < tc) ? tile[r * ld + c0 + e] : 0.f;
            v.y = (r < tr && c0 + e + 1 < tc) ? tile[r * ld + c0 + e + 1] : 0.f;
            v.z = (r < tr && c0 + e + 2 < tc) ? tile[r * ld + c0 + e + 2] : 0.f;
            v.w = (r < tr && c0 + e + 3 < tc) ? tile[r * ld + c0 + e + 3] : 0.f;
        }
        const float4 w = *reinterpret_cast<const float4 *>(inv + r * NPB + c0 + e);
        if (lpanel) {
            P[r][c0 + e] = v.x; P[r][c0 + e + 1] = v.y; P[r][c0 + e + 2] = v.z; P[r][c0 + e + 3] = v.w;
            Q[r][c0 + e] = w.x; Q[r][c0 + e + 1] = w.y; Q[r][c0 + e + 2] = w.z; Q[r][c0 + e + 3] = w.w;
        } else {
            Q[r][c0 + e] = v.x; Q[r][c0 + e + 1] = v.y; Q[r][c0 + e + 2] = v.z; Q[r][c0 + e + 3] = v.w;
            P[r][c0 + e] = w.x; P[r][c0 + e + 1] = w.y; P[r][c0 + e + 2] = w.z; P[r][c0 + e + 3] = w.w;
        }
    }
    __syncthreads();
    nf4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = nf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NPB; ks += 4) {
        const float pa = P[16 * wave + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa, Q[ks + (lane >> 4)][16 * q + (lane & 15)], acc[q], 0, 0,
                                                          0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int col = 16 * q + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = 16 * wave + 4 * (lane >> 4) + e;
            if (row < tr && col < tc) tile[row * ld + col] = acc[q][e];
        }
    }
}

// C -= A B on 64 x 64 output tiles (FP32, row-major), for the small products of
// the outer blocks (the in-block trailing updates, the U rows' substitution),
// where the 128 x 128 tiles of fp_gemm_sub leave most of the chip idle: blockIdx.x
// = tile (row-tile major), K in chunks of 64 through LDS, v_mfma_f32_16x16x4 as in
// k_nopiv_panel (wave w: output rows 16w..16w+15).  Loads of partial tiles use
// clamped addresses and a 0/1 factor (every load in bounds and unconditional).
__device__ __forceinline__ void gemm64_dev(int64_t m, int64_t n, int64_t kk, const float *__restrict__ A, int64_t lda,
                                           const float *__restrict__ B, int64_t ldb, float *__restrict__ C, int64_t ldc,
                                           int64_t tile)
{
    __shared__ float P[NPB][NPB + 4];
    __shared__ float Q[NPB][NPB + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t tn_count = (n + NPB - 1) / NPB;
    const int64_t i0 = (tile / tn_count) * NPB, j0 = (tile % tn_count) * NPB;
    const int tr = (int)min((int64_t)NPB, m - i0), tc = (int)min((int64_t)NPB, n - j0);
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    nf4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = nf4{0.f, 0.f, 0.f, 0.f};
    for (int64_t k0 = 0; k0 < kk; k0 += NPB) {
        const int kc = (int)min((int64_t)NPB, kk - k0);
        const float *ar = A + (i0 + min(r, tr - 1)) * lda + k0;
        const float *br = B + (k0 + min(r, kc - 1)) * ldb + j0;
        const float am = r < tr ? 1.f : 0.f, bm = r < kc ? 1.f : 0.f;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int c = c0 + e;
            P[r][c] = ar[min(c, kc - 1)] * (c < kc ? am : 0.f);
            Q[r][c] = br[min(c, tc - 1)] * (c < tc ? bm : 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < NPB; ks += 4) {
            const float pa = P[16 * wave + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa, Q[ks + (lane >> 4)][16 * q + (lane & 15)], acc[q], 0,
                                                              0, 0);
        }
    }
    // every load of C is issued before the first store (clamped addresses), then the stores
    float cv[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            cv[q][e] = C[(i0 + min(16 * wave + 4 * (lane >> 4) + e, tr - 1)) * ldc + j0 + min(16 * q + (lane & 15), tc - 1)];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int col = 16 * q + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = 16 * wave + 4 * (lane >> 4) + e;
            if (row < tr && col < tc) C[(i0 + row) * ldc + j0 + col] = cv[q][e] - acc[q][e];
        }
    }
}

__global__ __launch_bounds__(256) void k_gemm64_sub(int64_t m, int64_t n, int64_t kk, const float *__restrict__ A,
                                                   int64_t lda, const float *__restrict__ B, int64_t ldb,
                                                   float *__restrict__ C, int64_t ldc)
{
    gemm64_dev(m, n, kk, A, lda, B, ldb, C, ldc, blockIdx.x);
}

static int gemm64_sub(fp_ctx *ctx, int64_t m, int64_t n, int64_t k, const float *A, int64_t lda, const float *B,
                      int64_t ldb, float *C, int64_t ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return FP_OK;
    const int64_t tiles = ((m + NPB - 1) / NPB) * ((n + NPB - 1) / NPB);
    ProfScope ps(ctx, "gemm_f32");
    hipLaunchKernelGGL(k_gemm64_sub, dim3((unsigned)tiles), dim3(256), 0, ctx->stream, m, n, k, A, lda, B, ldb, C, ldc);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

__global__ __launch_bounds__(256) void k_nopiv_panel(float *__restrict__ A, int64_t ld, int64_t n, int64_t k0,
                                                    const float *__restrict__ Linv, const float *__restrict__ Uinv,
                                                    int nl, int64_t u_off)
{
    nopiv_panel_dev(A, ld, n, k0, Linv, Uinv, nl, u_off, blockIdx.x);
}

// U^T A V over the orbit blocks: thread (p, q) reads the 4 x 4 block of A at rows
// p + i m, columns q + j m (the identity beyond n), forms U_p^T B V_q in FP64 and
// stores it rounded to FP32.  W holds the 4 x 4 butterfly of every orbit,
// row-major: U's at W[16 p], V's at W[16 (m + q)].
template <typename Tin>
__global__ __launch_bounds__(256) void k_rbt_matrix(const Tin *__restrict__ A, int64_t n, int64_t lda,
                                                   float *__restrict__ Ar, int64_t m, int64_t ldr,
                                                   const double *__restrict__ W)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t p = blockIdx.y;
    if (q >= m) return;
    double B[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t gr = p + i * m;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t gc = q + j * m;
            B[i][j] = (gr < n && gc < n) ? (double)A[gr * lda + gc] : (gr == gc ? 1.0 : 0.0);
        }
    }
    const double *U = W + 16 * p, *V = W + 16 * (m + q);
    double T[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) T[i][j] = B[i][0] * V[j] + B[i][1] * V[4 + j] + B[i][2] * V[8 + j] + B[i][3] * V[12 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float *dst = Ar + (p + i * m) * ldr + q;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            dst[j * m] = (float)(U[i] * T[0][j] + U[4 + i] * T[1][j] + U[8 + i] * T[2][j] + U[12 + i] * T[3][j]);
    }
}

// t (n4 x k) = U^T b (n x k, zero-padded), or x (n x k) = V y (y n4 x k, the pad dropped)
__global__ void k_rbt_vec(const double *__restrict__ in, int64_t ldin, double *__restrict__ out, int64_t ldout,
                          int64_t n, int64_t m, int64_t k, const double *__restrict__ W, int transpose_u)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t col = blockIdx.y;
    if (p >= m || col >= k) return;
    const double *Wp = W + 16 * (transpose_u ? p : m + p);
    const double *src = in + col * ldin;
    double *dst = out + col * ldout;
    double v[4], o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t g = p + i * m;
        v[i] = (transpose_u && g >= n) ? 0.0 : src[g];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        o[i] = transpose_u ? Wp[i] * v[0] + Wp[4 + i] * v[1] + Wp[8 + i] * v[2] + Wp[12 + i] * v[3]
                           : Wp[4 * i] * v[0] + Wp[4 * i + 1] * v[1] + Wp[4 * i + 2] * v[2] + Wp[4 * i + 3] * v[3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t g = p + i * m;
        if (transpose_u || g < n) dst[g] = o[i];
    }
}

// splitmix64: deterministic random weights
uint64_t mix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One depth-2 recursive butterfly of size 4m as its m orbit blocks (row-major
// 4 x 4, acting on (x_p, x_p+m, x_p+2m, x_p+3m)): the outer butterfly pairs
// p with p + 2m, the two inner ones pair p with p + m in each half; every
// butterfly is (1/sqrt 2) [R S; R -S] with random diagonals exp(r / 10),
// r uniform in [-1/2, 1/2].
void butterfly_blocks(int64_t m, uint64_t seed, double *out)
{
    uint64_t s = seed;
    auto rnd = [&]() { return std::exp((((double)(mix64(s) >> 11) * 0x1.0p-53) - 0.5) / 10.0); };
    const double h = std::sqrt(0.5);
    for (int64_t p = 0; p < m; ++p) {
        const double R0 = rnd(), S0 = rnd(), R1 = rnd(), S1 = rnd();   // outer: positions p, p + m
        const double Ra = rnd(), Sa = rnd(), Rb = rnd(), Sb = rnd();   // inner: first half, second half
        // z = outer x: z_a = h (R0 x_a + S0 x_c), z_c = h (R0 x_a - S0 x_c),
        //              z_b = h (R1 x_b + S1 x_d), z_d = h (R1 x_b - S1 x_d)
        const double Z[4][4] = {{h * R0, 0, h * S0, 0}, {0, h * R1, 0, h * S1},
                                {h * R0, 0, -h * S0, 0}, {0, h * R1, 0, -h * S1}};
        // y = inner z: y_a = h (Ra z_a + Sa z_b), y_b = h (Ra z_a - Sa z_b),
        //              y_c = h (Rb z_c + Sb z_d), y_d = h (Rb z_c - Sb z_d)
        const double Y[4][4] = {{h * Ra, h * Sa, 0, 0}, {h * Ra, -h * Sa, 0, 0},
                                {0, 0, h * Rb, h * Sb}, {0, 0, h * Rb, -h * Sb}};
        double *o = out + 16 * p;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double acc = 0;
                for (int k = 0; k < 4; ++k) acc += Y[i][k] * Z[k][j];
                o[4 * i + j] = acc;
            }
    }
}

// The butterflies' orbit blocks on the device (the same construction as
// butterfly_blocks; the i-th draw of a splitmix64 stream is mix64 of seed +
// (i + 1) * golden, so every orbit computes its own eight draws).
__device__ __forceinline__ uint64_t mix64_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_rbt_weights(int64_t m, uint64_t seed, double *__restrict__ W)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= 2 * m) return;
    const bool v = p >= m;   // U's orbit blocks first, then V's
    const int64_t q = v ? p - m : p;
    const uint64_t sd = v ? seed ^ 0x5DEECE66Dull : seed;
    double r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = exp((((double)(mix64_at(sd, 8 * (uint64_t)q + j) >> 11) * 0x1.0p-53) - 0.5) / 10.0);
    const double h = sqrt(0.5);
    const double Z[4][4] = {{h * r[0], 0, h * r[1], 0}, {0, h * r[2], 0, h * r[3]},
                            {h * r[0], 0, -h * r[1], 0}, {0, h * r[2], 0, -h * r[3]}};
    const double Y[4][4] = {{h * r[4], h * r[5], 0, 0}, {h * r[4], -h * r[5], 0, 0},
                            {0, 0, h * r[6], h * r[7]}, {0, 0, h * r[6], -h * r[7]}};
    double *o = W + 16 * p;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) o[4 * i + j] = Y[i][0] * Z[0][j] + Y[i][1] * Z[1][j] + Y[i][2] * Z[2][j] + Y[i][3] * Z[3][j];
}

// E (n x n, FP64, lde) = -(V C U^T) restricted to n x n, C (4m x 4m, FP32, ldc):
// thread (p, q) forms -V_p B U_q^T of the orbit block B of C at rows p + i m,
// columns q + j m (the undoing of U^T A V for an inverse: A^-1 = V A_r^-1 U^T).
__global__ __launch_bounds__(256) void k_rbt_unmatrix(const float *__restrict__ C, int64_t ldc, int64_t m,
                                                     double *__restrict__ E, int64_t n, int64_t lde,
                                                     const double *__restrict__ W)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t p = blockIdx.y;
    if (q >= m) return;
    double B[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) B[i][j] = (double)C[(p + i * m) * ldc + q + j * m];
    const double *Vp = W + 16 * (m + p), *Uq = W + 16 * q;
    double T[4][4];   // B U_q^T
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            T[i][j] = B[i][0] * Uq[4 * j] + B[i][1] * Uq[4 * j + 1] + B[i][2] * Uq[4 * j + 2] + B[i][3] * Uq[4 * j + 3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t gr = p + i * m;
        if (gr >= n) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t gc = q + j * m;
            if (gc < n)
                E[gr * lde + gc] =
                    -(Vp[4 * i] * T[0][j] + Vp[4 * i + 1] * T[1][j] + Vp[4 * i + 2] * T[2][j] + Vp[4 * i + 3] * T[3][j]);
        }
    }
}

}  // namespace

int64_t fp_rbt_size(int64_t n) { return (n + 3) / 4 * 4; }

size_t fp_rbt_weights_bytes(int64_t n) { return (size_t)32 * (fp_rbt_size(n) / 4) * sizeof(double); }

// U's orbit blocks then V's, for a matrix of order n (padded to n4)
int fp_rbt_weights(fp_ctx *ctx, int64_t n, uint64_t seed, double *d_W)
{
    const int64_t m = fp_rbt_size(n) / 4;
    std::vector<double> W((size_t)32 * m);
    butterfly_blocks(m, seed, W.data());
    butterfly_blocks(m, seed ^ 0x5DEECE66Dull, W.data() + 16 * m);
    FP_HIP(ctx, hipMemcpyAsync(d_W, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

template <typename Tin>
int fp_rbt_matrix(fp_ctx *ctx, const Tin *A, int64_t n, int64_t lda, float *Ar, int64_t ldr, const double *d_W)
{
    const int64_t m = fp_rbt_size(n) / 4;
    if (m == 0) return FP_OK;
    ProfScope ps(ctx, "rbt");
    hipLaunchKernelGGL(k_rbt_matrix<Tin>, dim3((unsigned)((m + 255) / 256), (unsigned)m), dim3(256), 0, ctx->stream,
                       A, n, lda, Ar, m, ldr, d_W);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
template int fp_rbt_matrix<double>(fp_ctx *, const double *, int64_t, int64_t, float *, int64_t, const double *);
template int fp_rbt_matrix<float>(fp_ctx *, const float *, int64_t, int64_t, float *, int64_t, const double *);

// transpose_u: out (n4 x k, ldout) = U^T in (n x k, ldin); else out (n x k) = V in (n4 x k)
int fp_rbt_vec(fp_ctx *ctx, const double *in, int64_t ldin, double *out, int64_t ldout, int64_t n, int64_t k,
               const double *d_W, int transpose_u)
{
    const int64_t m = fp_rbt_size(n) / 4;
    if (m == 0 || k == 0) return FP_OK;
    hipLaunchKernelGGL(k_rbt_vec, dim3((unsigned)((m + 255) / 256), (unsigned)k), dim3(256), 0, ctx->stream, in, ldin,
                       out, ldout, n, m, k, d_W, transpose_u);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// Inner blocks of 64 columns per outer block of 256: the inner steps factor the
// outer block's columns only (their trailing updates stop at its last column),
// then the outer block's U rows right of it come from one blocked forward
// substitution with the four inner L11^-1, and the trailing matrix takes one
// rank-256 update: a rank-64 update per 64 columns reads and writes the whole
// trailing matrix four times as often (at 4,800 unknowns it bound the
// factorisation: 41 us per step against 32 us for the diagonal block).
constexpr int NOB = 4 * NPB, NOB_MAX = 8 * NPB;

// outer block: NOB columns, or FLOWPANEL_LU32_OUTER (a multiple of 64 up to NOB_MAX)
static int outer_block()
{
    static const int v = [] {
        const char *e = getenv("FLOWPANEL_LU32_OUTER");
        const int o = e ? atoi(e) : NOB;
        return (o >= NPB && o <= NOB_MAX && o % NPB == 0) ? o : NOB;
    }();
    return v;
}

// Linv and Uinv of up to NOB_MAX / NPB inner diagonal blocks
size_t fp_getrf_nopiv_work_bytes() { return (size_t)2 * (NOB_MAX / NPB) * NPB * NPB * sizeof(float); }

// LU without pivoting of the n x n FP32 matrix A (row-major, ld a multiple of 4);
// info_dev receives 1 + the first row of the diagonal block whose pivot was zero
// or not finite (0 if none).
int fp_getrf_nopiv_f32(fp_ctx *ctx, float *A, int64_t n, int64_t ld, int *info_dev, void *work, float *YL,
                       float *YU)
{
    if (n <= 0) return FP_OK;
    if (ld % 4) return fp_fail(ctx, FP_ERR_ARG, "internal: no-pivot LU needs a leading dimension divisible by 4");
    hipLaunchKernelGGL(k_fill_info, dim3(1), dim3(64), 0, ctx->stream, info_dev, 0);
    int st;
    const int nob = outer_block();
    for (int64_t k = 0; k < n; k += nob) {
        const int64_t ob = std::min<int64_t>(nob, n - k), kend = k + ob;
        const int T = (int)((ob + NPB - 1) / NPB);
        for (int t = 0; t < T; ++t) {
            const int64_t c = k + (int64_t)NPB * t;
            float *Linv = (float *)work + (int64_t)2 * t * NPB * NPB, *Uinv = Linv + NPB * NPB;
            {
                ProfScope ps(ctx, "lu_nopiv_diag");
                hipLaunchKernelGGL(k_nopiv_diag, dim3(1), dim3(512), 0, ctx->stream, A, ld, n, c, Linv, Uinv, info_dev,
                                   YL, YU);
            }
            const int64_t below = n - c - NPB, inpanel = kend - c - NPB;
            if (below <= 0) break;
            const int nl = (int)((below + NPB - 1) / NPB), nu = inpanel > 0 ? (int)((inpanel + NPB - 1) / NPB) : 0;
            {
                ProfScope ps(ctx, "lu_nopiv_panel");
                hipLaunchKernelGGL(k_nopiv_panel, dim3((unsigned)(nl + nu)), dim3(256), 0, ctx->stream, A, ld, n, c,
                                   (const float *)Linv, (const float *)Uinv, nl, c + NPB);
            }
            FP_HIP(ctx, hipGetLastError());
            if (inpanel > 0 &&
                (st = gemm64_sub(ctx, below, inpanel, NPB, A + (c + NPB) * ld + c, ld, A + c * ld + c + NPB, ld,
                                 A + (c + NPB) * ld + c + NPB, ld)))
                return st;
        }
        const int64_t right = n - kend;
        if (right <= 0) continue;
        // U12 = L11^-1 A12 for the columns right of the outer block: U_t = L_tt^-1 (A_t - sum_{s<t} L_ts U_s)
        const int nu = (int)((right + NPB - 1) / NPB);
        for (int t = 0; t < T; ++t) {
            const int64_t c = k + (int64_t)NPB * t;
            const float *Linv = (const float *)work + (int64_t)2 * t * NPB * NPB, *Uinv = Linv + NPB * NPB;
            if (t > 0 && (st = gemm64_sub(ctx, NPB, right, c - k, A + c * ld + k, ld, A + k * ld + kend, ld,
                                          A + c * ld + kend, ld)))
                return st;
            ProfScope ps(ctx, "lu_nopiv_panel");
            hipLaunchKernelGGL(k_nopiv_panel, dim3((unsigned)nu), dim3(256), 0, ctx->stream, A, ld, n, c, Linv, Uinv, 0,
                               kend);
            FP_HIP(ctx, hipGetLastError());
        }
        // the rank-ob trailing update
        if ((st = fp_gemm_sub<float>(ctx, right, right, ob, A + kend * ld + k, ld, A + k * ld + kend, ld,
                                     A + kend * ld + kend, ld)))
            return st;
    }
    return FP_OK;
}

// ---- the same factorisation for many bodies at once --------------------------
// One launch per step covers every body (blockIdx.y or blockIdx.x = body): the bodies'
// serial chains of 64-column diagonal blocks run side by side, and the small products
// of different bodies fill the chip together.  Global outer blocks [kk, kk + NOB);
// a body takes part in the steps inside its order.

__global__ void k_fill_info_lub(const fp_lu_desc *__restrict__ desc, int nb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) *desc[b].info = 0;
}

// inner step t (column c = kk + 64 t) of body blockIdx.x
__global__ __launch_bounds__(512) void k_nopiv_diag_lub(const fp_lu_desc *__restrict__ desc, int64_t c, int t)
{
    const fp_lu_desc d = desc[blockIdx.x];
    if (c >= d.n) return;   // uniform per workgroup
    float *Linv = d.work + (int64_t)2 * t * NPB * NPB;
    nopiv_diag_dev(d.A, d.ld, d.n, c, Linv, Linv + NPB * NPB, d.info, nullptr, nullptr);
}

// mode 0: the panel tiles of inner step t (L tiles below, U tiles inside the outer block);
// mode 1: the U-row tiles of inner block t right of the outer block
__global__ __launch_bounds__(256) void k_nopiv_panel_lub(const fp_lu_desc *__restrict__ desc, int64_t kk, int t, int mode)
{
    const fp_lu_desc d = desc[blockIdx.y];
    const int64_t c = kk + (int64_t)NPB * t;
    const float *Linv = d.work + (int64_t)2 * t * NPB * NPB, *Uinv = Linv + NPB * NPB;
    if (mode == 0) {
        const int64_t below = d.n - c - NPB;
        if (c >= d.n || below <= 0) return;
        const int64_t kend = min(kk + (int64_t)NOB, d.n), inpanel = kend - c - NPB;
        const int nl = (int)((below + NPB - 1) / NPB), nu = inpanel > 0 ? (int)((inpanel + NPB - 1) / NPB) : 0;
        if ((int)blockIdx.x >= nl + nu) return;
        nopiv_panel_dev(d.A, d.ld, d.n, c, Linv, Uinv, nl, c + NPB, blockIdx.x);
    } else {
        const int64_t kend = kk + NOB, right = d.n - kend;
        if (right <= 0 || (int64_t)blockIdx.x >= (right + NPB - 1) / NPB) return;
        nopiv_panel_dev(d.A, d.ld, d.n, c, Linv, Uinv, 0, kend, blockIdx.x);
    }
}

// mode 0: the in-block update after inner step t; mode 1: the U rows' update before inner block t
__global__ __launch_bounds__(256) void k_gemm64_lub(const fp_lu_desc *__restrict__ desc, int64_t kk, int t, int mode)
{
    const fp_lu_desc d = desc[blockIdx.y];
    const int64_t c = kk + (int64_t)NPB * t, ld = d.ld;
    float *A = d.A;
    if (mode == 0) {
        const int64_t below = d.n - c - NPB, kend = min(kk + (int64_t)NOB, d.n), inpanel = kend - c - NPB;
        if (c >= d.n || below <= 0 || inpanel <= 0) return;
        if ((int64_t)blockIdx.x >= ((below + NPB - 1) / NPB) * ((inpanel + NPB - 1) / NPB)) return;
        gemm64_dev(below, inpanel, NPB, A + (c + NPB) * ld + c, ld, A + c * ld + c + NPB, ld,
                   A + (c + NPB) * ld + c + NPB, ld, blockIdx.x);
    } else {
        const int64_t kend = kk + NOB, right = d.n - kend;
        if (right <= 0 || (int64_t)blockIdx.x >= (right + NPB - 1) / NPB) return;
        gemm64_dev(NPB, right, c - kk, A + c * ld + kk, ld, A + kk * ld + kend, ld, A + c * ld + kend, ld, blockIdx.x);
    }
}

int fp_getrf_nopiv_f32_batched(fp_ctx *ctx, const fp_lu_desc *d_desc, const int64_t *n_host, int nb)
{
    if (nb <= 0) return FP_OK;
    if (nb > 65535) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "batched LU: too many bodies");
    int64_t nmax = 0;
    for (int b = 0; b < nb; ++b) nmax = std::max(nmax, n_host[b]);
    hipLaunchKernelGGL(k_fill_info_lub, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, ctx->stream, d_desc, nb);
    auto tiles = [](int64_t x) { return x > 0 ? (x + NPB - 1) / NPB : (int64_t)0; };
    int st;
    for (int64_t kk = 0; kk < nmax; kk += NOB) {
        for (int t = 0; t < NOB / NPB; ++t) {
            const int64_t c = kk + (int64_t)NPB * t;
            if (c >= nmax) break;
            int64_t gp = 0, gg = 0;   // the largest grids over the bodies
            for (int b = 0; b < nb; ++b) {
                const int64_t n = n_host[b];
                if (c >= n) continue;
                const int64_t below = n - c - NPB, inpanel = std::min(kk + (int64_t)NOB, n) - c - NPB;
                gp = std::max(gp, tiles(below) + (below > 0 ? tiles(inpanel) : 0));
                if (below > 0 && inpanel > 0) gg = std::max(gg, tiles(below) * tiles(inpanel));
            }
            {
                ProfScope ps(ctx, "lu_nopiv_diag");
                hipLaunchKernelGGL(k_nopiv_diag_lub, dim3((unsigned)nb), dim3(512), 0, ctx->stream, d_desc, c, t);
            }
            if (gp > 0) {
                ProfScope ps(ctx, "lu_nopiv_panel");
                hipLaunchKernelGGL(k_nopiv_panel_lub, dim3((unsigned)gp, (unsigned)nb), dim3(256), 0, ctx->stream,
                                   d_desc, kk, t, 0);
            }
            if (gg > 0) {
                ProfScope ps(ctx, "gemm_f32");
                hipLaunchKernelGGL(k_gemm64_lub, dim3((unsigned)gg, (unsigned)nb), dim3(256), 0, ctx->stream, d_desc, kk,
                                   t, 0);
            }
            FP_HIP(ctx, hipGetLastError());
        }
        int64_t right = 0;
        for (int b = 0; b < nb; ++b) right = std::max(right, n_host[b] - kk - NOB);
        if (right <= 0) continue;
        for (int t = 0; t < NOB / NPB; ++t) {
            if (t > 0) {
                ProfScope ps(ctx, "gemm_f32");
                hipLaunchKernelGGL(k_gemm64_lub, dim3((unsigned)tiles(right), (unsigned)nb), dim3(256), 0, ctx->stream,
                                   d_desc, kk, t, 1);
            }
            ProfScope ps(ctx, "lu_nopiv_panel");
            hipLaunchKernelGGL(k_nopiv_panel_lub, dim3((unsigned)tiles(right), (unsigned)nb), dim3(256), 0, ctx->stream,
                               d_desc, kk, t, 1);
            FP_HIP(ctx, hipGetLastError());
        }
        if ((st = fp_gemm_sub_lub(ctx, d_desc, nb, kk, NOB, right))) return st;
    }
    return FP_OK;
}

// Explicit inverses of the no-pivot factors, for refinement sweeps that then cost
// two GEMVs instead of two blocked triangular solves.  YL = -L^-1 and YU = -U^-1
// (row-major, ld as A; zero outside their triangles on entry, the diagonal 64 x 64
// blocks written by fp_getrf_nopiv_f32).  With Y = -X the block recurrences
//   lower:  Y_JI = Y_JJ L_JI Y_II,   upper:  Y_IJ = Y_II U_IJ Y_JJ   (I before J)
// are two "C -= A B" products each through T = -L_JI Y_II (resp. -U_IJ Y_JJ);
// blocks merge pairwise, 64 -> 128 -> ..., one batched launch per product and
// level.  U^-1 L^-1 = YU YL: the signs cancel in the solve.  T: n x ld scratch.
int fp_trtri_nopiv_f32(fp_ctx *ctx, const float *A, int64_t n, int64_t ld, float *YL, float *YU, float *T)
{
    ProfScope ps(ctx, "lu_nopiv_trtri");
    int st;
    for (int64_t s = NPB; s < n; s *= 2) {
        const int64_t P = n / (2 * s), rem = n - 2 * s * P;
        const int64_t step = 2 * s * ld + 2 * s;   // between consecutive pairs
        if (P > 0) {
            if ((st = fp_fill_f32(ctx, T, P * s * ld, 0.f))) return st;
            // lower: T_p = -L_JI Y_II, Y_JI -= Y_JJ T_p
            if ((st = fp_gemm_sub_batched_f32(ctx, s, s, s, A + s * ld, ld, step, YL, ld, step, T, ld, s * ld, (int)P)))
                return st;
            if ((st = fp_gemm_sub_batched_f32(ctx, s, s, s, YL + s * ld + s, ld, step, T, ld, s * ld, YL + s * ld, ld,
                                              step, (int)P)))
                return st;
            if ((st = fp_fill_f32(ctx, T, P * s * ld, 0.f))) return st;
            // upper: T_p = -U_IJ Y_JJ, Y_IJ -= Y_II T_p
            if ((st = fp_gemm_sub_batched_f32(ctx, s, s, s, A + s, ld, step, YU + s * ld + s, ld, step, T, ld, s * ld,
                                              (int)P)))
                return st;
            if ((st = fp_gemm_sub_batched_f32(ctx, s, s, s, YU, ld, step, T, ld, s * ld, YU + s, ld, step, (int)P)))
                return st;
        }
        if (rem > s) {   // the last pair: I = [r, r + s), J = [r + s, n)
            const int64_t r = 2 * s * P, sj = rem - s, rj = r + s;
            if ((st = fp_fill_f32(ctx, T, s * ld, 0.f))) return st;
            if ((st = fp_gemm_sub_batched_f32(ctx, sj, s, s, A + rj * ld + r, ld, 0, YL + r * ld + r, ld, 0, T, ld, 0,
                                              1)))
                return st;
            if ((st = fp_gemm_sub_batched_f32(ctx, sj, s, sj, YL + rj * ld + rj, ld, 0, T, ld, 0, YL + rj * ld + r, ld,
                                              0, 1)))
                return st;
            if ((st = fp_fill_f32(ctx, T, s * ld, 0.f))) return st;
            if ((st = fp_gemm_sub_batched_f32(ctx, s, sj, sj, A + r * ld + rj, ld, 0, YU + rj * ld + rj, ld, 0, T, ld,
                                              0, 1)))
                return st;
            if ((st = fp_gemm_sub_batched_f32(ctx, s, sj, s, YU + r * ld + r, ld, 0, T, ld, 0, YU + r * ld + rj, ld, 0,
                                              1)))
                return st;
        }
    }
    return FP_OK;
}

// Workspace of fp_inverse_rbt_f32 for order n: the randomized copy (then the
// product), -L^-1, -U^-1, the triangular-inverse scratch (FP32, n4 x ld4 each),
// the butterflies and the LU's own work.
static int64_t inv_ld(int64_t n4) { return std::max<int64_t>(64, (n4 + 63) / 64 * 64); }
size_t fp_inverse_rbt_work_bytes(int64_t n)
{
    const int64_t n4 = fp_rbt_size(n), ld4 = inv_ld(n4);
    return (size_t)4 * n4 * ld4 * sizeof(float) + fp_rbt_weights_bytes(n) + fp_getrf_nopiv_work_bytes() + 1024;
}

// E^-1 in place (n x n, FP64, lde) by the LU without pivoting of the randomized
// FP32 matrix U^T E V, the explicit inverses of its factors and one product:
// E^-1 = V (U_r^-1 L_r^-1) U^T.  No chain of n / 64 Gauss-Jordan passes with
// their products: the factorisation's own chain is a few launches per 64
// columns and the rest are large MFMA products.  info_dev as fp_getrf_nopiv_f32.
// Nothing is allocated and the host is not synchronised (a side stream may run it).
int fp_inverse_rbt_f32(fp_ctx *ctx, double *E, int64_t n, int64_t lde, void *work, size_t work_bytes, int *info_dev,
                       uint64_t seed)
{
    if (n <= 0) return FP_OK;
    if (work_bytes < fp_inverse_rbt_work_bytes(n)) return fp_fail(ctx, FP_ERR_ARG, "internal: inverse workspace too small");
    ProfScope ps(ctx, "inverse_rbt");
    const int64_t n4 = fp_rbt_size(n), m = n4 / 4, ld4 = inv_ld(n4), sz = n4 * ld4;
    float *Ar = (float *)work, *YL = Ar + sz, *YU = YL + sz, *T = YU + sz;
    double *W = (double *)(T + sz);
    void *nwork = (char *)W + fp_rbt_weights_bytes(n);
    hipLaunchKernelGGL(k_rbt_weights, dim3((unsigned)((2 * m + 127) / 128)), dim3(128), 0, ctx->stream, m, seed, W);
    int st;
    if ((st = fp_rbt_matrix<double>(ctx, E, n, lde, Ar, ld4, W))) return st;
    if ((st = fp_fill_f32(ctx, YL, sz, 0.f)) || (st = fp_fill_f32(ctx, YU, sz, 0.f))) return st;
    if ((st = fp_getrf_nopiv_f32(ctx, Ar, n4, ld4, info_dev, nwork, YL, YU))) return st;
    if ((st = fp_trtri_nopiv_f32(ctx, Ar, n4, ld4, YL, YU, T))) return st;
    // Ar <- -(YU YL) = -U^-1 L^-1 (the factors are no longer needed)
    if ((st = fp_fill_f32(ctx, Ar, sz, 0.f))) return st;
    if ((st = fp_gemm_sub<float>(ctx, n4, n4, n4, YU, ld4, YL, ld4, Ar, ld4))) return st;
    hipLaunchKernelGGL(k_rbt_unmatrix, dim3((unsigned)((m + 255) / 256), (unsigned)m), dim3(256), 0, ctx->stream,
                       (const float *)Ar, ld4, m, E, n, lde, (const double *)W);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
