// fp_dense.hip -- dense LU with partial (tournament) pivoting on MFMA, triangular
// solves, GEMV and the block-Jacobi preconditioner (gfx950).
//
// Replaces LAPACK getrf/getrs behind `solve_ludiv!` (src/FLOWPanel_linearsolver.jl:57-78,
// calc_Alu :173-234).  Matrices are row-major (A[i*ld + j]).
//
// getrf: right-looking blocked LU, outer blocks of NB = 512 columns (the block
// row U12 = L11^-1 A12 and the trailing update are MFMA GEMMs with k = 512),
// each outer block factored in sub-panels of PW = 32 columns.  A sub-panel
// (m x 32) is factored with tournament pivoting (TSLU, Grigori-Demmel-Xiang
// 2011): every workgroup runs partial pivoting on a chunk of <= 256 rows (one
// row per thread, in registers) and proposes 32 candidate rows; candidate sets
// are merged 16 at a time by partial pivoting on their original values until
// 32 pivot rows remain.  This needs no grid-wide synchronisation: one
// sub-panel is a few short launches instead of one grid barrier per column.
#include "fp_internal.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <type_traits>

typedef float fp_f4 __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaOp;
template <> struct MfmaOp<float> {
    typedef fp_f4 acc;
    static __device__ __forceinline__ acc run(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator element r of lane l holds D(4*(l/16) + r, l%16)
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ----------------------------------------------------------------------------
// C[m x n] -= A[m x k] * B[k x n], all row-major.  128 x 128 tile per 256-thread
// workgroup (2 x 2 waves of 64 x 64), K step 32 through LDS with register
// prefetch of the next step, 16-byte global loads.  FP32 runs v_mfma_f32_16x16x4;
// FP64 runs v_mfma_f64_4x4x4_4b (k_gemm_sub_f64 below): on gfx950 the FP64
// 16x16x4 form issues at half the FP64 matrix rate (measured 36 of 78.6 TFLOP/s
// with every SIMD busy, register-only), the four-block 4x4x4 form at 67-73.
// FP32 operand layout (16x16x4): lane l supplies A(l%16, l/16) and B(l/16, l%16);
// accumulator rows as in MfmaOp<float>::row (measured on gfx950 with an
// in-bounds lane-layout probe).
// Tile order: workgroup b runs on XCD b % 8; every XCD is given a contiguous
// range of a grouped order (8 tile rows walked together), so the A rows of a
// group stay in that XCD's L2 while it sweeps across the B columns.
constexpr int GB_M = 128, GB_N = 128, GB_K = 32, GB_PAD = 4, GB_GROUP = 8, GB_XCD = 8;

template <typename T>
__device__ __forceinline__ void ld16(const T *__restrict__ p, bool full, bool vec, int valid, T *out)
{
    if (full && vec) {
        if (sizeof(T) == 8) {
#pragma unroll
            for (int e = 0; e < 16; e += 2) {
                const double2 v = *reinterpret_cast<const double2 *>(p + e);
                out[e] = (T)v.x; out[e + 1] = (T)v.y;
            }
        } else {
#pragma unroll
            for (int e = 0; e < 16; e += 4) {
                const float4 v = *reinterpret_cast<const float4 *>(p + e);
                out[e] = (T)v.x; out[e + 1] = (T)v.y; out[e + 2] = (T)v.z; out[e + 3] = (T)v.w;
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) out[e] = e < valid ? p[e] : T(0);
    }
}

// FP32: K steps of GB_KF through two LDS buffers, one barrier per step (a step's
// stores go to the buffer the previous step did not read; a wave reaching the
// store after next has passed the barrier that every wave reaches only after
// finishing its reads of that buffer).  Measured on 4,608^2 x k products:
// 87 against 80 TFLOP/s at k = 256 and 107 against 94 at k = 1,024 for K steps
// of 32 through one buffer (register-only 16x16x4 rate: 125 TFLOP/s).
constexpr int GB_KF = 16;

// n_ entries from p (16-byte aligned groups when vec and full), zero beyond `valid`
template <int N_>
__device__ __forceinline__ void ldf(const float *__restrict__ p, bool full, bool vec, int valid, float *out)
{
    if (full && vec) {
#pragma unroll
        for (int e = 0; e < N_; e += 4) {
            const float4 v = *reinterpret_cast<const float4 *>(p + e);
            out[e] = v.x; out[e + 1] = v.y; out[e + 2] = v.z; out[e + 3] = v.w;
        }
    } else {
        // clamped addresses: every load in bounds and unconditional, the rest zeroed by a product
#pragma unroll
        for (int e = 0; e < N_; ++e) out[e] = valid > 0 ? p[min(e, valid - 1)] * (e < valid ? 1.f : 0.f) : 0.f;
    }
}

// Tile of workgroup slot bx in the XCD-grouped order of an m x n product (GB_M x GB_N
// tiles): false when the slot has no tile (the grid may be larger than the product).
__device__ __forceinline__ bool gemm_tile_of(int64_t bx, int64_t m, int64_t n, int64_t &m0, int64_t &n0)
{
    const int64_t tiles_m = (m + GB_M - 1) / GB_M, tiles_n = (n + GB_N - 1) / GB_N;
    const int64_t ntiles = tiles_m * tiles_n;
    const int64_t per = (ntiles + GB_XCD - 1) / GB_XCD;
    if (bx >= per * GB_XCD) return false;
    const int64_t t = (bx % GB_XCD) * per + bx / GB_XCD;
    if (t >= ntiles) return false;
    const int64_t in_group = (int64_t)GB_GROUP * tiles_n;
    const int64_t first_m = (t / in_group) * GB_GROUP;
    const int64_t gsize = min((int64_t)GB_GROUP, tiles_m - first_m);
    const int64_t tm = first_m + (t % in_group) % gsize, tn = (t % in_group) / gsize;
    m0 = tm * GB_M;
    n0 = tn * GB_N;
    return true;
}

// One 128 x 128 tile (m0, n0) of C -= A B in FP32 (the body of k_gemm_sub, shared
// with the batched LU's trailing updates).
__device__ __forceinline__ void gemm_sub_f32_tile(int64_t m, int64_t n, int64_t k, const float *__restrict__ A,
                                                  int64_t lda, const float *__restrict__ B, int64_t ldb,
                                                  float *__restrict__ C, int64_t ldc, int vec, int64_t m0,
                                                  int64_t n0)
{
    typedef float T;
    typedef typename MfmaOp<T>::acc acc_t;
    __shared__ T As[2][GB_KF][GB_M + GB_PAD];
    __shared__ T Bs[2][GB_KF][GB_N + GB_PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

    // A: row tid/2, k (tid%2)*8..+8 ; B: k row tid/16, cols (tid%16)*8..+8
    constexpr int AE = GB_M * GB_KF / 256, BE = GB_KF * GB_N / 256;
    const int arow = tid / (GB_KF / AE), ak = (tid % (GB_KF / AE)) * AE;
    const int bk = tid / (GB_N / BE), bc = (tid % (GB_N / BE)) * BE;
    T ra[AE], rb[BE];
    auto load = [&](int64_t k0) {
        const int64_t gr = m0 + arow, gk = k0 + ak;
        const int av = gr < m ? (int)max((int64_t)0, min((int64_t)AE, k - gk)) : 0;
        ldf<AE>(A + min(gr, m - 1) * lda + gk, av == AE, vec, av, ra);
        const int64_t gkb = k0 + bk, gc = n0 + bc;
        const int bv = gkb < k ? (int)max((int64_t)0, min((int64_t)BE, n - gc)) : 0;
        ldf<BE>(B + min(gkb, k - 1) * ldb + gc, bv == BE, vec, bv, rb);
    };
    load(0);
    int buf = 0;
    for (int64_t k0 = 0; k0 < k; k0 += GB_KF) {
#pragma unroll
        for (int e = 0; e < AE; ++e) As[buf][ak + e][arow] = ra[e];
#pragma unroll
        for (int e = 0; e < BE; ++e) Bs[buf][bk][bc + e] = rb[e];
        __syncthreads();
        if (k0 + GB_KF < k) load(k0 + GB_KF);
#pragma unroll
        for (int ks = 0; ks < GB_KF; ks += 4) {
            T fa[4], fb[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) fa[a] = As[buf][ks + (lane >> 4)][wm * 64 + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 4; ++b) fb[b] = Bs[buf][ks + (lane >> 4)][wn * 64 + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = MfmaOp<T>::run(fa[a], fb[b], acc[a][b]);
        }
        buf ^= 1;
    }
    // epilogue: every load of the lane's 64 entries of C is issued before the first
    // store (as far as the compiler knows rows of C could alias, and a load after a
    // store would wait for it: one memory round trip per entry); edge tiles clamp
    // their addresses and store only inside C
    const bool interior = m0 + GB_M <= m && n0 + GB_N <= n;   // uniform
    T cv[4][4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm * 64 + a * 16 + MfmaOp<T>::row(lane, r);
                const int64_t col = n0 + wn * 64 + b * 16 + (lane & 15);
                cv[a][b][r] = interior ? C[row * ldc + col] : C[min(row, m - 1) * ldc + min(col, n - 1)];
            }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm * 64 + a * 16 + MfmaOp<T>::row(lane, r);
                const int64_t col = n0 + wn * 64 + b * 16 + (lane & 15);
                if (interior || (row < m && col < n)) C[row * ldc + col] = cv[a][b][r] - acc[a][b][r];
            }
}

// The batched LU's trailing update of outer block [kk, kk + ob) for body blockIdx.y:
// C = A[kend:, kend:] -= A[kend:, kk:kend] A[kk:kend, kend:] (bodies with n <= kend idle).
__global__ __launch_bounds__(256) void k_gemm_sub_lub(const fp_lu_desc *__restrict__ desc, int64_t kk, int ob)
{
    const fp_lu_desc d = desc[blockIdx.y];
    const int64_t kend = kk + ob, right = d.n - kend;
    if (right <= 0) return;
    int64_t m0, n0;
    if (!gemm_tile_of(blockIdx.x, right, right, m0, n0)) return;
    const float *Ap = d.A + kend * d.ld + kk, *Bp = d.A + kk * d.ld + kend;
    float *Cp = d.A + kend * d.ld + kend;
    const int vec = (((uintptr_t)Ap | (uintptr_t)Bp) & 15) == 0 && (d.ld * sizeof(float)) % 16 == 0;
    gemm_sub_f32_tile(right, right, ob, Ap, d.ld, Bp, d.ld, Cp, d.ld, vec, m0, n0);
}

int fp_gemm_sub_lub(fp_ctx *ctx, const fp_lu_desc *d_desc, int nb, int64_t kk, int ob, int64_t max_right)
{
    if (max_right <= 0 || nb <= 0) return FP_OK;
    const int64_t t = (max_right + GB_M - 1) / GB_M;
    const int64_t grid = (t * t + GB_XCD - 1) / GB_XCD * GB_XCD;
    ProfScope ps(ctx, "gemm_f32");
    hipLaunchKernelGGL(k_gemm_sub_lub, dim3((unsigned)grid, (unsigned)nb), dim3(256), 0, ctx->stream, d_desc, kk, ob);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(int64_t m, int64_t n, int64_t k, const T *__restrict__ A,
                                                  int64_t lda, const T *__restrict__ B, int64_t ldb,
                                                  T *__restrict__ C, int64_t ldc, int vec, int64_t sA = 0,
                                                  int64_t sB = 0, int64_t sC = 0)
{
    static_assert(sizeof(T) == 4, "FP64 products run k_gemm_sub_f64");
    // strided batch: grid row y runs product y (all products the same shape)
    int64_t m0, n0;
    if (!gemm_tile_of(blockIdx.x, m, n, m0, n0)) return;
    gemm_sub_f32_tile(m, n, k, A + (int64_t)blockIdx.y * sA, lda, B + (int64_t)blockIdx.y * sB, ldb,
                      C + (int64_t)blockIdx.y * sC, ldc, vec, m0, n0);
}

// FP64 tile kernel on v_mfma_f64_4x4x4_4b.  Operand layout of that instruction
// (in-bounds lane probe on gfx950): lane L supplies A(4*blk + i, k) and B(k, j)
// of block blk = (L>>2)&3 with k = L>>4, i = j = L&3, and receives
// D(4*blk + (L>>4), L&3).  Every block is given the same B columns, so one
// instruction is a 16 x 4 x 4 product; a wave's 64 x 64 tile is 4 row fragments
// x 16 column fragments, 64 independent accumulators per lane.
// LDS placement: slot a*16 + p of the wave's A rows holds tile row p*4 + a and
// slot b*4 + j of its B columns holds tile column j*16 + b.  Fragment reads are
// then 16 consecutive doubles (A) and 4 consecutive doubles per k row (B) --
// no bank conflicts -- and lane L owns the 4 x 16 block of C at rows P*4..P*4+3,
// P = 4*((L>>2)&3) + (L>>4), columns (L&3)*16..+15 of its wave tile, written
// back with 16-byte accesses.
__global__ __launch_bounds__(256) void k_gemm_sub_f64(int64_t m, int64_t n, int64_t k, const double *__restrict__ A,
                                                      int64_t lda, const double *__restrict__ B, int64_t ldb,
                                                      double *__restrict__ C, int64_t ldc, int vec)
{
    __shared__ double As[GB_K][GB_M + GB_PAD];
    __shared__ double Bs[GB_K][GB_N + GB_PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t tiles_m = (m + GB_M - 1) / GB_M, tiles_n = (n + GB_N - 1) / GB_N;
    const int64_t ntiles = tiles_m * tiles_n;
    const int64_t per = (ntiles + GB_XCD - 1) / GB_XCD;
    const int64_t t = (int64_t)(blockIdx.x % GB_XCD) * per + blockIdx.x / GB_XCD;
    if (t >= ntiles) return;
    const int64_t in_group = (int64_t)GB_GROUP * tiles_n;
    const int64_t first_m = (t / in_group) * GB_GROUP;
    const int64_t gsize = min((int64_t)GB_GROUP, tiles_m - first_m);
    const int64_t tm = first_m + (t % in_group) % gsize, tn = (t % in_group) / gsize;
    const int64_t m0 = tm * GB_M, n0 = tn * GB_N;

    double acc[4][16];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 16; ++b) acc[a][b] = 0.0;

    // A: thread loads 16 k of tile row arow into LDS slot aslot; B: 16 columns of k row bk
    const int aslot = tid >> 1, ak = (tid & 1) * 16;
    const int arow = (aslot & 64) + (aslot & 15) * 4 + ((aslot & 63) >> 4);
    const int bk = tid >> 3, bc = (tid & 7) * 16, bslot = (tid & 4) * 16 + (tid & 3);
    double ra[16], rb[16];
    auto load = [&](int64_t k0) {
        const int64_t gr = m0 + arow, gk = k0 + ak;
        const int av = gr < m ? (int)max((int64_t)0, min((int64_t)16, k - gk)) : 0;
        ld16<double>(A + gr * lda + gk, av == 16, vec, av, ra);
        const int64_t gkb = k0 + bk, gc = n0 + bc;
        const int bv = gkb < k ? (int)max((int64_t)0, min((int64_t)16, n - gc)) : 0;
        ld16<double>(B + gkb * ldb + gc, bv == 16, vec, bv, rb);
    };
    load(0);
    for (int64_t k0 = 0; k0 < k; k0 += GB_K) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) As[ak + e][aslot] = ra[e];
#pragma unroll
        for (int e = 0; e < 16; ++e) Bs[bk][bslot + e * 4] = rb[e];
        __syncthreads();
        if (k0 + GB_K < k) load(k0 + GB_K);
#pragma unroll
        for (int ks = 0; ks < GB_K; ks += 4) {
            double fa[4], fb[16];
#pragma unroll
            for (int a = 0; a < 4; ++a) fa[a] = As[ks + (lane >> 4)][wm * 64 + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 16; ++b) fb[b] = Bs[ks + (lane >> 4)][wn * 64 + b * 4 + (lane & 3)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 16; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_4x4x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
    }
    // epilogue: every load of the lane's 4 x 16 block of C is issued before the
    // first store (rows of C could alias as far as the compiler knows, which
    // would otherwise serialise one memory round trip per row)
    const int P = 4 * ((lane >> 2) & 3) + (lane >> 4);
    const int64_t col0 = n0 + wn * 64 + (lane & 3) * 16;
    const int64_t row0 = m0 + wm * 64 + P * 4;
    const bool cvec = vec && (ldc & 1) == 0 && ((uintptr_t)C & 15) == 0 && col0 + 16 <= n;
    double cv[4][16];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double *c = C + (row0 + a) * ldc + col0;
        if (row0 + a < m && cvec) {
#pragma unroll
            for (int b = 0; b < 16; b += 2) {
                const double2 v = *reinterpret_cast<const double2 *>(c + b);
                cv[a][b] = v.x;
                cv[a][b + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int b = 0; b < 16; ++b) cv[a][b] = (row0 + a < m && col0 + b < n) ? c[b] : 0.0;
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        if (row0 + a >= m) continue;
        double *c = C + (row0 + a) * ldc + col0;
        if (cvec) {
#pragma unroll
            for (int b = 0; b < 16; b += 2)
                *reinterpret_cast<double2 *>(c + b) = double2{cv[a][b] - acc[a][b], cv[a][b + 1] - acc[a][b + 1]};
        } else {
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if (col0 + b < n) c[b] = cv[a][b] - acc[a][b];
        }
    }
}

template <typename T>
static int gemm_sub(fp_ctx *ctx, int64_t m, int64_t n, int64_t k, const T *A, int64_t lda, const T *B, int64_t ldb,
                    T *C, int64_t ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return FP_OK;
    const int64_t ntiles = ((m + GB_M - 1) / GB_M) * ((n + GB_N - 1) / GB_N);
    const int64_t grid = (ntiles + GB_XCD - 1) / GB_XCD * GB_XCD;
    // 16-byte loads need 16-byte aligned rows and offsets
    const int vec = (((uintptr_t)A | (uintptr_t)B) & 15) == 0 && (lda * sizeof(T)) % 16 == 0 &&
                    (ldb * sizeof(T)) % 16 == 0;
    ProfScope ps(ctx, sizeof(T) == 8 ? "gemm_f64" : "gemm_f32");
    if constexpr (sizeof(T) == 8)
        hipLaunchKernelGGL(k_gemm_sub_f64, dim3((unsigned)grid), dim3(256), 0, ctx->stream, m, n, k,
                           (const double *)A, lda, (const double *)B, ldb, (double *)C, ldc, vec);
    else
        hipLaunchKernelGGL(k_gemm_sub<T>, dim3((unsigned)grid), dim3(256), 0, ctx->stream, m, n, k, A, lda, B, ldb,
                           C, ldc, vec);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

template <typename T>
int fp_gemm_sub(fp_ctx *ctx, int64_t m, int64_t n, int64_t k, const T *A, int64_t lda, const T *B, int64_t ldb, T *C,
                int64_t ldc)
{
    return gemm_sub<T>(ctx, m, n, k, A, lda, B, ldb, C, ldc);
}

// C_b -= A_b B_b for b < batch, operand b at element offset b * s? from the
// first (FP32): one launch for a level of small products.
int fp_gemm_sub_batched_f32(fp_ctx *ctx, int64_t m, int64_t n, int64_t k, const float *A, int64_t lda, int64_t sA,
                            const float *B, int64_t ldb, int64_t sB, float *C, int64_t ldc, int64_t sC, int batch)
{
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return FP_OK;
    if (batch > 65535) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "batched gemm: too many products");
    const int64_t ntiles = ((m + GB_M - 1) / GB_M) * ((n + GB_N - 1) / GB_N);
    const int64_t grid = (ntiles + GB_XCD - 1) / GB_XCD * GB_XCD;
    const int vec = (((uintptr_t)A | (uintptr_t)B) & 15) == 0 && (lda * 4) % 16 == 0 && (ldb * 4) % 16 == 0 &&
                    (sA * 4) % 16 == 0 && (sB * 4) % 16 == 0;
    ProfScope ps(ctx, "gemm_f32");
    hipLaunchKernelGGL(k_gemm_sub<float>, dim3((unsigned)grid, (unsigned)batch), dim3(256), 0, ctx->stream, m, n, k,
                       A, lda, B, ldb, C, ldc, vec, sA, sB, sC);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
template int fp_gemm_sub<double>(fp_ctx *, int64_t, int64_t, int64_t, const double *, int64_t, const double *, int64_t,
                                 double *, int64_t);
template int fp_gemm_sub<float>(fp_ctx *, int64_t, int64_t, int64_t, const float *, int64_t, const float *, int64_t,
                                float *, int64_t);

// Least-squares system of the watertight formulation from C = -(G^T G) (as left
// by fp_gemm_sub on a zero C): C := G^T G, then every prescribed strength p gets
// the row and column of the identity and the right-hand side entry val_p, which
// leaves the other unknowns' normal equations (G_red^T G_red) untouched.
__global__ void k_lsq_fix(double *__restrict__ C, int64_t n, int64_t ld, const int64_t *__restrict__ pidx, int np,
                          double *__restrict__ b, const double *__restrict__ pval)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= n) return;
    double v = -C[i * ld + j];
    bool hit = false;
    for (int q = 0; q < np; ++q) {
        if (pidx[q] == i || pidx[q] == j) hit = true;
    }
    if (hit) v = (i == j) ? 1.0 : 0.0;
    C[i * ld + j] = v;
    if (j == 0)
        for (int q = 0; q < np; ++q)
            if (pidx[q] == i) b[i] = pval[q];
}

int fp_lsq_fix(fp_ctx *ctx, double *C, int64_t n, int64_t ld, const int64_t *pidx, int np, double *b,
               const double *pval)
{
    if (n <= 0) return FP_OK;
    hipLaunchKernelGGL(k_lsq_fix, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, ctx->stream, C, n, ld,
                       pidx, np, b, pval);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// rhs[i] -= sum_q val_q G[i, p_q]: the prescribed strengths moved to the right-hand side
__global__ void k_lsq_rhs(const double *__restrict__ G, int64_t n, int64_t ld, const int64_t *__restrict__ pidx,
                          const double *__restrict__ pval, int np, double *__restrict__ rhs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = rhs[i];
    for (int q = 0; q < np; ++q) s -= pval[q] * G[i * ld + pidx[q]];
    rhs[i] = s;
}

int fp_lsq_rhs(fp_ctx *ctx, const double *G, int64_t n, int64_t ld, const int64_t *pidx, const double *pval, int np,
               double *rhs)
{
    if (n <= 0) return FP_OK;
    hipLaunchKernelGGL(k_lsq_rhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, G, n, ld, pidx, pval,
                       np, rhs);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// Bordered matrix of the least-squares solve: column p_q of B gets the border
// vector w_q added (w: np x n, row q contiguous).
__global__ void k_lsq_border(double *__restrict__ B, int64_t n, int64_t ld, const int64_t *__restrict__ pidx, int np,
                             const double *__restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int q = 0; q < np; ++q) B[i * ld + pidx[q]] += w[(int64_t)q * n + i];
}

int fp_lsq_border(fp_ctx *ctx, double *B, int64_t n, int64_t ld, const int64_t *pidx, int np, const double *w)
{
    if (n <= 0 || np <= 0) return FP_OK;
    hipLaunchKernelGGL(k_lsq_border, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, B, n, ld, pidx, np,
                       w);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// X (nb x ncols, row-major ld) := L^-1 X with L (nb x nb, ld) unit lower, nb <= 256.
// One workgroup per 32 columns; X in LDS; diagonal 32 x 32 blocks solved by one
// lane per column in registers, the rows below updated by the whole workgroup.
constexpr int TRSM_TC = 32;

template <typename T>
__global__ __launch_bounds__(256) void k_trsm_llu(int nb, int64_t ncols, const T *__restrict__ L, int64_t ldl,
                                                  T *__restrict__ X, int64_t ldx)
{
    __shared__ T Xs[256][TRSM_TC + 1];
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * TRSM_TC;
    const int tc = (int)min((int64_t)TRSM_TC, ncols - c0);
    for (int e = tid; e < nb * TRSM_TC; e += 256) {
        const int r = e / TRSM_TC, c = e % TRSM_TC;
        Xs[r][c] = c < tc ? X[(int64_t)r * ldx + c0 + c] : T(0);
    }
    __syncthreads();
    for (int b = 0; b < nb; b += 32) {
        const int bs = min(32, nb - b);
        if (tid < TRSM_TC) {
            T x[32];
#pragma unroll
            for (int r = 0; r < 32; ++r) x[r] = r < bs ? Xs[b + r][tid] : T(0);
#pragma unroll
            for (int r = 1; r < 32; ++r) {
                if (r < bs) {
                    T acc = x[r];
                    const T *lr = L + (int64_t)(b + r) * ldl + b;
#pragma unroll
                    for (int q = 0; q < 32; ++q)
                        if (q < r) acc -= lr[q] * x[q];
                    x[r] = acc;
                }
            }
#pragma unroll
            for (int r = 0; r < 32; ++r)
                if (r < bs) Xs[b + r][tid] = x[r];
        }
        __syncthreads();
        const int rows = nb - b - bs;
        for (int e = tid; e < rows * TRSM_TC; e += 256) {
            const int r = b + bs + e / TRSM_TC, c = e % TRSM_TC;
            const T *lr = L + (int64_t)r * ldl + b;
            T acc = Xs[r][c];
            for (int q = 0; q < bs; ++q) acc -= lr[q] * Xs[b + q][c];
            Xs[r][c] = acc;
        }
        __syncthreads();
    }
    for (int e = tid; e < nb * TRSM_TC; e += 256) {
        const int r = e / TRSM_TC, c = e % TRSM_TC;
        if (c < tc) X[(int64_t)r * ldx + c0 + c] = Xs[r][c];
    }
}

template <typename T>
static int trsm_llu(fp_ctx *ctx, int nb, int64_t ncols, const T *L, int64_t ldl, T *X, int64_t ldx)
{
    if (ncols <= 0 || nb <= 0) return FP_OK;
    ProfScope ps(ctx, "trsm");
    hipLaunchKernelGGL(k_trsm_llu<T>, dim3((unsigned)((ncols + TRSM_TC - 1) / TRSM_TC)), dim3(256), 0,
                       ctx->stream, nb, ncols, L, ldl, X, ldx);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// Row interchanges: rows i in [k1, k2) swapped with ipiv[i] (sequentially), for
// every column; one lane per column (row-major rows are contiguous).
template <typename T>
__global__ void k_laswp(T *__restrict__ A, int64_t ld, int64_t ncols, int64_t ld_rows, const int *__restrict__ ipiv,
                        int k1, int k2)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    for (int i = k1; i < k2; ++i) {
        const int p = ipiv[i];
        if (p != i && p >= 0 && (int64_t)p < ld_rows) {
            const T a = A[(int64_t)i * ld + c];
            A[(int64_t)i * ld + c] = A[(int64_t)p * ld + c];
            A[(int64_t)p * ld + c] = a;
        }
    }
}

// ----------------------------------------------------------------------------
// Tournament pivoting.
constexpr int PW = 32;      // sub-panel width
constexpr int CH = 256;     // rows per candidate chunk
constexpr int MR = 512;     // rows per merge (16 candidate sets)
constexpr int MERGE_F = MR / PW;

// Partial pivoting of a tile of up to NT rows x cols (cols <= PW) with one row
// per thread held in registers (v[], plus its key and tag).  The pivot of each
// column is the largest packed key (float(|v|) bits, ~row): largest magnitude
// first, the lower tile row on ties (a NaN counts as zero).  Each wave finds its
// largest float(|v|) with the DPP wave reduction and its lowest row holding it
// with a ballot; one pass over the per-wave keys picks the pivot.  The pivot
// row and the row it displaces are exchanged through LDS; the elimination
// itself touches registers only.  Two barriers per column.  *sing receives
// 1 + the first column whose pivot was exactly zero.
template <typename T, int NT>
__device__ __forceinline__ void gepp_rows(T (&v)[PW], int &key, int &tag, int rows, int cols, int *sing)
{
    __shared__ unsigned long long wmax[2][NT / 64];   // per-wave maxima, alternating by column parity
    __shared__ T xrow[2][PW];                          // [0] pivot row, [1] the row it displaces
    __shared__ int xkt[2][2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nc = min(rows, cols);
#pragma unroll
    for (int c = 0; c < PW; ++c) {
        if (c < nc) {
            const bool valid = tid >= c && tid < rows;
            unsigned a = 0u;
            if (valid) {
                float f = (float)fabs((double)v[c]);
                if (!(f == f)) f = 0.f;
                a = __float_as_uint(f);
            }
            const unsigned amax = __ockl_wfred_max_u32(a);
            const unsigned long long ball = __ballot(valid && a == amax);
            if (lane == 0) {
                unsigned long long kk = 0ull;
                if (ball) {
                    const unsigned row = (unsigned)(wave * 64 + __ffsll((long long)ball) - 1);
                    kk = ((unsigned long long)amax << 32) | (unsigned long long)(0xFFFFFFFFu - row);
                }
                wmax[c & 1][wave] = kk;
            }
            __syncthreads();
            unsigned long long best = 0ull;
#pragma unroll
            for (int w = 0; w < NT / 64; ++w) best = wmax[c & 1][w] > best ? wmax[c & 1][w] : best;
            int p = (int)(0xFFFFFFFFu - (unsigned)(best & 0xFFFFFFFFull));
            if (best == 0ull || p < c || p >= rows) p = c;
            if (tid == p) {
#pragma unroll
                for (int e = 0; e < PW; ++e) xrow[0][e] = v[e];
                xkt[0][0] = key; xkt[0][1] = tag;
            }
            if (tid == c && p != c) {
#pragma unroll
                for (int e = 0; e < PW; ++e) xrow[1][e] = v[e];
                xkt[1][0] = key; xkt[1][1] = tag;
            }
            __syncthreads();
            if (p != c && (tid == c || tid == p)) {
                const int from = tid == c ? 0 : 1;
#pragma unroll
                for (int e = 0; e < PW; ++e) v[e] = xrow[from][e];
                key = xkt[from][0]; tag = xkt[from][1];
            }
            const T piv = xrow[0][c];
            if (piv == T(0)) {
                if (tid == 0 && *sing == 0) *sing = c + 1;
            } else if (tid > c && tid < rows) {
                const T l = v[c] * (T(1) / piv);
                v[c] = l;
#pragma unroll
                for (int e = c + 1; e < PW; ++e)
                    if (e < cols) v[e] -= l * xrow[0][e];
            }
        }
    }
}

// Level 0: each workgroup proposes PW candidate rows out of CH rows of the
// sub-panel A[s:n, s:s+pw] (thread t holds tile row t).  Output: original
// values + global row index.
template <typename T>
__global__ __launch_bounds__(CH) void k_tslu_local(const T *__restrict__ A, int64_t ld, int n, int s, int pw,
                                                   T *__restrict__ cval, int *__restrict__ cidx,
                                                   int *__restrict__ info)
{
    __shared__ int sing;
    const int g = blockIdx.x, tid = threadIdx.x;
    const int r0 = s + g * CH;
    const int rows = min(CH, n - r0);
    if (tid == 0) sing = 0;
    const bool has = tid < rows;
    T v[PW];
#pragma unroll
    for (int c = 0; c < PW; ++c) v[c] = (has && c < pw) ? A[(int64_t)(r0 + tid) * ld + s + c] : T(0);
    int key = has ? r0 + tid : INT_MAX, tag = tid;
    __syncthreads();
    gepp_rows<T, CH>(v, key, tag, rows, pw, &sing);
    // candidates: the first min(rows, pw) rows, original values re-read
    if (tid < PW) {
        const bool ok = tid < rows && tid < pw;
        T o[PW];
#pragma unroll
        for (int c = 0; c < PW; ++c) o[c] = (ok && c < pw) ? A[(int64_t)(r0 + tag) * ld + s + c] : T(0);
#pragma unroll
        for (int c = 0; c < PW; ++c) cval[((int64_t)g * PW + tid) * PW + c] = o[c];
        cidx[g * PW + tid] = ok ? key : INT_MAX;
    }
    (void)info;
}

// Merge MERGE_F candidate sets per workgroup (thread t holds candidate row t).
// When `final_lu` is set (one workgroup left) it also writes the factored
// pw x pw block and the winners.
template <typename T>
__global__ __launch_bounds__(MR) void k_tslu_merge(int nsets, int pw, const T *__restrict__ ival,
                                                   const int *__restrict__ iidx, T *__restrict__ oval,
                                                   int *__restrict__ oidx, T *__restrict__ final_lu,
                                                   int *__restrict__ info, int s)
{
    __shared__ int sing;
    const int g = blockIdx.x, tid = threadIdx.x;
    const int set0 = g * MERGE_F;
    const int nset = min(MERGE_F, nsets - set0);
    const int rows = nset * PW;
    if (tid == 0) sing = 0;
    const bool has = tid < rows;
    T v[PW];
#pragma unroll
    for (int c = 0; c < PW; ++c) v[c] = has ? ival[((int64_t)set0 * PW + tid) * PW + c] : T(0);
    int key = has ? iidx[set0 * PW + tid] : INT_MAX, tag = tid;
    __syncthreads();
    gepp_rows<T, MR>(v, key, tag, rows, pw, &sing);
    if (tid < PW) {
        T o[PW];
#pragma unroll
        for (int c = 0; c < PW; ++c) o[c] = ival[((int64_t)set0 * PW + tag) * PW + c];
#pragma unroll
        for (int c = 0; c < PW; ++c) {
            oval[((int64_t)g * PW + tid) * PW + c] = o[c];
            if (final_lu) final_lu[tid * PW + c] = v[c];
        }
        oidx[g * PW + tid] = key;
    }
    __syncthreads();
    if (final_lu && tid == 0 && sing && *info == 0) *info = s + sing;
}

// Winners (in pivot order) -> row map: position s+i <- row w_i, and the top
// rows [s, s+pw) that are not winners move to the positions the winners left
// outside the top block (paired in order).  One wave; map[0] = count, then
// (dst, src) pairs.  The same map is applied to the running permutation
// perm (perm[pos] = original row now at pos).
__global__ __launch_bounds__(64) void k_tslu_rowmap(const int *__restrict__ win, int s, int pw, int n,
                                                    int *__restrict__ map, int *__restrict__ perm)
{
    __shared__ int w[PW], intop[PW], vac[PW], dis[PW];
    __shared__ int nv, nd;
    const int lane = threadIdx.x;
    if (lane < PW) intop[lane] = 0;
    if (lane == 0) { nv = 0; nd = 0; }
    __syncthreads();
    if (lane < pw) {
        int wi = win[lane];
        if (wi < s || wi >= n) wi = s + lane;  // cannot happen for m >= pw
        w[lane] = wi;
        if (wi < s + pw) intop[wi - s] = 1;
    }
    __syncthreads();
    if (lane == 0) {
        for (int i = 0; i < pw; ++i)
            if (w[i] >= s + pw) vac[nv++] = w[i];
        for (int t = 0; t < pw; ++t)
            if (!intop[t]) dis[nd++] = s + t;
    }
    __syncthreads();
    // pairs: (s+i <- w_i) for w_i != s+i, then (vac_k <- dis_k)
    __shared__ int dst[2 * PW], src[2 * PW];
    __shared__ int cnt;
    if (lane == 0) {
        int c = 0;
        for (int i = 0; i < pw; ++i)
            if (w[i] != s + i) { dst[c] = s + i; src[c] = w[i]; ++c; }
        for (int k = 0; k < nv && k < nd; ++k) { dst[c] = vac[k]; src[c] = dis[k]; ++c; }
        cnt = c;
        map[0] = c;
    }
    __syncthreads();
    int pv = 0;
    if (lane < cnt) {
        map[1 + 2 * lane] = dst[lane];
        map[2 + 2 * lane] = src[lane];
        pv = perm[src[lane]];
    }
    __syncthreads();
    if (lane < cnt) perm[dst[lane]] = pv;
}

__global__ void k_iota(int *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

// Apply a row map (A[dst] <- A[src] for every pair) to every column: one lane
// per column, all loads issued before the stores.
template <typename T>
__global__ __launch_bounds__(256) void k_rowmap(T *__restrict__ A, int64_t ld, int64_t ncols, const int *__restrict__ map)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = map[0];
    if (c >= ncols) return;
    T v[2 * PW];
#pragma unroll
    for (int e = 0; e < 2 * PW; ++e)
        if (e < cnt) v[e] = A[(int64_t)map[2 + 2 * e] * ld + c];
#pragma unroll
    for (int e = 0; e < 2 * PW; ++e)
        if (e < cnt) A[(int64_t)map[1 + 2 * e] * ld + c] = v[e];
}

// Rows [s, s+pw) of the sub-panel <- factored block; rows below <- A21 U11^-1.
template <typename T>
__global__ __launch_bounds__(64) void k_tslu_l21(T *__restrict__ A, int64_t ld, int n, int s, int pw,
                                                 const T *__restrict__ lu)
{
    __shared__ T U[PW][PW + 1];
    for (int e = threadIdx.x; e < PW * PW; e += blockDim.x) U[e / PW][e % PW] = lu[e];
    __syncthreads();
    const int64_t r = s + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    T *row = A + r * ld + s;
    if (r < s + pw) {
        const int i = (int)(r - s);
        for (int c = 0; c < pw; ++c) row[c] = U[i][c];
        return;
    }
    T x[PW];
#pragma unroll
    for (int c = 0; c < PW; ++c) x[c] = c < pw ? row[c] : T(0);
#pragma unroll
    for (int c = 0; c < PW; ++c) {
        if (c < pw) {
            T acc = x[c];
#pragma unroll
            for (int q = 0; q < PW; ++q)
                if (q < c) acc -= x[q] * U[q][c];
            x[c] = U[c][c] != T(0) ? acc / U[c][c] : acc;
        }
    }
#pragma unroll
    for (int c = 0; c < PW; ++c)
        if (c < pw) row[c] = x[c];
}

__global__ void k_fill_i32(int *p, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Outer-block triangular update U12 := L11^-1 A12 (nb <= TRI_NB rows) as one MFMA
// GEMM: M = I - L11^-1 (strictly lower, written by k_trtri_llu), W = a copy of
// A12, then A12 -= M W.  Inverting the unit lower L11 explicitly is the
// inverted-diagonal-block trsm used by GPU LAPACK codes; with pivoted
// multipliers |l| <= 1 its entries stay modest.
constexpr int TRI_NB = 512;   // outer block of the LU (columns per trailing update)

// Column j = blockIdx.x of L^-1 by forward substitution: thread r owns x_r and
// keeps 32 entries of its own row of L in registers at a time; x_q is published
// in LDS the step it becomes final.  Computed in FP64 for both storage types.
template <typename T>
__global__ __launch_bounds__(TRI_NB) void k_trtri_llu(int nb, const T *__restrict__ L, int64_t ldl,
                                                      T *__restrict__ M, int64_t ldm)
{
    __shared__ double xs[TRI_NB];
    const int j = blockIdx.x, r = threadIdx.x;
    double x = r == j ? 1.0 : 0.0;
    for (int q0 = j & ~31; q0 < nb; q0 += 32) {
        double lr[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int q = q0 + u;
            lr[u] = (r < nb && q < r && q >= j) ? (double)L[(int64_t)r * ldl + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int q = q0 + u;
            if (q >= j && q < nb) {
                if (r == q) xs[q] = x;
                __syncthreads();
                if (r > q) x -= lr[u] * xs[q];
            }
        }
    }
    if (r < nb) M[(int64_t)r * ldm + j] = r > j ? (T)(-x) : T(0);
}

template <typename T>
__global__ void k_copy_rows(const T *__restrict__ src, int64_t lds, T *__restrict__ dst, int64_t ldd, int rows,
                            int64_t cols)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c < cols && r < rows) dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
}

static int64_t tri_ldw(int64_t n) { return (n + 3) & ~(int64_t)3; }

size_t fp_getrf_work_bytes(int64_t n, int elem_bytes)
{
    const int64_t g = (n + CH - 1) / CH;
    // two candidate buffers (values + indices) + final block + row map, then
    // the outer-block inverse M (TRI_NB x TRI_NB) and the A12 copy W (TRI_NB x ldw)
    const size_t tslu = (size_t)(2 * g * PW * PW) * elem_bytes + (size_t)(2 * g * PW) * sizeof(int) +
                        PW * PW * elem_bytes + (4 * PW + 16) * sizeof(int);
    const size_t head = (tslu + 255) & ~(size_t)255;
    return head + (size_t)TRI_NB * TRI_NB * elem_bytes + (size_t)TRI_NB * tri_ldw(n) * elem_bytes + 256;
}

template <typename T>
static int trsm_outer(fp_ctx *ctx, int nb, int64_t ncols, const T *L, int64_t ld, T *X, T *M, T *W, int64_t ldw)
{
    if (ncols <= 0 || nb <= 0) return FP_OK;
    {
        ProfScope ps(ctx, "trsm");
        hipLaunchKernelGGL(k_trtri_llu<T>, dim3((unsigned)nb), dim3(TRI_NB), 0, ctx->stream, nb, L, ld, M,
                           (int64_t)TRI_NB);
        hipLaunchKernelGGL(k_copy_rows<T>, dim3((unsigned)((ncols + 255) / 256), (unsigned)nb), dim3(256), 0,
                           ctx->stream, (const T *)X, ld, W, ldw, nb, ncols);
        FP_HIP(ctx, hipGetLastError());
    }
    return gemm_sub<T>(ctx, nb, ncols, nb, M, (int64_t)TRI_NB, W, ldw, X, ld);
}

// Factors P A = L U in place; perm[i] is the original row now at row i.
template <typename T>
int fp_getrf(fp_ctx *ctx, T *A, int64_t n, int64_t ld, int *perm, int *info_dev, void *work, size_t work_bytes)
{
    if (n <= 0) return FP_OK;
    if (work_bytes < fp_getrf_work_bytes(n, sizeof(T))) return fp_fail(ctx, FP_ERR_ARG, "getrf workspace too small");
    if (n > INT_MAX / 2) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "matrix too large for 32-bit pivots");
    const int64_t gmax = (n + CH - 1) / CH;
    char *w = (char *)work;
    T *cv0 = (T *)w; w += gmax * PW * PW * sizeof(T);
    T *cv1 = (T *)w; w += gmax * PW * PW * sizeof(T);
    int *ci0 = (int *)w; w += gmax * PW * sizeof(int);
    int *ci1 = (int *)w; w += gmax * PW * sizeof(int);
    T *flu = (T *)w; w += PW * PW * sizeof(T);
    int *rmap = (int *)w;
    const size_t tslu = (size_t)(w - (char *)work) + (4 * PW + 16) * sizeof(int);
    T *tri_m = (T *)((char *)work + ((tslu + 255) & ~(size_t)255));
    T *tri_w = tri_m + TRI_NB * TRI_NB;
    const int64_t ldw = tri_ldw(n);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, info_dev, (int64_t)1, 0);
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, perm, n);
    const int NB = TRI_NB;
    for (int64_t k = 0; k < n; k += NB) {
        const int nb = (int)std::min<int64_t>(NB, n - k);
        for (int64_t s = k; s < k + nb; s += PW) {
            const int pw = (int)std::min<int64_t>(PW, k + nb - s);
            const int m = (int)(n - s);
            {
            int g = (m + CH - 1) / CH;
            ProfScope ps(ctx, "panel_tslu");
            hipLaunchKernelGGL(k_tslu_local<T>, dim3(g), dim3(256), 0, ctx->stream, (const T *)A, ld, (int)n, (int)s,
                               pw, cv0, ci0, info_dev);
            T *iv = cv0, *ov = cv1;
            int *ii = ci0, *oi = ci1;
            for (;;) {
                const int go = (g + MERGE_F - 1) / MERGE_F;
                const bool last = go == 1;
                hipLaunchKernelGGL(k_tslu_merge<T>, dim3(go), dim3(512), 0, ctx->stream, g, pw, (const T *)iv,
                                   (const int *)ii, ov, oi, last ? flu : (T *)nullptr, info_dev, (int)s);
                std::swap(iv, ov);
                std::swap(ii, oi);
                g = go;
                if (last) break;
            }
            // ii now holds the pw winners in pivot order
            hipLaunchKernelGGL(k_tslu_rowmap, dim3(1), dim3(64), 0, ctx->stream, (const int *)ii, (int)s, pw, (int)n,
                               rmap, perm);
            hipLaunchKernelGGL(k_rowmap<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, A, ld, n,
                               (const int *)rmap);
            hipLaunchKernelGGL(k_tslu_l21<T>, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, ctx->stream, A, ld,
                               (int)n, (int)s, pw, (const T *)flu);
            FP_HIP(ctx, hipGetLastError());
            }
            // update the rest of the outer block
            const int64_t rest = k + nb - (s + pw);
            if (rest > 0) {
                int st = trsm_llu<T>(ctx, pw, rest, A + s * ld + s, ld, A + s * ld + s + pw, ld);
                if (st) return st;
                st = gemm_sub<T>(ctx, n - s - pw, rest, pw, A + (s + pw) * ld + s, ld, A + s * ld + s + pw, ld,
                                 A + (s + pw) * ld + s + pw, ld);
                if (st) return st;
            }
        }
        const int64_t right = n - k - nb;
        if (right > 0) {
            int st = trsm_outer<T>(ctx, nb, right, A + k * ld + k, ld, A + k * ld + k + nb, tri_m, tri_w, ldw);
            if (st) return st;
            st = gemm_sub<T>(ctx, right, right, nb, A + (k + nb) * ld + k, ld, A + k * ld + k + nb, ld,
                             A + (k + nb) * ld + k + nb, ld);
            if (st) return st;
        }
    }
    return FP_OK;
}
template int fp_getrf<double>(fp_ctx *, double *, int64_t, int64_t, int *, int *, void *, size_t);
template int fp_getrf<float>(fp_ctx *, float *, int64_t, int64_t, int *, int *, void *, size_t);

// ----------------------------------------------------------------------------
// Triangular solves with the LU factors, FP64 right-hand sides (n x k,
// column-major with ldx, i.e. each RHS contiguous).  Blocked by TB rows with
// explicit inverses of the TB x TB diagonal blocks (k_trtri_diag, once per
// call).  One launch per block step: every workgroup subtracts the solved
// block's contribution from the TB rows it owns, and the workgroup owning the
// next block -- final after that subtraction -- solves it with one small GEMV
// (y = D^-1 x) for the next launch.  No grid-wide synchronisation; the running
// right-hand side and the solution live in separate arrays.
constexpr int TB = 128;
constexpr int TRSV_ROWS = TB;   // rows of the update per workgroup: workgroup 0 owns the next block

// Inverse of diagonal block bi = blockIdx.y, column j = blockIdx.x: thread r
// owns x_r, 32 entries of its own row are staged in registers at a time.
// LOWER: unit lower L (the strictly lower part of LU).  Upper: U incl. diagonal.
template <typename T, bool LOWER>
__global__ __launch_bounds__(TB) void k_trtri_diag(const T *__restrict__ LU, int64_t ld, int64_t n,
                                                   double *__restrict__ D)
{
    __shared__ double xs[TB];
    const int j = blockIdx.x, r = threadIdx.x;
    const int64_t b0 = (int64_t)blockIdx.y * TB;
    const int bs = (int)min((int64_t)TB, n - b0);
    if (j >= bs) return;   // uniform per workgroup
    const T *row = LU + (b0 + r) * ld + b0;
    double x = r == j ? 1.0 : 0.0;
    if (LOWER) {
        for (int q0 = j & ~31; q0 < bs; q0 += 32) {
            double lr[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int q = q0 + u;
                lr[u] = (r < bs && q < r && q >= j) ? (double)row[q] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int q = q0 + u;
                if (q >= j && q < bs) {
                    if (r == q) xs[q] = x;
                    __syncthreads();
                    if (r > q) x -= lr[u] * xs[q];
                }
            }
        }
    } else {
        // x = U^-1 e_j: rows q = j, j-1, ..., 0 become final in turn
        for (int q1 = j | 31; q1 >= 0; q1 -= 32) {
            double ur[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int q = q1 - u;
                ur[u] = (r < bs && q >= r && q <= j && q >= 0) ? (double)row[q] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int q = q1 - u;
                if (q <= j && q >= 0) {
                    if (r == q) {
                        x /= ur[u];
                        xs[q] = x;
                    }
                    __syncthreads();
                    if (r < q) x -= ur[u] * xs[q];
                }
            }
        }
    }
    if (r < bs) D[((int64_t)blockIdx.y * TB + r) * TB + j] = x;   // row-major block
}

// Row t of D in[0:bs] for a row-major TB x TB block inverse D; 256 threads,
// threads t and t+128 each take half of the columns.  dv[] holds this thread's
// half of row t, loaded ahead by load_block_row as 16-byte loads from one base
// (entries at or beyond bs are never written by k_trtri_diag and are skipped by
// block_gemv_row); the sum is meaningful for threads t < bs of the first half.
__device__ __forceinline__ void load_block_row(const double *__restrict__ D, int tid, double (&dv)[TB / 2])
{
    const int t = tid & (TB - 1), cb = (tid >> 7) * (TB / 2);
    const double2 *p = reinterpret_cast<const double2 *>(D + (int64_t)t * TB + cb);
#pragma unroll
    for (int c = 0; c < TB / 4; ++c) {
        const double2 w = p[c];
        dv[2 * c] = w.x;
        dv[2 * c + 1] = w.y;
    }
}

__device__ __forceinline__ double block_gemv_row(const double (&dv)[TB / 2], const double *in, int bs, int tid,
                                                 double *part)
{
    const int t = tid & (TB - 1), h = tid >> 7, cb = h * (TB / 2);
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < TB / 2; ++c)
        if (cb + c < bs) acc += dv[c] * in[cb + c];
    if (h == 1) part[t] = acc;
    __syncthreads();
    return acc + part[t];
}

// First block of a sweep: out[b0:b0+bs] = D x[b0:b0+bs] (one workgroup per RHS).
__global__ __launch_bounds__(256) void k_trsv_first(const double *__restrict__ D, int64_t b0, int bs,
                                                    const double *__restrict__ x, int64_t ldx,
                                                    double *__restrict__ out, int64_t ldo)
{
    __shared__ double xb[TB], part[TB];
    const int tid = threadIdx.x;
    double dv[TB / 2];
    load_block_row(D, tid, dv);
    const double *xv = x + (int64_t)blockIdx.y * ldx + b0;
    if (tid < bs) xb[tid] = xv[tid];
    __syncthreads();
    const double v = block_gemv_row(dv, xb, bs, tid, part);
    if (tid < bs) out[(int64_t)blockIdx.y * ldo + b0 + tid] = v;
}

// Step for the solved block [b0, b0+bs) of right-hand side blockIdx.y (its
// solution sol[b0:b0+bs] written by the previous launch): x[r] -= LU[r, b0:b0+bs]
// . sol[b0:b0+bs] for the TB rows this workgroup owns -- rows [rlo + w*TB, ...)
// going down (L), rows [.., rhi - w*TB) going up (U).  Workgroup 0 owns exactly
// the next block [nb0, nb0+nbs) and, when Dn is given, writes its solution
// sol[nb0:nb0+nbs] = Dn x[nb0:nb0+nbs].  Every global load (LU rows, x rows,
// the solved block, workgroup 0's rows of Dn) is issued before the first use.
template <typename T, bool DOWN>
__global__ __launch_bounds__(256) void k_trsv_step(const T *__restrict__ LU, int64_t ld, int64_t b0, int bs,
                                                   double *__restrict__ x, int64_t ldx, double *__restrict__ sol,
                                                   int64_t lds, int64_t rlo, int64_t rhi,
                                                   const double *__restrict__ Dn, int64_t nb0, int nbs)
{
    __shared__ double yb[TB], xn[TB], part[TB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *xv = x + (int64_t)blockIdx.y * ldx;
    double *sv = sol + (int64_t)blockIdx.y * lds;
    int64_t r0, r1;
    if (DOWN) {
        r0 = rlo + (int64_t)blockIdx.x * TRSV_ROWS;
        r1 = min(rhi, r0 + TRSV_ROWS);
    } else {
        r1 = rhi - (int64_t)blockIdx.x * TRSV_ROWS;
        r0 = max(rlo, r1 - TRSV_ROWS);
    }
    const bool lead = blockIdx.x == 0 && Dn != nullptr;   // uniform per workgroup
    const int sub = lane & 15, c0 = sub * 8;
    constexpr int NIT = TRSV_ROWS / 16;
    double v[NIT][8], xr[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int64_t r = r0 + it * 16 + wave * 4 + (lane >> 4);
        const T *lr = LU + r * ld + b0 + c0;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[it][e] = (r < r1 && c0 + e < bs) ? (double)lr[e] : 0.0;
        xr[it] = (sub == 0 && r < r1) ? xv[r] : 0.0;
    }
    double dv[TB / 2];
    if (lead) load_block_row(Dn, tid, dv);
    if (tid < bs) yb[tid] = sv[b0 + tid];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int64_t r = r0 + it * 16 + wave * 4 + (lane >> 4);
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < 8; ++e) acc += v[it][e] * (c0 + e < bs ? yb[c0 + e] : 0.0);
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (sub == 0 && r < r1) {
            const double nv = xr[it] - acc;
            xv[r] = nv;
            if (lead) xn[r - r0] = nv;
        }
    }
    if (!lead) return;
    __syncthreads();
    const double vn = block_gemv_row(dv, xn, nbs, tid, part);
    if (tid < nbs) sv[nb0 + tid] = vn;
}

// Sync-free sweep: the whole triangular solve of one right-hand side in one
// launch.  One workgroup per TB-row block and right-hand side (blockIdx.y); a
// workgroup takes its block from a per-RHS ticket counter in solve order
// (first block first), so every block it waits for belongs to a workgroup
// that is already running -- no reliance on the dispatch order, no deadlock
// however few workgroups are resident.  The solution array starts filled with
// a sentinel (a signalling-NaN bit pattern that no arithmetic produces) and
// each solved entry is published by an agent-scope atomic store: the entries
// are their own ready flags, so a waiting workgroup polls the 128 entries it
// needs (agent-scope atomic loads, coherent across the XCDs' L2s) and needs
// no flag, fence or cache write-back.  The workgroup owning block b applies
// the solved blocks' contributions one TB x TB tile at a time (each tile's
// loads are issued before the wait), each thread keeping partial sums over its
// 8 columns of every tile; then one 16-lane reduction and the block solve
// y_b = D_b x_b with the explicit inverse of the diagonal block (its rows
// loaded ahead of the last wait).  DOWN: forward sweep with the unit lower L
// (tiles left of the block); else the backward sweep with U (tiles to the
// right).  rhs is not modified; tickets (k per sweep) are zeroed before the
// launch.  A wait that exceeds its spin limit takes NaN for the entry, so
// every workgroup ends and the caller's residual check reports the failure.
constexpr int TRSV_SPIN_LIMIT = 1 << 20;
constexpr unsigned long long TRSV_PENDING = 0xFFF4F00DF00DF00Dull;

__device__ __forceinline__ double trsv_pending() { return __longlong_as_double((long long)TRSV_PENDING); }

// One thread's 8 x 8 share of the TB x TB tile at rows r0.., columns cb..:
// rows it * 16 + wave * 4 + lane / 16, columns c0..c0+7, as stored (no widening).
// Full tiles (VEC: 16-byte aligned groups of 8) by vector loads; ragged tiles by
// clamped addresses (every load in bounds, none behind a branch) with the
// entries outside the tile zeroed by a product with 0/1 (not a select, which
// the compiler turns back into a branch around the load).
template <typename T, bool VEC>
__device__ __forceinline__ void load_tile(const T *__restrict__ LU, int64_t ld, int64_t r0, int bs, int64_t cb, int cs,
                                          int wave, int lane, int c0, T (&v)[TB / 16][8])
{
    constexpr int NIT = TB / 16;
    if (VEC && bs == TB && cs == TB) {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const T *p = LU + (r0 + it * 16 + wave * 4 + (lane >> 4)) * ld + cb + c0;
            if constexpr (sizeof(T) == 4) {
                const float4 x = *reinterpret_cast<const float4 *>(p), y = *reinterpret_cast<const float4 *>(p + 4);
                v[it][0] = x.x; v[it][1] = x.y; v[it][2] = x.z; v[it][3] = x.w;
                v[it][4] = y.x; v[it][5] = y.y; v[it][6] = y.z; v[it][7] = y.w;
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double2 x = reinterpret_cast<const double2 *>(p)[q];
                    v[it][2 * q] = x.x;
                    v[it][2 * q + 1] = x.y;
                }
            }
        }
    } else {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int rr = it * 16 + wave * 4 + (lane >> 4);
            const T *lr = LU + (r0 + min(rr, bs - 1)) * ld + cb;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[it][e] = lr[min(c0 + e, cs - 1)] * ((rr < bs && c0 + e < cs) ? T(1) : T(0));
        }
    }
}

// VEC: ld and LU allow 16-byte loads of 8-entry groups (ld a multiple of 8, LU 16-byte aligned)
template <typename T, bool DOWN, bool VEC>
__global__ __launch_bounds__(256) void k_trsv_sweep(const T *__restrict__ LU, int64_t ld, int64_t n, int nblk,
                                                    const double *__restrict__ D, const double *__restrict__ rhs,
                                                    int64_t ldr, double *__restrict__ sol, int64_t lds,
                                                    int *__restrict__ tickets)
{
    __shared__ int s_t;
    __shared__ double yb[TB], xn[TB], part[TB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = blockIdx.y;
    if (tid == 0) s_t = atomicAdd(tickets + col, 1);
    __syncthreads();
    const int t = s_t;
    const int bi = DOWN ? t : nblk - 1 - t;
    const int64_t r0 = (int64_t)bi * TB;
    const int bs = (int)min((int64_t)TB, n - r0);
    const int sub = lane & 15, c0 = sub * 8;
    constexpr int NIT = TB / 16;
    const double *rv = rhs + (int64_t)col * ldr;
    double *sv = sol + (int64_t)col * lds;
    auto blk = [&](int s) { return DOWN ? s : nblk - 1 - s; };   // the s-th block solved
    double acc[NIT], dv[TB / 2];
#pragma unroll
    for (int it = 0; it < NIT; ++it) acc[it] = 0.0;
    const double *Db = D + (int64_t)bi * TB * TB;
    // FP32 factors: the inverse's rows are loaded ahead of the last wait; FP64 tiles leave no room
    // for them in registers beside two tiles in flight, so they are loaded after the last tile
    constexpr bool DV_AHEAD = sizeof(T) == 4;
    if (DV_AHEAD && t == 0) load_block_row(Db, tid, dv);
    // two tiles in flight: tile s+1 is requested before the wait for block s
    T cur[NIT][8], nxt[NIT][8];
    if (t > 0) {
        const int64_t cb = (int64_t)blk(0) * TB;
        load_tile<T, VEC>(LU, ld, r0, bs, cb, (int)min((int64_t)TB, n - cb), wave, lane, c0, cur);
    }
    for (int s = 0; s < t; ++s) {
        const int64_t cb = (int64_t)blk(s) * TB;
        const int cs = (int)min((int64_t)TB, n - cb);
        if (s + 1 < t) {
            const int64_t cn = (int64_t)blk(s + 1) * TB;
            load_tile<T, VEC>(LU, ld, r0, bs, cn, (int)min((int64_t)TB, n - cn), wave, lane, c0, nxt);
        } else if (DV_AHEAD) {
            load_block_row(Db, tid, dv);   // the block's own inverse, ahead of the last wait
        }
        if (tid < TB) {
            double y = 0.0;
            if (tid < cs) {
                int spins = 0;
                for (;;) {
                    y = __hip_atomic_load(sv + cb + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((unsigned long long)__double_as_longlong(y) != TRSV_PENDING) break;
                    if (++spins > TRSV_SPIN_LIMIT) {
                        y = __builtin_nan("");
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            yb[tid] = y;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < NIT; ++it)
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[it] += (double)cur[it][e] * yb[c0 + e];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < NIT; ++it)
#pragma unroll
            for (int e = 0; e < 8; ++e) cur[it][e] = nxt[it][e];
    }
    if (!DV_AHEAD) load_block_row(Db, tid, dv);
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) acc[it] += __shfl_xor(acc[it], o);
        const int rr = it * 16 + wave * 4 + (lane >> 4);
        if (sub == 0 && rr < bs) xn[rr] = rv[r0 + rr] - acc[it];
    }
    __syncthreads();
    const double y = block_gemv_row(dv, xn, bs, tid, part);
    if (tid < bs) __hip_atomic_store(sv + r0 + tid, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// x[c * ldx + i] = the pending sentinel, i < n, c < k
__global__ void k_trsv_pending(double *__restrict__ x, int64_t n, int64_t ldx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[(int64_t)blockIdx.y * ldx + i] = trsv_pending();
}

__global__ void k_zero_i32(int *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_permute(const int *__restrict__ perm, int64_t n, const double *__restrict__ src, int64_t lds,
                          double *__restrict__ dst, int64_t ldd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[(int64_t)blockIdx.y * ldd + i] = src[(int64_t)blockIdx.y * lds + perm[i]];
}

// running right-hand sides (2 n k), the diagonal-block inverses of L and U,
// and the sweeps' tickets (2 k ints; the former per-block flags' room is kept)
size_t fp_getrs_work_bytes(int64_t n, int64_t k)
{
    const int64_t nblk = (n + TB - 1) / TB;
    return (size_t)2 * n * k * sizeof(double) + (size_t)2 * nblk * TB * TB * sizeof(double) +
           (size_t)2 * (k + k * nblk) * sizeof(int) + 256;
}

// FLOWPANEL_TRSV=steps: one launch per block step (k_trsv_step) instead of the sync-free sweeps
static bool trsv_steps()
{
    static const bool v = [] {
        const char *e = getenv("FLOWPANEL_TRSV");
        return e && !strcmp(e, "steps");
    }();
    return v;
}

// `ipiv` here is the row permutation perm (row i of P*A is row perm[i] of A).
template <typename T>
int fp_getrs(fp_ctx *ctx, const T *LU, int64_t n, int64_t ld, const int *perm, double *x, int64_t k, int64_t ldx,
             void *work, size_t work_bytes, bool reuse_diag)
{
    if (n <= 0 || k <= 0) return FP_OK;
    if (work_bytes < fp_getrs_work_bytes(n, k)) return fp_fail(ctx, FP_ERR_ARG, "getrs workspace too small");
    if (k > 65535) return fp_fail(ctx, FP_ERR_UNSUPPORTED, "getrs: too many right-hand sides in one call");
    ProfScope ps(ctx, "getrs");
    const int64_t nblk = (n + TB - 1) / TB;
    auto bsz = [&](int64_t bi) { return (int)std::min<int64_t>(TB, n - bi * TB); };
    double *b = (double *)work;          // P rhs: running right-hand side of the forward sweep
    double *y = b + n * k;               // L^-1 P rhs: running right-hand side of the backward sweep
    double *Dl = y + n * k, *Du = Dl + nblk * TB * TB;
    hipLaunchKernelGGL(k_permute, dim3((unsigned)((n + 255) / 256), (unsigned)k), dim3(256), 0, ctx->stream, perm, n,
                       (const double *)x, ldx, b, n);
    if (!reuse_diag) {
        hipLaunchKernelGGL((k_trtri_diag<T, true>), dim3(TB, (unsigned)nblk), dim3(TB), 0, ctx->stream, LU, ld, n, Dl);
        hipLaunchKernelGGL((k_trtri_diag<T, false>), dim3(TB, (unsigned)nblk), dim3(TB), 0, ctx->stream, LU, ld, n, Du);
    }
    if (!trsv_steps()) {
        int *sync = (int *)(Du + nblk * TB * TB);
        hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((2 * k + 255) / 256)), dim3(256), 0, ctx->stream, sync, 2 * k);
        const dim3 gp((unsigned)((n + 255) / 256), (unsigned)k);
        hipLaunchKernelGGL(k_trsv_pending, gp, dim3(256), 0, ctx->stream, y, n, n);
        hipLaunchKernelGGL(k_trsv_pending, gp, dim3(256), 0, ctx->stream, x, n, ldx);
        // forward: y = L^-1 b, then backward: x = U^-1 y
        const dim3 g((unsigned)nblk, (unsigned)k);
        int *sync2 = sync + k;
        if (ld % 8 == 0 && ((uintptr_t)LU & 15) == 0) {
            hipLaunchKernelGGL((k_trsv_sweep<T, true, true>), g, dim3(256), 0, ctx->stream, LU, ld, n, (int)nblk,
                               (const double *)Dl, (const double *)b, n, y, n, sync);
            hipLaunchKernelGGL((k_trsv_sweep<T, false, true>), g, dim3(256), 0, ctx->stream, LU, ld, n, (int)nblk,
                               (const double *)Du, (const double *)y, n, x, ldx, sync2);
        } else {
            hipLaunchKernelGGL((k_trsv_sweep<T, true, false>), g, dim3(256), 0, ctx->stream, LU, ld, n, (int)nblk,
                               (const double *)Dl, (const double *)b, n, y, n, sync);
            hipLaunchKernelGGL((k_trsv_sweep<T, false, false>), g, dim3(256), 0, ctx->stream, LU, ld, n, (int)nblk,
                               (const double *)Du, (const double *)y, n, x, ldx, sync2);
        }
        FP_HIP(ctx, hipGetLastError());
        return FP_OK;
    }
    // forward: y = L^-1 b
    hipLaunchKernelGGL(k_trsv_first, dim3(1, (unsigned)k), dim3(256), 0, ctx->stream, (const double *)Dl, (int64_t)0,
                       bsz(0), (const double *)b, n, y, n);
    for (int64_t bi = 0; bi + 1 < nblk; ++bi) {
        const int64_t b0 = bi * TB, rlo = b0 + TB;
        const int64_t g = (n - rlo + TRSV_ROWS - 1) / TRSV_ROWS;
        hipLaunchKernelGGL((k_trsv_step<T, true>), dim3((unsigned)g, (unsigned)k), dim3(256), 0, ctx->stream, LU, ld,
                           b0, TB, b, n, y, n, rlo, n, (const double *)(Dl + (bi + 1) * TB * TB), rlo, bsz(bi + 1));
    }
    // backward: x = U^-1 y
    hipLaunchKernelGGL(k_trsv_first, dim3(1, (unsigned)k), dim3(256), 0, ctx->stream,
                       (const double *)(Du + (nblk - 1) * TB * TB), (nblk - 1) * TB, bsz(nblk - 1), (const double *)y,
                       n, x, ldx);
    for (int64_t bi = nblk - 1; bi >= 1; --bi) {
        const int64_t b0 = bi * TB;
        const int64_t g = (b0 + TRSV_ROWS - 1) / TRSV_ROWS;
        hipLaunchKernelGGL((k_trsv_step<T, false>), dim3((unsigned)g, (unsigned)k), dim3(256), 0, ctx->stream, LU,
                           ld, b0, bsz(bi), y, n, x, ldx, (int64_t)0, b0, (const double *)(Du + (bi - 1) * TB * TB),
                           b0 - TB, TB);
    }
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
template int fp_getrs<double>(fp_ctx *, const double *, int64_t, int64_t, const int *, double *, int64_t, int64_t,
                              void *, size_t, bool);
template int fp_getrs<float>(fp_ctx *, const float *, int64_t, int64_t, const int *, double *, int64_t, int64_t,
                             void *, size_t, bool);

// ----------------------------------------------------------------------------
// y = alpha A x + beta y (row-major A, FP64 vectors).  One wave per 4 rows:
// every lane loads 4 consecutive x values once and reuses them for the 4 rows,
// A is read with 16-byte (float4) or 32-byte (2 x double2) vector loads.
// 16-byte loads of 4 matrix entries; `stream` marks them non-temporal (each
// entry of a Krylov matvec is read once per launch and the matrix is far
// larger than the caches: measured 0.558 against 0.616 ms for the 30k FP32 AIC).
typedef float fp_f4v __attribute__((ext_vector_type(4)));
typedef double fp_d2v __attribute__((ext_vector_type(2)));
template <typename T> struct Vec4;
template <> struct Vec4<float> {
    static __device__ __forceinline__ void load(const float *p, double *v)
    {
        const float4 q = *reinterpret_cast<const float4 *>(p);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    }
    // raw non-temporal load kept in the storage type until the dot product
    typedef fp_f4v raw;
    static __device__ __forceinline__ raw stream(const float *p)
    {
        return __builtin_nontemporal_load(reinterpret_cast<const fp_f4v *>(p));
    }
    static __device__ __forceinline__ double dot(const raw &q, const double2 &xa, const double2 &xb)
    {
        return (double)q.x * xa.x + (double)q.y * xa.y + (double)q.z * xb.x + (double)q.w * xb.y;
    }
};
template <> struct Vec4<double> {
    static __device__ __forceinline__ void load(const double *p, double *v)
    {
        const double2 a = *reinterpret_cast<const double2 *>(p);
        const double2 b = *reinterpret_cast<const double2 *>(p + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
    // FP64 rows keep ordinary loads: non-temporal 16-byte loads measured slower
    // here (2.13 against 1.29 ms for the 30k FP64 residual GEMV)
    struct raw { fp_d2v a, b; };
    static __device__ __forceinline__ raw stream(const double *p)
    {
        raw q;
        q.a = *reinterpret_cast<const fp_d2v *>(p);
        q.b = *reinterpret_cast<const fp_d2v *>(p + 2);
        return q;
    }
    static __device__ __forceinline__ double dot(const raw &q, const double2 &xa, const double2 &xb)
    {
        return q.a.x * xa.x + q.a.y * xa.y + q.b.x * xb.x + q.b.y * xb.y;
    }
};

// Rows per wave R of the Krylov GEMV: each x load is reused by R rows.  At
// small n two rows keep more waves in flight (measured on gfx950 at n = 10,000:
// 0.076 ms with R = 2 against 0.110 ms with R = 4); at n = 30,000 there are
// waves enough and, with non-temporal loads, R = 8 is fastest (back-to-back
// launches: 0.532 ms against 0.540 for R = 6, 0.561 for R = 4, 0.609 for R = 2;
// inside the Krylov loop 0.59 ms, 0.71 -> 0.77 of 8 TB/s with the 0.636 ms of
// plain loads).
constexpr int GEMVM_R = 4;  // rows per wave of the multi-vector GEMV

template <typename T, bool XVEC, int GEMV_R>
__global__ __launch_bounds__(256) void k_gemv(const T *__restrict__ A, int64_t n, int64_t ld,
                                              const double *__restrict__ x, double *__restrict__ y, double alpha,
                                              double beta)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * GEMV_R;
    if (r0 >= n) return;
    const int nr = (int)min((int64_t)GEMV_R, n - r0);
    double acc[GEMV_R];
#pragma unroll
    for (int r = 0; r < GEMV_R; ++r) acc[r] = 0;
    const int64_t n4 = (n / 4) * 4;  // ld is a multiple of 64, rows are 256-byte aligned
    for (int64_t c = 4 * lane; c < n4; c += 256) {
        double2 xa, xb;
        if (XVEC) {
            xa = *reinterpret_cast<const double2 *>(x + c);
            xb = *reinterpret_cast<const double2 *>(x + c + 2);
        } else {
            xa.x = x[c]; xa.y = x[c + 1]; xb.x = x[c + 2]; xb.y = x[c + 3];
        }
        // rows past the end re-read row r0 (in bounds, result discarded): no branches around the loads
        typename Vec4<T>::raw v[GEMV_R];
#pragma unroll
        for (int r = 0; r < GEMV_R; ++r) v[r] = Vec4<T>::stream(A + (r0 + (r < nr ? r : 0)) * ld + c);
#pragma unroll
        for (int r = 0; r < GEMV_R; ++r) acc[r] += Vec4<T>::dot(v[r], xa, xb);
    }
    for (int64_t c = n4 + lane; c < n; c += 64)
        for (int r = 0; r < nr; ++r) acc[r] += (double)A[(r0 + r) * ld + c] * x[c];
#pragma unroll
    for (int r = 0; r < GEMV_R; ++r) {
        double a = acc[r];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
        if (lane == 0 && r < nr) y[r0 + r] = alpha * a + (beta == 0.0 ? 0.0 : beta * y[r0 + r]);
    }
}

template <typename T, int R>
static int gemv_launch(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, const double *x, double *y, double alpha,
                       double beta)
{
    const dim3 grid((unsigned)((n + 4 * R - 1) / (4 * R)));
    if (((uintptr_t)x & 15) == 0 && ((uintptr_t)A & 15) == 0)
        hipLaunchKernelGGL((k_gemv<T, true, R>), grid, dim3(256), 0, ctx->stream, A, n, ld, x, y, alpha, beta);
    else if (((uintptr_t)A & 15) == 0)
        hipLaunchKernelGGL((k_gemv<T, false, R>), grid, dim3(256), 0, ctx->stream, A, n, ld, x, y, alpha, beta);
    else
        return fp_fail(ctx, FP_ERR_ARG, "gemv needs a 16-byte aligned matrix");
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

template <typename T>
int fp_gemv(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, const double *x, double *y, double alpha, double beta)
{
    if (n <= 0) return FP_OK;
    if (ld % 4) return fp_fail(ctx, FP_ERR_ARG, "gemv needs ld % 4 == 0");
    ProfScope ps(ctx, sizeof(T) == 8 ? "gemv_f64" : "gemv_f32");
    if (n <= 20000) return gemv_launch<T, 2>(ctx, A, n, ld, x, y, alpha, beta);
    // FP64 rows are twice the registers per row: 4 rows (1.29 ms at 30k against 1.88 with 8)
    return sizeof(T) == 4 ? gemv_launch<T, 8>(ctx, A, n, ld, x, y, alpha, beta)
                          : gemv_launch<T, 4>(ctx, A, n, ld, x, y, alpha, beta);
}
template int fp_gemv<double>(fp_ctx *, const double *, int64_t, int64_t, const double *, double *, double, double);
template int fp_gemv<float>(fp_ctx *, const float *, int64_t, int64_t, const double *, double *, double, double);

// Y[:, c] = A X[:, c] for c < k (X, Y column-major with leading dimensions
// ldx, ldy): GEMVM_NV vectors per grid row, so A is streamed ceil(k/GEMVM_NV)
// times instead of k times.
constexpr int GEMVM_NV = 8;

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_multi(const T *__restrict__ A, int64_t n, int64_t ld,
                                                    const double *__restrict__ X, int64_t ldx, int k,
                                                    double *__restrict__ Y, int64_t ldy)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * GEMVM_R;
    if (r0 >= n) return;
    const int nr = (int)min((int64_t)GEMVM_R, n - r0);
    const int c0 = blockIdx.y * GEMVM_NV;
    const int nv = min(GEMVM_NV, k - c0);
    double acc[GEMVM_R][GEMVM_NV];
#pragma unroll
    for (int r = 0; r < GEMVM_R; ++r)
#pragma unroll
        for (int q = 0; q < GEMVM_NV; ++q) acc[r][q] = 0;
    const int64_t n4 = (n / 4) * 4;
    for (int64_t c = 4 * lane; c < n4; c += 256) {
        double v[GEMVM_R][4];
#pragma unroll
        for (int r = 0; r < GEMVM_R; ++r) {
            if (r < nr) Vec4<T>::load(A + (r0 + r) * ld + c, v[r]);
            else v[r][0] = v[r][1] = v[r][2] = v[r][3] = 0;
        }
#pragma unroll
        for (int q = 0; q < GEMVM_NV; ++q) {
            if (q < nv) {
                const double *x = X + (int64_t)(c0 + q) * ldx + c;
                const double x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3];
#pragma unroll
                for (int r = 0; r < GEMVM_R; ++r) acc[r][q] += v[r][0] * x0 + v[r][1] * x1 + v[r][2] * x2 + v[r][3] * x3;
            }
        }
    }
    for (int64_t c = n4 + lane; c < n; c += 64) {
#pragma unroll
        for (int r = 0; r < GEMVM_R; ++r) {
            if (r < nr) {
                const double a = (double)A[(r0 + r) * ld + c];
#pragma unroll
                for (int q = 0; q < GEMVM_NV; ++q)
                    if (q < nv) acc[r][q] += a * X[(int64_t)(c0 + q) * ldx + c];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < GEMVM_R; ++r)
#pragma unroll
        for (int q = 0; q < GEMVM_NV; ++q) {
            double a = acc[r][q];
            for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
            if (lane == 0 && r < nr && q < nv) Y[(int64_t)(c0 + q) * ldy + r0 + r] = a;
        }
}

template <typename T>
int fp_gemv_multi(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, const double *X, int64_t ldx, int64_t k, double *Y,
                  int64_t ldy)
{
    if (n <= 0 || k <= 0) return FP_OK;
    if (ld % 4 || ((uintptr_t)A & 15)) return fp_fail(ctx, FP_ERR_ARG, "gemv needs an aligned matrix");
    ProfScope ps(ctx, sizeof(T) == 8 ? "gemv_multi_f64" : "gemv_multi_f32");
    const dim3 grid((unsigned)((n + 4 * GEMVM_R - 1) / (4 * GEMVM_R)), (unsigned)((k + GEMVM_NV - 1) / GEMVM_NV));
    hipLaunchKernelGGL(k_gemv_multi<T>, grid, dim3(256), 0, ctx->stream, A, n, ld, X, ldx, (int)k, Y, ldy);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
template int fp_gemv_multi<double>(fp_ctx *, const double *, int64_t, int64_t, const double *, int64_t, int64_t,
                                   double *, int64_t);
template int fp_gemv_multi<float>(fp_ctx *, const float *, int64_t, int64_t, const double *, int64_t, int64_t,
                                  double *, int64_t);

__global__ void k_fill(double *__restrict__ x, int64_t n, double v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

int fp_fill(fp_ctx *ctx, double *x, int64_t n, double v)
{
    if (n <= 0) return FP_OK;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, x, n, v);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

__global__ void k_fill_f32(float *__restrict__ x, int64_t n, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

int fp_fill_f32(fp_ctx *ctx, float *x, int64_t n, float v)
{
    if (n <= 0) return FP_OK;
    const int64_t g = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)g), dim3(256), 0, ctx->stream, x, n, v);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// column norms^2 of a column-major n x k block: out[c] = sum_i R[c*ld + i]^2
__global__ __launch_bounds__(256) void k_colnorm2(const double *__restrict__ R, int64_t n, int64_t ld,
                                                  double *__restrict__ out)
{
    __shared__ double red[4];
    const int c = blockIdx.x;
    double s = 0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double v = R[(int64_t)c * ld + i];
        s += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[c] = red[0] + red[1] + red[2] + red[3];
}

int fp_colnorm2(fp_ctx *ctx, const double *R, int64_t n, int64_t ld, int64_t k, double *out)
{
    if (k <= 0) return FP_OK;
    hipLaunchKernelGGL(k_colnorm2, dim3((unsigned)k), dim3(256), 0, ctx->stream, R, n, ld, out);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// Block-Jacobi inverses by Gauss-Jordan with partial pivoting, the block held
// in registers (TG x TG threads, thread (tr, tc) holding rows tr + TG i and
// columns tc + TG j, 8 x 8 FP32 values: blocks up to 8 TG rows).  Rows are never swapped: the pivot of step k is the unused
// row p_k with the largest |a(p, k)| (lowest row on ties, NaN counted as 0),
// and the in-place result W satisfies inverse(pinv(r), p_c) = W(r, c), so the
// final store scatters through the two index maps.  Per step: column k and
// the scaled pivot row go through LDS, everything else stays in registers.

//
// Block b of the grid inverts the square sub-block of hs rows at offset
// (off, off) of its diagonal block (hs < 0: the whole block, bs = min(block,
// n - b * block) rows).  Source element (r, c) of block b is
// A[b * a_bstride + (off + r) * a_ld + off + c]; the result goes to the same
// place of inv (i_bstride, i_ld).  A and inv may be the same buffer: every
// element is read into registers before the first barrier and written after
// the last one.
template <typename T, int TG>
__global__ __launch_bounds__(TG * TG) void k_bj_gj(const T *A, int64_t a_bstride, int64_t a_ld, float *inv,
                                                int64_t i_bstride, int64_t i_ld, int64_t n, int block, int off,
                                                int hs, int *__restrict__ info)
{
    constexpr int NR = 8 * TG, NW = TG * TG / 64;
    __shared__ float colk[2][NR], prow[NR];
    __shared__ int used[NR], prow_of[NR], step_of[NR];
    __shared__ float wv[NW];
    __shared__ int wr[NW];
    const int b = blockIdx.x;
    const int64_t b0 = (int64_t)b * block;
    const int bs = hs < 0 ? (int)min((int64_t)block, n - b0) : hs;
    const T *src = A + (int64_t)b * a_bstride + off * a_ld + off;
    const int tid = threadIdx.x, tr = tid / TG, tc = tid % TG, lane = tid & 63, wave = tid >> 6;
    float w[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int r = tr + TG * i, c = tc + TG * j;
            w[i][j] = (r < bs && c < bs) ? (float)src[r * a_ld + c] : 0.f;
        }
    for (int r = tid; r < NR; r += TG * TG) {
        used[r] = 0;
        colk[0][r] = 0.f;
        colk[1][r] = 0.f;
        prow[r] = 0.f;
    }
    __syncthreads();
    // Two barriers per step.  Column k goes to colk[k & 1] together with the
    // pivot candidates of the TG threads holding it, so that writing column
    // k + 1 never races with the update of step k still reading column k.
    // The register column kb of column k is a compile-time index (unrolled kb
    // loop); only the pivot row p / TG is selected at run time.
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        if (kb * TG >= bs) break;
        for (int kk = 0; kk < TG; ++kk) {
            const int k = kb * TG + kk;
            if (k >= bs) break;
            float *ck = colk[k & 1];
            float v = -1.f;
            int rr = NR;
            if (tc == kk) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float val = w[i][kb];
                    const int r = tr + TG * i;
                    if (r < bs) {
                        ck[r] = val;
                        float a = fabsf(val);
                        if (!(a == a)) a = 0.f;
                        if (!used[r] && (a > v || (a == v && r < rr))) { v = a; rr = r; }
                    }
                }
            }
            // the holders of column k in this wave are lanes kk, kk + TG, ...
#pragma unroll
            for (int o = TG; o < 64; o <<= 1) {
                const float v2 = __shfl_xor(v, o);
                const int r2 = __shfl_xor(rr, o);
                if (v2 > v || (v2 == v && r2 < rr)) { v = v2; rr = r2; }
            }
            if (lane == kk) { wv[wave] = v; wr[wave] = rr; }
            __syncthreads();
            float bv = wv[0];
            int p = wr[0];
            for (int q = 1; q < NW; ++q)
                if (wv[q] > bv || (wv[q] == bv && wr[q] < p)) { bv = wv[q]; p = wr[q]; }
            const float piv = ck[p];
            if (piv == 0.f && tid == 0) atomicCAS(info, 0, (int)(b0 + off + k + 1));
            const float ip = 1.f / piv;
            const bool prow_thread = tr == p % TG;
            const int ip5 = p / TG;
            if (prow_thread) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    float val = 0.f;
#pragma unroll
                    for (int i = 0; i < 8; ++i)
                        if (i == ip5) val = w[i][j];
                    const int c = tc + TG * j;
                    if (c < bs) prow[c] = c == k ? ip : val * ip;
                }
            }
            if (tid == 0) { used[p] = 1; prow_of[k] = p; step_of[p] = k; }
            __syncthreads();
            // w -= f pr everywhere, then column k = -f ip and row p = the scaled
            // pivot row (entries at or beyond bs hold zeros of colk/prow and are
            // never stored)
            float f[8], pr[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) f[i] = ck[tr + TG * i];
#pragma unroll
            for (int j = 0; j < 8; ++j) pr[j] = prow[tc + TG * j];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) w[i][j] -= f[i] * pr[j];
            if (tc == kk) {
#pragma unroll
                for (int i = 0; i < 8; ++i) w[i][kb] = -f[i] * ip;
            }
            if (prow_thread) {
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j) w[i][j] = i == ip5 ? pr[j] : w[i][j];
            }
        }
    }
    float *I = inv + (int64_t)b * i_bstride + off * i_ld + off;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int r = tr + TG * i, c = tc + TG * j;
            if (r < bs && c < bs) I[(int64_t)step_of[r] * i_ld + prow_of[c]] = w[i][j];
        }
}

// FP32 copy of every diagonal block, padded to block x block with the
// identity (rows and columns at or beyond bs of a short last block), so that
// every block has the same shape: the padded inverse is diag(inverse, I).
// Workgroup (r, b) writes row r of block b.
template <typename T>
__global__ __launch_bounds__(256) void k_bj_copy(const T *__restrict__ A, int64_t n, int64_t ld, int block,
                                                 float *__restrict__ X)
{
    // every thread four consecutive columns per pass: all loads of a pass issued together
    // (clamped addresses, so each is in bounds and unconditional), then the stores
    const int b = blockIdx.y, r = blockIdx.x;
    const int64_t b0 = (int64_t)b * block;
    const int bs = (int)min((int64_t)block, n - b0);
    float *x = X + (int64_t)b * block * block + (int64_t)r * block;
    if (r >= bs) {   // uniform: a row of the identity padding
        for (int c = threadIdx.x; c < block; c += blockDim.x) x[c] = c == r ? 1.f : 0.f;
        return;
    }
    const T *a = A + (b0 + r) * ld + b0;
    for (int c0 = 4 * threadIdx.x; c0 < block; c0 += 4 * blockDim.x) {
        T w[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e] = a[min(c0 + e, bs - 1)];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (c0 + e < block) x[c0 + e] = (float)w[e] * (c0 + e < bs ? 1.f : 0.f);
    }
}

// Operand of the batched block products: block b's matrix starts at
// p + b * bstride + r0 * ld + c0.
struct BjOp {
    float *p;
    int64_t bstride;
    int ld, r0, c0;
};

// C_b = alpha A_b B_b (+ C_b when beta) for every diagonal block b (m x n,
// inner dimension k), leaving untouched the C entries whose row or column lies
// in [skip0, skip1) and those whose row and column both lie in [sq0, sq1).
// 64 x 64 tiles of C per workgroup on v_mfma_f32_16x16x4f32 (each wave a
// 32 x 32 quarter, 2 x 2 instruction tiles), K in LDS steps of 16.  C never
// overlaps A or B.
__global__ __launch_bounds__(256) void k_bj_gemm(BjOp Ao, BjOp Bo, BjOp Co, int m, int nn, int kk, float alpha,
                                                 int beta, int skip0, int skip1, int sq0, int sq1)
{
    __shared__ float As[16][68], Bs[16][68];
    const int b = blockIdx.z;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const float *Ap = Ao.p + b * Ao.bstride + (int64_t)Ao.r0 * Ao.ld + Ao.c0;
    const float *Bp = Bo.p + b * Bo.bstride + (int64_t)Bo.r0 * Bo.ld + Bo.c0;
    float *Cp = Co.p + b * Co.bstride + (int64_t)Co.r0 * Co.ld + Co.c0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave >> 1, wn = wave & 1;
    fp_f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = fp_f4{0.f, 0.f, 0.f, 0.f};
    const int ar = tid >> 2, ak = (tid & 3) * 4, bk = tid >> 4, bc = (tid & 15) * 4;
    for (int k0 = 0; k0 < kk; k0 += 16) {
        // clamped addresses (every load in bounds and unconditional, so the eight loads are
        // in flight together) and a 0/1 factor for the entries outside the operands
        float av[4], bv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = m0 + ar, kx = k0 + ak + e;
            av[e] = Ap[(int64_t)min(r, m - 1) * Ao.ld + min(kx, kk - 1)] * ((r < m && kx < kk) ? 1.f : 0.f);
            const int kb = k0 + bk, c = n0 + bc + e;
            bv[e] = Bp[(int64_t)min(kb, kk - 1) * Bo.ld + min(c, nn - 1)] * ((kb < kk && c < nn) ? 1.f : 0.f);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            As[ak + e][ar] = av[e];
            Bs[bk][bc + e] = bv[e];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            float fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) fa[i] = As[ks + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) fb[j] = Bs[ks + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = MfmaOp<float>::run(fa[i], fb[j], acc[i][j]);
        }
        __syncthreads();
    }
    // every load of C (beta) is issued before the first store: a load after a store to a
    // row of C that might alias would wait for it, one memory round trip per entry
    float cv[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = m0 + wm * 32 + i * 16 + MfmaOp<float>::row(lane, q);
                cv[i][j][q] = beta ? Cp[(int64_t)min(r, m - 1) * Co.ld + min(c, nn - 1)] : 0.f;
            }
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = m0 + wm * 32 + i * 16 + MfmaOp<float>::row(lane, q);
                const bool skip = (r >= skip0 && r < skip1) || (c >= skip0 && c < skip1) ||
                                  (r >= sq0 && r < sq1 && c >= sq0 && c < sq1);
                if (r < m && c < nn && !skip) Cp[(int64_t)r * Co.ld + c] = alpha * acc[i][j][q] + cv[i][j][q];
            }
        }
}

// Row chunk J = [j0, j0 + hj) of block b takes W (hj x block) outside the
// chunk's own columns, the other rows take V (block x hj) in J's columns.
// Workgroup (r, b) handles row r.
__global__ __launch_bounds__(256) void k_bj_panels(float *__restrict__ X, int block, int j0, int hj,
                                                   const float *__restrict__ W, const float *__restrict__ V)
{
    const int b = blockIdx.y, r = blockIdx.x;
    float *x = X + (int64_t)b * block * block + (int64_t)r * block;
    if (r >= j0 && r < j0 + hj) {
        const float *w = W + (int64_t)b * hj * block + (int64_t)(r - j0) * block;
        for (int c = threadIdx.x; c < block; c += blockDim.x)
            if (c < j0 || c >= j0 + hj) x[c] = w[c];
    } else {
        const float *v = V + (int64_t)b * block * hj + (int64_t)r * hj;
        for (int c = threadIdx.x; c < hj; c += blockDim.x) x[j0 + c] = v[c];
    }
}

// Wave-wide maximum of v (lane 63 after row prefix maxima and the two row
// broadcasts of the DPP network), returned to every lane.
__device__ __forceinline__ int wave_max_i32(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x111, 0xf, 0xf, false));   // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x112, 0xf, 0xf, false));   // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x114, 0xf, 0xf, false));   // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x118, 0xf, 0xf, false));   // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x142, 0xa, 0xf, false));   // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x143, 0xc, 0xf, false));   // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

// Gauss-Jordan with partial pivoting of a chunk of at most 64 rows in one wave:
// lane r holds row r of the chunk in 64 registers (the same element layout and
// result as k_bj_gj: source element (r, c) at A[b * a_bstride + (off + r) * a_ld +
// off + c], rows never swapped, inverse(pinv(r), p_c) = W(r, c) scattered by the
// final store).  Step k: the pivot is the unused row with the largest |a(r, k)|
// (NaN counted as 0, the lowest row on ties): a DPP wave maximum of the FP32 bit
// patterns (monotonic for non-negative values) and the lowest lane that holds
// it (ballot); the pivot row reaches every lane through LDS (written by its lane,
// read back as broadcasts); k and the columns are compile-time indices; no
// barrier.  Rows at or beyond the chunk's size are never stored.
// One pivot step k of k_bj_gj64 (k a compile-time value once inlined and unrolled).
// The pivot row is not divided by its pivot: it is eliminated from the other rows
// as it stands (multiplier a(r, k) / pivot) and keeps 1 in column k, so that
// every row that has been a pivot row holds its final values times its own
// pivot (later eliminations are linear in it) and is divided once at the end;
// the pivot lane takes a zero multiplier, so one FMA per entry serves every lane.
__device__ __forceinline__ void gj64_step(float (&a)[64], const int k, const int lane, const int bs, bool &used,
                                          int &step, float &mypiv, int *prow_of, float4 *prow, int *__restrict__ info,
                                          int64_t b0, int off)
{
    float av = fabsf(a[k]);
    if (!(av == av)) av = 0.f;
    const int key = used ? -1 : __float_as_int(av);
    const int kmax = wave_max_i32(key);
    const uint64_t hold = __builtin_amdgcn_ballot_w64(key == kmax);
    const int p = (int)__builtin_ctzll(hold);
    const float piv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[k]), p));
    if (piv == 0.f && lane == 0 && k < bs) atomicCAS(info, 0, (int)(b0 + off + k + 1));
    const bool isp = lane == p;
    const float g = isp ? 0.f : a[k] / piv;
    // the pivot row to every lane: lane p writes it to LDS, every lane reads it back
    // as broadcasts (the LDS operations of one wave complete in order)
    if (isp) {
#pragma unroll
        for (int c = 0; c < 16; ++c) prow[c] = make_float4(a[4 * c], a[4 * c + 1], a[4 * c + 2], a[4 * c + 3]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int c4 = 0; c4 < 16; ++c4) {
        const float4 v = prow[c4];
        const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int c = 4 * c4 + t;
            if (c != k) a[c] = fmaf(-g, e[t], a[c]);
        }
    }
    __builtin_amdgcn_wave_barrier();
    a[k] = isp ? 1.f : -g;
    used = used || isp;
    step = isp ? k : step;
    mypiv = isp ? piv : mypiv;
    if (lane == 0) prow_of[k] = p;
}

template <typename T>
__global__ __launch_bounds__(64) void k_bj_gj64(const T *A, int64_t a_bstride, int64_t a_ld, float *inv,
                                                int64_t i_bstride, int64_t i_ld, int64_t n, int block, int off,
                                                int hs, int *__restrict__ info)
{
    __shared__ int prow_of[64];
    __shared__ float4 prow[16];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int64_t b0 = (int64_t)b * block;
    const int bs = hs < 0 ? (int)min((int64_t)block, n - b0) : hs;
    const T *src = A + (int64_t)b * a_bstride + (int64_t)off * a_ld + off;
    // rows and columns at or beyond bs: the identity's, so that steps k >= bs pivot on
    // row k with 1 and change nothing (every step runs, the row stays in registers)
    float a[64];
#pragma unroll
    for (int c = 0; c < 64; ++c)
        a[c] = (lane < bs && c < bs) ? (float)src[(int64_t)lane * a_ld + c] : (lane == c ? 1.f : 0.f);
    bool used = false;
    int step = 0;
    float mypiv = 1.f;
    // two halves, each within the compiler's limit for a fully unrolled loop
#pragma unroll
    for (int k = 0; k < 32; ++k) gj64_step(a, k, lane, bs, used, step, mypiv, prow_of, prow, info, b0, off);
#pragma unroll
    for (int k = 32; k < 64; ++k) gj64_step(a, k, lane, bs, used, step, mypiv, prow_of, prow, info, b0, off);
    const float rp = 1.f / mypiv;
#pragma unroll
    for (int c = 0; c < 64; ++c) a[c] *= rp;
    __syncthreads();
    float *I = inv + (int64_t)b * i_bstride + (int64_t)off * i_ld + off;
    if (lane < bs) {
        float *row = I + (int64_t)step * i_ld;
#pragma unroll
        for (int c = 0; c < 64; ++c)
            if (c < bs) row[prow_of[c]] = a[c];
    }
    if (hs < 0 && bs < block && lane < block) {   // a short whole block: the rest is the identity's
        if (lane < bs) {
            float *row = I + (int64_t)step * i_ld;
            for (int c = bs; c < block; ++c) row[c] = 0.f;
        } else {
            float *row = I + (int64_t)lane * i_ld;
            for (int c = 0; c < block; ++c) row[c] = c == lane ? 1.f : 0.f;
        }
    }
}

constexpr int GJ_TG = 16;               // Gauss-Jordan thread grid: 16 x 16 threads, 8 x 8 values each
// Rows per Gauss-Jordan chunk: 64 (one wave, k_bj_gj64), or 128 with the
// 256-thread register grid of k_bj_gj (FLOWPANEL_BJ_GJ=grid).
static bool gj_grid()
{
    const char *e = getenv("FLOWPANEL_BJ_GJ");
    return e && strcmp(e, "grid") == 0;
}
static int gj_chunk() { return gj_grid() ? 8 * GJ_TG : 64; }
static int bj_chunks(int block) { return (block + gj_chunk() - 1) / gj_chunk(); }

size_t fp_bj_gj_scratch_floats(int64_t n, int block)
{
    if (block <= gj_chunk()) return 0;
    const int64_t nb = (n + block - 1) / block, h = (block + bj_chunks(block) - 1) / bj_chunks(block);
    return (size_t)(2 * nb * h * block);
}

// Inverses of the diagonal blocks straight from A, block <= FP_COARSE_MAX.  Up to
// GJ_CHUNK rows: one Gauss-Jordan pass per block.  Larger blocks: blocked
// Gauss-Jordan over q = ceil(block / GJ_CHUNK) chunks J of h = ceil(block / q)
// rows (pivoting within each chunk), on the identity-padded FP32 copy X:
//   X_JJ = X_JJ^-1;  W = X_JJ X_J*;  X_RC -= X_RJ W_C (R, C outside J);
//   V = -X_*J X_JJ;  X_J* = W, X_*J = V (outside X_JJ).
// scratch holds W and V (fp_bj_gj_scratch_floats).
// copied (optional): recorded on the context stream once A's diagonal blocks
// have been read (other work that reads A can start then without slowing the copy).
// lookahead: run each next chunk's pass on the context's third stream (false:
// everything in order on the context stream).
template <typename T>
int fp_bj_gj_inverse(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, int block, float *inv, int *info_dev,
                     float *scratch, hipEvent_t copied, bool lookahead, bool prefilled)
{
    // the block-Jacobi product takes blocks up to 1024 rows (its callers check); the
    // coarse-space inverse, one block of nc rows, up to FP_COARSE_MAX
    if (block > FP_COARSE_MAX) return fp_fail(ctx, FP_ERR_ARG, "Gauss-Jordan block larger than the coarse-space limit");
    const int nb = (int)((n + block - 1) / block);
    const int64_t bb = (int64_t)block * block;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, info_dev, (int64_t)1, 0);
    ProfScope ps(ctx, "bj_invert");
    const bool grid = gj_grid();
    if (block <= gj_chunk()) {
        if (copied) FP_HIP(ctx, hipEventRecord(copied, ctx->stream));
        if (prefilled && grid)
            hipLaunchKernelGGL((k_bj_gj<float, GJ_TG>), dim3(nb), dim3(GJ_TG * GJ_TG), 0, ctx->stream, (const float *)inv,
                               bb, (int64_t)block, inv, bb, (int64_t)block, n, block, 0, -1, info_dev);
        else if (prefilled)
            hipLaunchKernelGGL((k_bj_gj64<float>), dim3(nb), dim3(64), 0, ctx->stream, (const float *)inv, bb,
                               (int64_t)block, inv, bb, (int64_t)block, n, block, 0, -1, info_dev);
        else if (grid)
            hipLaunchKernelGGL((k_bj_gj<T, GJ_TG>), dim3(nb), dim3(GJ_TG * GJ_TG), 0, ctx->stream, A,
                               (int64_t)block * (ld + 1), ld, inv, bb, (int64_t)block, n, block, 0, -1, info_dev);
        else
            hipLaunchKernelGGL((k_bj_gj64<T>), dim3(nb), dim3(64), 0, ctx->stream, A, (int64_t)block * (ld + 1), ld,
                               inv, bb, (int64_t)block, n, block, 0, -1, info_dev);
        FP_HIP(ctx, hipGetLastError());
        return FP_OK;
    }
    if (!scratch) return fp_fail(ctx, FP_ERR_ARG, "internal: blocked Gauss-Jordan inverse needs scratch");
    const int q = bj_chunks(block), h = (block + q - 1) / q;
    float *W = scratch, *V = scratch + (int64_t)nb * h * block;
    // look-ahead: once X_{J+1,J+1} has its update from chunk J, its Gauss-Jordan
    // pass runs on a second stream beside the rest of chunk J's products
    if (lookahead && !ctx->side2) FP_HIP(ctx, hipStreamCreateWithFlags(&ctx->side2, hipStreamNonBlocking));
    if (lookahead && !ctx->gj_fork) FP_HIP(ctx, hipEventCreateWithFlags(&ctx->gj_fork, hipEventDisableTiming));
    if (lookahead && !ctx->gj_join) FP_HIP(ctx, hipEventCreateWithFlags(&ctx->gj_join, hipEventDisableTiming));
    struct {
        hipEvent_t fork, join;
    } ev{ctx->gj_fork, ctx->gj_join};
    // prefilled: inv already holds the identity-padded diagonal blocks (A unused)
    if (!prefilled) hipLaunchKernelGGL(k_bj_copy<T>, dim3(block, nb), dim3(256), 0, ctx->stream, A, n, ld, block, inv);
    if (copied) FP_HIP(ctx, hipEventRecord(copied, ctx->stream));
    auto gj = [&](int j, hipStream_t s) {
        const int j0 = j * h, hj = min(h, block - j0);
        if (grid)
            hipLaunchKernelGGL((k_bj_gj<float, GJ_TG>), dim3(nb), dim3(GJ_TG * GJ_TG), 0, s, (const float *)inv, bb,
                               (int64_t)block, inv, bb, (int64_t)block, n, block, j0, hj, info_dev);
        else
            hipLaunchKernelGGL((k_bj_gj64<float>), dim3(nb), dim3(64), 0, s, (const float *)inv, bb, (int64_t)block,
                               inv, bb, (int64_t)block, n, block, j0, hj, info_dev);
    };
    const int tb = (block + 63) / 64;
    gj(0, ctx->stream);
    for (int j = 0; j < q; ++j) {
        const int j0 = j * h, hj = min(h, block - j0), th = (hj + 63) / 64;
        const bool ahead = j + 1 < q;
        const int n0 = ahead ? j0 + hj : 0, hn = ahead ? min(h, block - n0) : 0;
        const BjOp XJJ{inv, bb, block, j0, j0}, XJrow{inv, bb, block, j0, 0}, XcolJ{inv, bb, block, 0, j0},
            Xall{inv, bb, block, 0, 0}, Wo{W, (int64_t)hj * block, block, 0, 0}, Vo{V, (int64_t)block * hj, hj, 0, 0};
        // W = X_JJ X_J*  (hj x block)
        hipLaunchKernelGGL(k_bj_gemm, dim3(tb, th, nb), dim3(256), 0, ctx->stream, XJJ, XJrow, Wo, hj, block, hj,
                           1.f, 0, 0, 0, 0, 0);
        if (ahead) {
            // X_{J+1,J+1} -= X_{J+1,J} W_{J+1}, then its inverse on the second stream
            const BjOp XnJ{inv, bb, block, n0, j0}, Wn{W, (int64_t)hj * block, block, 0, n0}, Xnn{inv, bb, block, n0, n0};
            const int tn = (hn + 63) / 64;
            hipLaunchKernelGGL(k_bj_gemm, dim3(tn, tn, nb), dim3(256), 0, ctx->stream, XnJ, Wn, Xnn, hn, hn, hj, -1.f,
                               1, 0, 0, 0, 0);
            if (lookahead) {
                FP_HIP(ctx, hipEventRecord(ev.fork, ctx->stream));
                FP_HIP(ctx, hipStreamWaitEvent(ctx->side2, ev.fork, 0));
                gj(j + 1, ctx->side2);
                FP_HIP(ctx, hipEventRecord(ev.join, ctx->side2));
            } else {
                gj(j + 1, ctx->stream);
            }
        }
        // X_RC -= X_RJ W_C outside row and column chunk J (and outside X_{J+1,J+1}, done above)
        hipLaunchKernelGGL(k_bj_gemm, dim3(tb, tb, nb), dim3(256), 0, ctx->stream, XcolJ, Wo, Xall, block, block, hj,
                           -1.f, 1, j0, j0 + hj, n0, n0 + hn);
        // V = -X_*J X_JJ  (block x hj)
        hipLaunchKernelGGL(k_bj_gemm, dim3(th, tb, nb), dim3(256), 0, ctx->stream, XcolJ, XJJ, Vo, block, hj, hj,
                           -1.f, 0, 0, 0, 0, 0);
        hipLaunchKernelGGL(k_bj_panels, dim3(block, nb), dim3(256), 0, ctx->stream, inv, block, j0, hj,
                           (const float *)W, (const float *)V);
        if (ahead && lookahead) FP_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev.join, 0));
    }
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
template int fp_bj_gj_inverse<double>(fp_ctx *, const double *, int64_t, int64_t, int, float *, int *, float *,
                                      hipEvent_t, bool, bool);
template int fp_bj_gj_inverse<float>(fp_ctx *, const float *, int64_t, int64_t, int, float *, int *, float *,
                                     hipEvent_t, bool, bool);

// y_b = inv_b x_b for every diagonal block.
template <int KM>
__global__ __launch_bounds__(256) void k_bj_apply_inv(const float *__restrict__ inv, int64_t n, int block,
                                                      const double *__restrict__ x, double *__restrict__ y,
                                                      int accumulate)
{
    // four rows per wave (each row in its own diagonal block), up to KM entries per lane
    // and row (block <= 64 KM): 4 KM independent loads per lane in flight
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4;
    if (r0 >= n) return;
    double acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t r = min(r0 + q, n - 1);
        const int64_t b = r / block, b0 = b * block;
        const int bs = (int)min((int64_t)block, n - b0);
        const float *row = inv + b * block * block + (r - b0) * block;
        double a = 0;
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            const int c = lane + 64 * k;
            if (c < bs) a += (double)row[c] * x[b0 + c];
        }
        acc[q] = a;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
    if (lane < 4 && r0 + lane < n) {
        double a = acc[0];
#pragma unroll
        for (int q = 1; q < 4; ++q)
            if (lane == q) a = acc[q];
        const int64_t r = r0 + lane;
        y[r] = accumulate ? y[r] + a : a;
    }
}

// The same with 16-byte loads (block % 4 == 0, 16-byte aligned inv and x):
// lane l takes entries 4 l .. 4 l + 3 of every 256 of a row, KM4 steps per row.
template <int KM4>
__global__ __launch_bounds__(256) void k_bj_apply_inv4(const float *__restrict__ inv, int64_t n, int block,
                                                       const double *__restrict__ x, double *__restrict__ y,
                                                       int accumulate)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4;
    if (r0 >= n) return;
    double acc[4];
    float4 v[4][KM4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t r = min(r0 + q, n - 1);
        const int64_t b = r / block, b0 = b * block;
        const float *row = inv + b * block * block + (r - b0) * block;
#pragma unroll
        for (int k = 0; k < KM4; ++k) {
            const int c = 4 * lane + 256 * k;
            v[q][k] = c < block ? *reinterpret_cast<const float4 *>(row + c) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t r = min(r0 + q, n - 1);
        const int64_t b = r / block, b0 = b * block;
        const int bs = (int)min((int64_t)block, n - b0);
        double a = 0;
#pragma unroll
        for (int k = 0; k < KM4; ++k) {
            const int c = 4 * lane + 256 * k;
            // entries at or beyond bs of a short last block are the identity padding's zeros
            if (c + 3 < bs) {
                const double2 xa = *reinterpret_cast<const double2 *>(x + b0 + c);
                const double2 xb = *reinterpret_cast<const double2 *>(x + b0 + c + 2);
                a += (double)v[q][k].x * xa.x + (double)v[q][k].y * xa.y + (double)v[q][k].z * xb.x +
                     (double)v[q][k].w * xb.y;
            } else {
                const float e[4] = {v[q][k].x, v[q][k].y, v[q][k].z, v[q][k].w};
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (c + t < bs) a += (double)e[t] * x[b0 + c + t];
            }
        }
        acc[q] = a;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
    if (lane < 4 && r0 + lane < n) {
        double a = acc[0];
#pragma unroll
        for (int q = 1; q < 4; ++q)
            if (lane == q) a = acc[q];
        const int64_t r = r0 + lane;
        y[r] = accumulate ? y[r] + a : a;
    }
}

int fp_bj_apply_inv(fp_ctx *ctx, const float *inv, int64_t n, int block, const double *x, double *y, int accumulate)
{
    if (block > 1024) return fp_fail(ctx, FP_ERR_ARG, "block-Jacobi block larger than 1024");
    ProfScope ps(ctx, "bj_apply");
    const dim3 g((unsigned)((n + 15) / 16));
    const char *sc = getenv("FLOWPANEL_BJ_APPLY_SCALAR");
    const bool vec = block % 4 == 0 && (uintptr_t)inv % 16 == 0 && (uintptr_t)x % 16 == 0 && !(sc && sc[0] == '1');
    if (vec && block <= 256)
        hipLaunchKernelGGL(k_bj_apply_inv4<1>, g, dim3(256), 0, ctx->stream, inv, n, block, x, y, accumulate);
    else if (vec && block <= 512)
        hipLaunchKernelGGL(k_bj_apply_inv4<2>, g, dim3(256), 0, ctx->stream, inv, n, block, x, y, accumulate);
    else if (vec)
        hipLaunchKernelGGL(k_bj_apply_inv4<4>, g, dim3(256), 0, ctx->stream, inv, n, block, x, y, accumulate);
    else if (block <= 512)
        hipLaunchKernelGGL(k_bj_apply_inv<8>, g, dim3(256), 0, ctx->stream, inv, n, block, x, y, accumulate);
    else
        hipLaunchKernelGGL(k_bj_apply_inv<16>, g, dim3(256), 0, ctx->stream, inv, n, block, x, y, accumulate);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// Coarse space of chunk indicators Z (column k = 1 on rows [k*cs, (k+1)*cs)).
// AZ = A Z: segmented row sums of A, one wave per row.
template <typename T>
__global__ __launch_bounds__(256) void k_chunk_rowsum(const T *__restrict__ A, int64_t n, int64_t ld, int cs, int nc,
                                                      float *__restrict__ AZ)
{
    // four rows per wave: four independent loads and reductions in flight
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4;
    if (r0 >= n) return;
    const int nr = (int)min((int64_t)4, n - r0);
    for (int k = 0; k < nc; ++k) {
        const int64_t c0 = (int64_t)k * cs, c1 = min(n, c0 + cs);
        double acc[4] = {0, 0, 0, 0};
        for (int64_t c = c0 + lane; c < c1; c += 64) {
            T v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = A[(r0 + (q < nr ? q : 0)) * ld + c];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += (double)v[q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
        if (lane < nr) {
            double a = acc[0];
#pragma unroll
            for (int q = 1; q < 4; ++q)
                if (lane == q) a = acc[q];
            AZ[(int64_t)k * n + r0 + lane] = (float)a;
        }
    }
}

// The same for chunks shorter than a wave: one lane per chunk, four rows per
// workgroup (a wave-wide reduction per chunk would cost more than the chunk's
// loads).
template <typename T>
__global__ __launch_bounds__(256) void k_chunk_rowsum_lane(const T *__restrict__ A, int64_t n, int64_t ld, int cs,
                                                           int nc, float *__restrict__ AZ)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const T *a = A + r * ld;
    for (int k0 = 0; k0 < nc; k0 += 64) {
        const int k = k0 + lane;
        if (k >= nc) break;
        const int64_t c0 = (int64_t)k * cs, c1 = min(n, c0 + cs);
        double acc = 0;
#pragma unroll 8
        for (int64_t c = c0; c < c1; ++c) acc += (double)a[c];
        AZ[(int64_t)k * n + r] = (float)acc;
    }
}

// Chunk sums of rows for chunks shorter than a wave, read coalesced:
// out[k * nrows + r] = sum of A[r * ld + c] over the columns c of chunk k
// (columns k cs .. k cs + cs - 1, the last chunk cut at ncols), summed in FP64 in
// column order.  Each wave takes one row and walks it in segments of 64 chunks:
// the segment goes into the wave's slice of LDS with 16-byte loads, then lane j
// sums chunk j of the segment out of LDS.  The four waves of a workgroup take
// the same number of segments (the barriers are uniform); dynamic LDS of
// 4 * 64 * cs elements.
template <typename T, typename OT>
__global__ __launch_bounds__(256) void k_chunk_rowsum_seg(const T *__restrict__ A, int64_t nrows, int64_t ncols,
                                                          int64_t ld, int cs, int vec, OT *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char seg_raw[];
    constexpr int V = 16 / sizeof(T);
    using VT = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * 4 + wave;
    const bool live = r < nrows;
    T *seg = reinterpret_cast<T *>(seg_raw) + (int64_t)wave * 64 * cs;
    const int64_t span = (int64_t)64 * cs;
    for (int64_t s0 = 0; s0 < ncols; s0 += span) {
        const int L = (int)min(span, ncols - s0);
        if (live) {
            const T *a = A + r * ld + s0;
            const int LV = vec ? L / V * V : 0;   // 16-byte loads where the rows are 16-byte aligned
            for (int i = lane * V; i < LV; i += 64 * V)
                *reinterpret_cast<VT *>(seg + i) = *reinterpret_cast<const VT *>(a + i);
            for (int i = LV + lane; i < L; i += 64) seg[i] = a[i];
        }
        __syncthreads();
        const int c0 = lane * cs;
        if (live && c0 < L) {
            const int e = min(cs, L - c0);
            double acc = 0;
            for (int c = 0; c < e; ++c) acc += (double)seg[c0 + c];
            out[(s0 / cs + lane) * nrows + r] = (OT)acc;
        }
        __syncthreads();
    }
}

template <typename T, typename OT>
static int chunk_rowsum_seg(fp_ctx *ctx, const T *A, int64_t nrows, int64_t ncols, int64_t ld, int cs, OT *out)
{
    const size_t lds = (size_t)4 * 64 * cs * sizeof(T);
    static bool raised = false;   // per instantiation
    if (lds > 65536 && !raised) {
        FP_HIP(ctx, hipFuncSetAttribute((const void *)k_chunk_rowsum_seg<T, OT>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        raised = true;
    }
    const int vec = ((uintptr_t)A % 16 == 0 && (ld * sizeof(T)) % 16 == 0) ? 1 : 0;
    hipLaunchKernelGGL((k_chunk_rowsum_seg<T, OT>), dim3((unsigned)((nrows + 3) / 4)), dim3(256), lds, ctx->stream, A,
                       nrows, ncols, ld, cs, vec, out);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// E = Z^T AZ (nc x nc, row-major): E[k'][k] = sum over chunk k' of AZ[:, k]
__global__ void k_coarse_galerkin(const float *__restrict__ AZ, int64_t n, int cs, int nc, double *__restrict__ E)
{
    const int kp = blockIdx.x, k = blockIdx.y;
    __shared__ double red[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)kp * cs, r1 = min(n, r0 + cs);
    double acc = 0;
    for (int64_t r = r0 + tid; r < r1; r += blockDim.x) acc += AZ[(int64_t)k * n + r];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (tid == 0) E[(int64_t)kp * nc + k] = red[0] + red[1] + red[2] + red[3];
}

// Larger of two (|value|, row) candidates, the lower row on ties; a NaN never wins.
__device__ __forceinline__ void argmax_pair(double &v, int &r, double v2, int r2)
{
    if (v2 > v || (v2 == v && r2 < r)) { v = v2; r = r2; }
}

// In-place inverse of a small dense matrix (m <= SMALL_INV_LDS, row-major) in
// LDS by Gauss-Jordan elimination with partial pivoting (row interchanges, the
// pivot row scaled by the reciprocal pivot, the columns swapped back in reverse
// order at the end); one workgroup.  The pivot of each column is found by all
// lanes at once (wave shuffles, then the 16 wave winners in LDS); ties go to the
// lowest row, and a column without a finite candidate keeps its diagonal row.
constexpr int SMALL_INV_LDS = 128;

__global__ __launch_bounds__(1024) void k_small_inverse_lds(double *__restrict__ E, int m, int *__restrict__ info)
{
    __shared__ double a[SMALL_INV_LDS][SMALL_INV_LDS + 1];
    __shared__ int piv[SMALL_INV_LDS];
    __shared__ double col[SMALL_INV_LDS];
    __shared__ double wv[16];
    __shared__ int wr[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nt = blockDim.x;
    for (int e = tid; e < m * m; e += nt) a[e / m][e % m] = E[e];
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        double v = -1.0;
        int r = m;
        for (int rr = k + tid; rr < m; rr += nt) argmax_pair(v, r, fabs(a[rr][k]), rr);
        for (int o = 32; o > 0; o >>= 1) {
            const double v2 = __shfl_xor(v, o);
            const int r2 = __shfl_xor(r, o);
            argmax_pair(v, r, v2, r2);
        }
        if (lane == 0) { wv[wave] = v; wr[wave] = r; }
        __syncthreads();
        double bv = wv[0];
        int p = wr[0];
        for (int q = 1; q < (nt >> 6); ++q) argmax_pair(bv, p, wv[q], wr[q]);
        if (p >= m) p = k;
        if (tid == 0) piv[k] = p;
        if (p != k)
            for (int e = tid; e < m; e += nt) {
                const double t = a[k][e];
                a[k][e] = a[p][e];
                a[p][e] = t;
            }
        __syncthreads();
        const double pv = a[k][k];
        if (pv == 0.0) {
            if (tid == 0 && *info == 0) *info = k + 1;
            __syncthreads();
            continue;
        }
        const double inv = 1.0 / pv;
        // the multipliers (column k of every row) are saved before any entry changes
        for (int i = tid; i < m; i += nt) col[i] = a[i][k];
        __syncthreads();
        for (int e = tid; e < m; e += nt) a[k][e] = e == k ? inv : a[k][e] * inv;
        __syncthreads();
        // 32 x 32 thread grid over the (<= 128)^2 entries: no index division
#pragma unroll
        for (int ii = 0; ii < SMALL_INV_LDS / 32; ++ii) {
            const int i = (tid >> 5) + 32 * ii;
            if (i >= m || i == k) continue;
            const double f = col[i];
#pragma unroll
            for (int jj = 0; jj < SMALL_INV_LDS / 32; ++jj) {
                const int j = (tid & 31) + 32 * jj;
                if (j < m) a[i][j] = j == k ? -f * inv : a[i][j] - f * a[k][j];
            }
        }
        __syncthreads();
    }
    for (int k = m - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < m; i += nt) {
                const double t = a[i][k];
                a[i][k] = a[i][p];
                a[i][p] = t;
            }
        __syncthreads();
    }
    for (int e = tid; e < m * m; e += nt) E[e] = a[e / m][e % m];
}

__global__ __launch_bounds__(256) void k_widen_f32(const float *__restrict__ x, int64_t m, double *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) y[i] = (double)x[i];
}

// FLOWPANEL_COARSE_INV=lu: coarse spaces from this size up are inverted through the
// randomized no-pivot LU and the inverses of its factors instead of the blocked
// Gauss-Jordan.  Measured on the 30,000-panel BWB with 3,000 coarse vectors: 7.5 ms
// on the side stream (about 4 nc^3 flop of FP32 MFMA products at ~50 TFLOP/s, and
// ~280 launches whose host time delays the block inverses queued after them)
// against 8.3 ms for the 47 Gauss-Jordan passes, and no shorter set-up; the pivoted
// Gauss-Jordan also survives Galerkin matrices too close to singular for a
// factorisation without pivoting, so it stays the default.
constexpr int COARSE_INV_LU_MIN = 512;
constexpr uint64_t COARSE_RBT_SEED = 0x434F415253455242ull;

size_t fp_coarse_work_bytes(int64_t nc)
{
    return std::max<size_t>((size_t)2 * nc * nc * sizeof(double), fp_inverse_rbt_work_bytes(nc));
}

// lookahead: each next Gauss-Jordan chunk's pass runs on ctx->side2 (with the
// ctx->gj_fork / gj_join events) beside the current chunk's products
template <typename T>
int fp_coarse_setup(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, int cs, float *AZ, double *Einv, double *work,
                    int *info_dev, bool lookahead)
{
    const int nc = (int)((n + cs - 1) / cs);
    if (nc > FP_COARSE_MAX) return fp_fail(ctx, FP_ERR_ARG, "coarse space larger than 4096 vectors");
    ProfScope ps(ctx, "coarse_setup");
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, info_dev, (int64_t)1, 0);
    // AZ: chunk sums of the rows of A; E = Z^T AZ: chunk sums of the columns of AZ
    // (AZ column k is contiguous, so E[k'][k] is chunk k' of "row" k of AZ)
    const char *lanev = getenv("FLOWPANEL_COARSE_LANE");
    if (cs < 64 && !(lanev && lanev[0] == '1')) {
        int st = chunk_rowsum_seg<T, float>(ctx, A, n, n, ld, cs, AZ);
        if (!st) st = chunk_rowsum_seg<float, double>(ctx, AZ, nc, n, n, cs, Einv);
        if (st) return st;
    } else {
        if (cs < 64)
            hipLaunchKernelGGL(k_chunk_rowsum_lane<T>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, A, n,
                               ld, cs, nc, AZ);
        else
            hipLaunchKernelGGL(k_chunk_rowsum<T>, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, ctx->stream, A, n, ld,
                               cs, nc, AZ);
        hipLaunchKernelGGL(k_coarse_galerkin, dim3(nc, nc), dim3(256), 0, ctx->stream, (const float *)AZ, n, cs, nc,
                           Einv);
    }
    const char *ci = getenv("FLOWPANEL_COARSE_INV");
    const bool gj = !(ci && !strcmp(ci, "lu"));
    if (nc <= SMALL_INV_LDS) {
        hipLaunchKernelGGL(k_small_inverse_lds, dim3(1), dim3(1024), 0, ctx->stream, Einv, nc, info_dev);
    } else if (nc >= COARSE_INV_LU_MIN && !gj) {
        // large coarse spaces: the randomized no-pivot FP32 LU and the inverses of its
        // factors (the blocked Gauss-Jordan below is a chain of nc / 64 passes)
        int st = fp_inverse_rbt_f32(ctx, Einv, nc, nc, work, fp_coarse_work_bytes(nc), info_dev, COARSE_RBT_SEED);
        if (st) return st;
    } else {
        // larger coarse spaces: FP32 blocked Gauss-Jordan of E (it only feeds the
        // preconditioner), widened back in place; work (2 nc^2 doubles) holds the
        // FP32 inverse and the Gauss-Jordan scratch
        float *inv32 = reinterpret_cast<float *>(work), *scr = inv32 + (int64_t)nc * nc;
        if (fp_bj_gj_scratch_floats(nc, nc) > (size_t)3 * nc * nc)
            return fp_fail(ctx, FP_ERR_ARG, "internal: coarse inverse scratch too small");
        int st = fp_bj_gj_inverse<double>(ctx, Einv, nc, nc, nc, inv32, info_dev, scr, nullptr, lookahead);
        if (st) return st;
        hipLaunchKernelGGL(k_widen_f32, dim3((unsigned)(((int64_t)nc * nc + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const float *)inv32, (int64_t)nc * nc, Einv);
    }
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
template int fp_coarse_setup<double>(fp_ctx *, const double *, int64_t, int64_t, int, float *, double *, double *, int *,
                                     bool);
template int fp_coarse_setup<float>(fp_ctx *, const float *, int64_t, int64_t, int, float *, double *, double *, int *,
                                    bool);

// t = Z^T v: one wave per chunk (chunk sums of v)
__global__ __launch_bounds__(256) void k_coarse_restrict(const double *__restrict__ v, int64_t n, int cs, int nc,
                                                         double *__restrict__ t)
{
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= nc) return;
    const int64_t r0 = (int64_t)k * cs, r1 = min(n, r0 + cs);
    double acc = 0;
    for (int64_t r = r0 + lane; r < r1; r += 64) acc += v[r];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) t[k] = acc;
}

// c = Einv t: one wave per coarse row
__global__ __launch_bounds__(256) void k_coarse_solve(const double *__restrict__ Einv, int nc,
                                                      const double *__restrict__ t, double *__restrict__ c)
{
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= nc) return;
    double acc = 0;
    for (int q = lane; q < nc; q += 64) acc += Einv[(int64_t)k * nc + q] * t[q];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) c[k] = acc;
}

// y = Z c, r2 = v - AZ c: 64 rows per workgroup, the nc columns of AZ split
// over the 4 waves (coalesced rows of AZ, summed through LDS)
__global__ __launch_bounds__(256) void k_coarse_resid(const double *__restrict__ v, const float *__restrict__ AZ,
                                                      int64_t n, int cs, int nc, const double *__restrict__ c,
                                                      double *__restrict__ y, double *__restrict__ r2)
{
    __shared__ double part[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    double acc = 0;
    if (i < n) {
        int k = wave;
        for (; k + 28 < nc; k += 32) {   // 8 independent loads per step
            float a[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) a[q] = AZ[(int64_t)(k + 4 * q) * n + i];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc += (double)a[q] * c[k + 4 * q];
        }
        for (; k < nc; k += 4) acc += (double)AZ[(int64_t)k * n + i] * c[k];
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && i < n) {
        r2[i] = v[i] - (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
        y[i] = c[i / cs];
    }
}

// The same with four rows per lane and 16-byte loads of A Z (n % 4 == 0, 16-byte
// aligned A Z): 256 rows per workgroup, the nc columns split over the 4 waves.
__global__ __launch_bounds__(256) void k_coarse_resid4(const double *__restrict__ v, const float *__restrict__ AZ,
                                                       int64_t n, int cs, int nc, const double *__restrict__ c,
                                                       double *__restrict__ y, double *__restrict__ r2)
{
    __shared__ double part[4][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * 256 + 4 * lane;
    double acc[4] = {0, 0, 0, 0};
    if (i0 < n) {
        int k = wave;
        for (; k + 12 < nc; k += 16) {   // 4 independent 16-byte loads per step
            float4 a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = *reinterpret_cast<const float4 *>(AZ + (int64_t)(k + 4 * q) * n + i0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double cq = c[k + 4 * q];
                acc[0] += (double)a[q].x * cq;
                acc[1] += (double)a[q].y * cq;
                acc[2] += (double)a[q].z * cq;
                acc[3] += (double)a[q].w * cq;
            }
        }
        for (; k < nc; k += 4) {
            const float4 a = *reinterpret_cast<const float4 *>(AZ + (int64_t)k * n + i0);
            const double ck = c[k];
            acc[0] += (double)a.x * ck;
            acc[1] += (double)a.y * ck;
            acc[2] += (double)a.z * ck;
            acc[3] += (double)a.w * ck;
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) part[wave][t][lane] = acc[t];
    __syncthreads();
    // wave w finalises rows i0 + w of every lane's group of four
    const int64_t i = i0 + wave;
    if (i0 < n && i < n) {
        const double a = part[0][wave][lane] + part[1][wave][lane] + part[2][wave][lane] + part[3][wave][lane];
        r2[i] = v[i] - a;
        y[i] = c[i / cs];
    }
}

// k_coarse_resid4 with the coarse vectors split over grid.y = ks slices: at 30,000
// rows one slice is 118 workgroups for 256 CUs (360 MB of A Z at 3 TB/s); the
// slices' partial sums go to part[s * n + i] and k_coarse_finish adds them.
__global__ __launch_bounds__(256) void k_coarse_resid4s(const float *__restrict__ AZ, int64_t n, int nc,
                                                        const double *__restrict__ c, double *__restrict__ part)
{
    __shared__ double ps[4][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * 256 + 4 * lane;
    const int ks = gridDim.y, sl = blockIdx.y;
    const int kb = (int)((int64_t)nc * sl / ks), ke = (int)((int64_t)nc * (sl + 1) / ks);
    double acc[4] = {0, 0, 0, 0};
    if (i0 < n) {
        int k = kb + wave;
        for (; k + 12 < ke; k += 16) {   // 4 independent 16-byte loads per step
            float4 a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = *reinterpret_cast<const float4 *>(AZ + (int64_t)(k + 4 * q) * n + i0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double cq = c[k + 4 * q];
                acc[0] += (double)a[q].x * cq;
                acc[1] += (double)a[q].y * cq;
                acc[2] += (double)a[q].z * cq;
                acc[3] += (double)a[q].w * cq;
            }
        }
        for (; k < ke; k += 4) {
            const float4 a = *reinterpret_cast<const float4 *>(AZ + (int64_t)k * n + i0);
            const double ck = c[k];
            acc[0] += (double)a.x * ck;
            acc[1] += (double)a.y * ck;
            acc[2] += (double)a.z * ck;
            acc[3] += (double)a.w * ck;
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) ps[wave][t][lane] = acc[t];
    __syncthreads();
    const int64_t i = i0 + wave;
    if (i0 < n && i < n) part[(int64_t)sl * n + i] = ps[0][wave][lane] + ps[1][wave][lane] + ps[2][wave][lane] + ps[3][wave][lane];
}

__global__ __launch_bounds__(256) void k_coarse_finish(const double *__restrict__ v, int64_t n, int cs, int ks,
                                                       const double *__restrict__ c, const double *__restrict__ part,
                                                       double *__restrict__ y, double *__restrict__ r2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = 0.0;
    for (int s = 0; s < ks; ++s) a += part[(int64_t)s * n + i];
    r2[i] = v[i] - a;
    y[i] = c[i / cs];
}

// slices of k_coarse_resid4s for n rows: about 1,024 workgroups, at most COARSE_KS
constexpr int COARSE_KS = 8;
static int coarse_slices(int64_t n, int nc)
{
    const int64_t rb = (n + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)COARSE_KS, (1024 + rb - 1) / rb, (int64_t)nc / 16}));
}

size_t fp_coarse_tc_bytes(int64_t n, int nc) { return ((size_t)2 * nc + (size_t)COARSE_KS * n) * sizeof(double); }

int fp_coarse_apply(fp_ctx *ctx, const double *v, int64_t n, int cs, const float *AZ, const double *Einv, double *tc,
                    double *y, double *r2)
{
    const int nc = (int)((n + cs - 1) / cs);
    ProfScope ps(ctx, "coarse_apply");
    const unsigned g4 = (unsigned)((nc + 3) / 4);
    hipLaunchKernelGGL(k_coarse_restrict, dim3(g4), dim3(256), 0, ctx->stream, v, n, cs, nc, tc);
    hipLaunchKernelGGL(k_coarse_solve, dim3(g4), dim3(256), 0, ctx->stream, Einv, nc, (const double *)tc, tc + nc);
    const int ks = coarse_slices(n, nc);
    if (n % 4 == 0 && (uintptr_t)AZ % 16 == 0 && ks > 1) {
        double *part = tc + 2 * nc;   // tc holds fp_coarse_tc_bytes
        hipLaunchKernelGGL(k_coarse_resid4s, dim3((unsigned)((n + 255) / 256), (unsigned)ks), dim3(256), 0, ctx->stream,
                           AZ, n, nc, (const double *)(tc + nc), part);
        hipLaunchKernelGGL(k_coarse_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, v, n, cs, ks,
                           (const double *)(tc + nc), (const double *)part, y, r2);
    } else if (n % 4 == 0 && (uintptr_t)AZ % 16 == 0)
        hipLaunchKernelGGL(k_coarse_resid4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, v, AZ, n, cs,
                           nc, (const double *)(tc + nc), y, r2);
    else
        hipLaunchKernelGGL(k_coarse_resid, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, ctx->stream, v, AZ, n, cs,
                           nc, (const double *)(tc + nc), y, r2);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// conversions and BLAS-1
__global__ void k_convert(const double *__restrict__ src, int64_t n, int64_t lds, void *__restrict__ dst, int64_t ldd,
                          int prec, int transpose)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= n) return;
    // transpose: src column-major (element (i,j) at j*lds + i), dst row-major
    const double v = transpose ? src[j * lds + i] : src[i * lds + j];
    if (prec == FP_F64) ((double *)dst)[i * ldd + j] = v;
    else ((float *)dst)[i * ldd + j] = (float)v;
}

int fp_mat_convert(fp_ctx *ctx, const double *src, int64_t n, int64_t lds, void *dst, int64_t ldd, int dst_prec,
                   int transpose)
{
    if (n <= 0) return FP_OK;
    hipLaunchKernelGGL(k_convert, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, ctx->stream, src, n,
                       lds, dst, ldd, dst_prec, transpose);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

__global__ void k_narrow(const double *__restrict__ src, float *__restrict__ dst, int64_t n, int64_t ld)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j < n) dst[i * ld + j] = (float)src[i * ld + j];
}

int fp_mat_narrow(fp_ctx *ctx, const double *src, float *dst, int64_t n, int64_t ld)
{
    if (n <= 0) return FP_OK;
    hipLaunchKernelGGL(k_narrow, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, ctx->stream, src, dst,
                       n, ld);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

__global__ void k_resid(const double *__restrict__ b, double *__restrict__ r, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - r[i];
}

int fp_residual_update(fp_ctx *ctx, const double *b, double *r, int64_t n)
{
    if (n <= 0) return FP_OK;
    hipLaunchKernelGGL(k_resid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, b, r, n);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// multi-dot: out[c] = sum_i V[c*ldv + i] * w[i], c < nv; two passes (partials
// per workgroup in a fixed order, then a single-workgroup sum) so the result is
// deterministic.
// out[c] = V[:, c] . w for c < nv: MDOT_G columns per grid row accumulated in
// registers, MDOT_PARTS row chunks; a second kernel sums the chunks (one
// thread per column, fixed order).
constexpr int MDOT_G = 8;
constexpr int MDOT_PARTS = 128;

__global__ __launch_bounds__(256) void k_mdot_partial(const double *__restrict__ V, int64_t ldv, int nv,
                                                      const double *__restrict__ w, int64_t n,
                                                      double *__restrict__ part)
{
    __shared__ double red[4][MDOT_G];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.y * MDOT_G;
    const int nc = min(MDOT_G, nv - c0);
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double acc[MDOT_G];
#pragma unroll
    for (int q = 0; q < MDOT_G; ++q) acc[q] = 0;
    for (int64_t i = i0 + tid; i < i1; i += 256) {
        const double wi = w[i];
#pragma unroll
        for (int q = 0; q < MDOT_G; ++q)
            if (q < nc) acc[q] += V[(int64_t)(c0 + q) * ldv + i] * wi;
    }
#pragma unroll
    for (int q = 0; q < MDOT_G; ++q) {
        double a = acc[q];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
        if (lane == 0) red[wave][q] = a;
    }
    __syncthreads();
    if (tid < nc) part[(int64_t)(c0 + tid) * gridDim.x + blockIdx.x] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// one wave per column: every partial loaded at once, then a shuffle sum
__global__ __launch_bounds__(256) void k_mdot_final(const double *__restrict__ part, int nparts, int nv,
                                                    double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nv) return;
    const double *pc = part + (int64_t)c * nparts;
    double acc = 0;
#pragma unroll
    for (int k = 0; k < MDOT_PARTS; k += 64)
        if (k + lane < nparts) acc += pc[k + lane];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) out[c] = acc;
}

int fp_mdot(fp_ctx *ctx, const double *V, int64_t ldv, int nv, const double *w, int64_t n, double *part,
            double *out)
{
    if (nv <= 0) return FP_OK;
    const int parts = (int)std::max<int64_t>(1, std::min<int64_t>(MDOT_PARTS, (n + 255) / 256));
    hipLaunchKernelGGL(k_mdot_partial, dim3(parts, (nv + MDOT_G - 1) / MDOT_G), dim3(256), 0, ctx->stream, V, ldv, nv,
                       w, n, part);
    hipLaunchKernelGGL(k_mdot_final, dim3((nv + 3) / 4), dim3(256), 0, ctx->stream, (const double *)part, parts, nv,
                       out);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// w -= V h (h on the device, nv entries)
__global__ void k_vupdate(const double *__restrict__ V, int64_t ldv, int nv, const double *__restrict__ h,
                          double *__restrict__ w, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = w[i];
    for (int c = 0; c < nv; ++c) acc -= V[c * ldv + i] * h[c];
    w[i] = acc;
}

int fp_vupdate(fp_ctx *ctx, const double *V, int64_t ldv, int nv, const double *h, double *w, int64_t n)
{
    hipLaunchKernelGGL(k_vupdate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, V, ldv, nv, h, w, n);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// y = a*x + b*y
__global__ void k_axpby(double a, const double *__restrict__ x, double b, double *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

int fp_axpby(fp_ctx *ctx, double a, const double *x, double b, double *y, int64_t n)
{
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a, x, b, y, n);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// Arnoldi bookkeeping of restarted GMRES kept on the device, so that the
// iterations are queued back to back without a host round trip; the host only
// reads the residual estimates every few iterations.  One thread: the work is
// O(j) per iteration.  H row-major (m+1) x m.
__global__ void k_arnoldi_start(double *__restrict__ g, int m, double beta)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= m) g[i] = i == 0 ? beta : 0.0;
}

__global__ void k_arnoldi_step(const double *__restrict__ hbuf, int m, int j, double bnorm, double *__restrict__ H,
                               double *__restrict__ cs, double *__restrict__ sn, double *__restrict__ g,
                               double *__restrict__ relh, double *__restrict__ hnv)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double *h1 = hbuf, *h2 = hbuf + (m + 2);
    const double hn = sqrt(fmax(0.0, hbuf[2 * (m + 2) - 1]));
    for (int i = 0; i <= j; ++i) H[(int64_t)i * m + j] = h1[i] + h2[i];
    for (int i = 0; i < j; ++i) {
        const double a = H[(int64_t)i * m + j], b = H[(int64_t)(i + 1) * m + j];
        H[(int64_t)i * m + j] = cs[i] * a + sn[i] * b;
        H[(int64_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
    }
    const double hj = H[(int64_t)j * m + j];
    const double den = hypot(hj, hn);
    cs[j] = den > 0 ? hj / den : 1.0;
    sn[j] = den > 0 ? hn / den : 0.0;
    H[(int64_t)j * m + j] = den;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    relh[j] = fabs(g[j + 1]) / bnorm;
    hnv[j] = hn;
}

__global__ void k_arnoldi_normalize(const double *__restrict__ w, const double *__restrict__ hnv, int j,
                                    double *__restrict__ v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double hn = hnv[j];
    if (i < n && hn > 0) v[i] = w[i] / hn;
}

// y = H(0:jc, 0:jc)^-1 g(0:jc), upper triangular
__global__ void k_arnoldi_solve(const double *__restrict__ H, int m, int jc, const double *__restrict__ g,
                                double *__restrict__ y)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = jc - 1; i >= 0; --i) {
        double acc = g[i];
        for (int q = i + 1; q < jc; ++q) acc -= H[(int64_t)i * m + q] * y[q];
        y[i] = acc / H[(int64_t)i * m + i];
    }
}

int fp_arnoldi_start(fp_ctx *ctx, double *g, int m, double beta)
{
    hipLaunchKernelGGL(k_arnoldi_start, dim3((m + 256) / 256), dim3(256), 0, ctx->stream, g, m, beta);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

int fp_arnoldi_step(fp_ctx *ctx, const double *hbuf, int m, int j, double bnorm, double *H, double *cs, double *sn,
                    double *g, double *relh, double *hnv, const double *w, double *vnext, int64_t n)
{
    hipLaunchKernelGGL(k_arnoldi_step, dim3(1), dim3(64), 0, ctx->stream, hbuf, m, j, bnorm, H, cs, sn, g, relh, hnv);
    hipLaunchKernelGGL(k_arnoldi_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, w,
                       (const double *)hnv, j, vnext, n);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// ----------------------------------------------------------------------------
// Fused classical Gram-Schmidt passes for the Arnoldi step (fewer launches per
// Krylov iteration).  A pass is k_mdot_partial with CGS_PARTS row chunks, then
// k_cgs_update: every workgroup sums the CGS_PARTS partials of each coefficient
// itself (independent loads), subtracts V h from its rows of w and, on the second
// pass, leaves the sum of squares of its rows of the new w.  k_arnoldi_step2
// finishes both coefficient vectors and the norm in one workgroup and applies the
// Givens rotations.  Requires nv <= CGS_MAXV.
constexpr int CGS_PARTS = FP_CGS_PARTS, CGS_MAXV = FP_CGS_MAXV;

size_t fp_cgs_work_doubles(int64_t n, int m) { return (size_t)2 * CGS_PARTS * (m + 2) + (n + 255) / 256 + 64; }

template <bool NORM>
__global__ __launch_bounds__(256) void k_cgs_update(const double *__restrict__ V, int64_t ldv, int nv,
                                                    const double *__restrict__ part, int P,
                                                    double *__restrict__ w, int64_t n, double *__restrict__ normpart)
{
    __shared__ double hl[CGS_MAXV];
    __shared__ double red[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int c = tid; c < nv; c += 256) {
        const double *pc = part + (int64_t)c * P;
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < CGS_PARTS; ++q)
            if (q < P) a += pc[q];
        hl[c] = a;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + tid;
    double acc = 0.0;
    if (i < n) {
        acc = w[i];
#pragma unroll 8
        for (int c = 0; c < nv; ++c) acc -= V[(int64_t)c * ldv + i] * hl[c];
        w[i] = acc;
    }
    if (NORM) {
        double sq = acc * acc;
        for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
        if (lane == 0) red[wave] = sq;
        __syncthreads();
        if (tid == 0) normpart[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

int fp_cgs_pass(fp_ctx *ctx, const double *V, int64_t ldv, int nv, double *w, int64_t n, double *part,
                double *normpart)
{
    if (nv > CGS_MAXV) return fp_fail(ctx, FP_ERR_ARG, "Gram-Schmidt pass: too many basis vectors");
    const int P = (int)std::max<int64_t>(1, std::min<int64_t>(CGS_PARTS, (n + 255) / 256));
    hipLaunchKernelGGL(k_mdot_partial, dim3(P, (nv + MDOT_G - 1) / MDOT_G), dim3(256), 0, ctx->stream, V, ldv, nv,
                       (const double *)w, n, part);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (normpart)
        hipLaunchKernelGGL(k_cgs_update<true>, grid, dim3(256), 0, ctx->stream, V, ldv, nv, (const double *)part, P, w,
                           n, normpart);
    else
        hipLaunchKernelGGL(k_cgs_update<false>, grid, dim3(256), 0, ctx->stream, V, ldv, nv, (const double *)part, P,
                           w, n, normpart);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// Hessenberg column j from the two passes' partials (h1 + h2) and the norm of
// what is left, then the Givens rotations as in k_arnoldi_step.
__global__ __launch_bounds__(256) void k_arnoldi_step2(const double *__restrict__ part1,
                                                       const double *__restrict__ part2, int P,
                                                       const double *__restrict__ normpart, int nblk, int m, int j,
                                                       double bnorm, double *__restrict__ H, double *__restrict__ cs,
                                                       double *__restrict__ sn, double *__restrict__ g,
                                                       double *__restrict__ relh, double *__restrict__ hnv)
{
    __shared__ double hc[CGS_MAXV], lcs[CGS_MAXV], lsn[CGS_MAXV];
    __shared__ double red[256];
    const int tid = threadIdx.x;
    for (int i = tid; i < j; i += 256) {
        lcs[i] = cs[i];
        lsn[i] = sn[i];
    }
    for (int i = tid; i <= j; i += 256) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int q = 0; q < CGS_PARTS; ++q)
            if (q < P) {
                a += part1[(int64_t)i * P + q];
                b += part2[(int64_t)i * P + q];
            }
        hc[i] = a + b;
    }
    double s = 0.0;
    for (int q = tid; q < nblk; q += 256) s += normpart[q];
    red[tid] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    // the earlier rotations applied to the column in LDS (one thread: they are sequential)
    if (tid == 0) {
        const double hn = sqrt(fmax(0.0, red[0]));
        for (int i = 0; i < j; ++i) {
            const double a = hc[i], b = hc[i + 1];
            hc[i] = lcs[i] * a + lsn[i] * b;
            hc[i + 1] = -lsn[i] * a + lcs[i] * b;
        }
        const double hj = hc[j];
        const double den = hypot(hj, hn);
        const double c = den > 0 ? hj / den : 1.0, sj = den > 0 ? hn / den : 0.0;
        cs[j] = c;
        sn[j] = sj;
        hc[j] = den;
        const double gj = g[j];
        g[j + 1] = -sj * gj;
        g[j] = c * gj;
        relh[j] = fabs(sj * gj) / bnorm;
        hnv[j] = hn;
    }
    __syncthreads();
    for (int i = tid; i <= j; i += 256) H[(int64_t)i * m + j] = hc[i];
}

int fp_arnoldi_step2(fp_ctx *ctx, const double *part1, const double *part2, const double *normpart, int m, int j,
                     double bnorm, double *H, double *cs, double *sn, double *g, double *relh, double *hnv,
                     const double *w, double *vnext, int64_t n)
{
    const int P = (int)std::max<int64_t>(1, std::min<int64_t>(CGS_PARTS, (n + 255) / 256));
    const int nblk = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_arnoldi_step2, dim3(1), dim3(256), 0, ctx->stream, part1, part2, P, normpart, nblk, m, j,
                       bnorm, H, cs, sn, g, relh, hnv);
    hipLaunchKernelGGL(k_arnoldi_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, w,
                       (const double *)hnv, j, vnext, n);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

int fp_arnoldi_solve(fp_ctx *ctx, const double *H, int m, int jc, const double *g, double *y)
{
    hipLaunchKernelGGL(k_arnoldi_solve, dim3(1), dim3(64), 0, ctx->stream, H, m, jc, g, y);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

// x += V z (z on the device, nv entries): Krylov solution update
__global__ void k_vcombine(const double *__restrict__ V, int64_t ldv, int nv, const double *__restrict__ z,
                           double *__restrict__ x, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0;
    for (int c = 0; c < nv; ++c) acc += V[c * ldv + i] * z[c];
    x[i] = acc;
}

int fp_vcombine(fp_ctx *ctx, const double *V, int64_t ldv, int nv, const double *z, double *x, int64_t n)
{
    hipLaunchKernelGGL(k_vcombine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, V, ldv, nv, z, x, n);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
