// fp_post.hip -- surface post-processing kernels (gfx950, FP64).
//
// Surface velocity from the gradient of the doublet strength, cell-centred
// Green-Gauss scheme (src/FLOWPanel_postprocess.jl:139-244, with the trailing-
// edge treatment of RigidWakeBody at :245-497).  One lane per cell walks its
// edges; the neighbour across edge k of cell j is nbr[k*N + j] (-1 if none),
// found once on the host from the connectivity.  For triangles this is the
// reference's formula; other polygons use all their edges.
#include "fp_internal.h"

namespace {

template <int NVP>
__device__ __forceinline__ void ugm_cell(const double *__restrict__ pv, const int *__restrict__ nvv, int64_t N,
                                         const double *__restrict__ geo, const int *__restrict__ nbr,
                                         const double *__restrict__ G, double sgn, int64_t j, double *acc)
{
    const int nv = nvv[j];
    const double nx = geo[j], ny = geo[N + j], nz = geo[2 * N + j];
    const double cx = geo[3 * N + j], cy = geo[4 * N + j], cz = geo[5 * N + j];
    const double area = geo[6 * N + j];
    acc[0] = acc[1] = acc[2] = 0;
    for (int k = 0; k < NVP; ++k) {
        if (k >= nv) break;
        const int m = nbr[(int64_t)k * N + j];
        if (m < 0) continue;
        const int q = (k + 1) % nv;
        const double ax = pv[(3 * k) * N + j], ay = pv[(3 * k + 1) * N + j], az = pv[(3 * k + 2) * N + j];
        const double rx = pv[(3 * q) * N + j] - ax, ry = pv[(3 * q + 1) * N + j] - ay, rz = pv[(3 * q + 2) * N + j] - az;
        double dx = ry * nz - rz * ny, dy = rz * nx - rx * nz, dz = rx * ny - ry * nx;
        const double dm = sqrt(dx * dx + dy * dy + dz * dz);
        dx /= dm; dy /= dm; dz /= dm;
        const double mx = ax - cx, my = ay - cy, mz = az - cz;
        const double qx = ax - geo[3 * N + m], qy = ay - geo[4 * N + m], qz = az - geo[5 * N + m];
        const double c1 = my * rz - mz * ry, c2 = mz * rx - mx * rz, c3 = mx * ry - my * rx;
        const double e1 = qy * rz - qz * ry, e2 = qz * rx - qx * rz, e3 = qx * ry - qy * rx;
        const double dmain = sqrt(c1 * c1 + c2 * c2 + c3 * c3), dnear = sqrt(e1 * e1 + e2 * e2 + e3 * e3);
        const double f = dnear / (dmain + dnear);
        const double face = f * G[j] + (1.0 - f) * G[m];
        const double mag = face * sqrt(rx * rx + ry * ry + rz * rz) / area;
        acc[0] += 0.5 * sgn * dx * mag;
        acc[1] += 0.5 * sgn * dy * mag;
        acc[2] += 0.5 * sgn * dz * mag;
    }
}

// out (3 x N point-major) -= sum of the cell's edge terms
template <int NVP>
__global__ __launch_bounds__(256) void k_ugradmu_cell(const double *__restrict__ pv, const int *__restrict__ nvv,
                                                      int64_t N, const double *__restrict__ geo,
                                                      const int *__restrict__ nbr, const double *__restrict__ G,
                                                      double sgn, double *__restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double acc[3];
    ugm_cell<NVP>(pv, nvv, N, geo, nbr, G, sgn, j, acc);
    out[3 * j] -= acc[0];
    out[3 * j + 1] -= acc[1];
    out[3 * j + 2] -= acc[2];
}

// trailing-edge panels: their edge terms added back, once per appearance in
// the shedding list, in list order
template <int NVP>
__global__ void k_ugradmu_te(const double *__restrict__ pv, const int *__restrict__ nvv, int64_t N,
                             const double *__restrict__ geo, const int *__restrict__ nbr, const double *__restrict__ G,
                             double sgn, const int64_t *__restrict__ sh, int64_t ns, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int64_t e = 0; e < ns; ++e)
        for (int side = 0; side < 2; ++side) {
            const int64_t c = sh[6 * e + 3 * side];
            if (c < 0) continue;
            double acc[3];
            ugm_cell<NVP>(pv, nvv, N, geo, nbr, G, sgn, c, acc);
            out[3 * c] += acc[0];
            out[3 * c + 1] += acc[1];
            out[3 * c + 2] += acc[2];
        }
}

// Node-based scheme (src/FLOWPanel_postprocess.jl:622-748): nodal values of
// Gamma as the area-weighted mean over the cells holding each node (node -> cell
// lists in CSR form), then per cell the in-plane gradient of the plane through
// its first three vertices carrying those values.
__global__ __launch_bounds__(256) void k_nodal_mean(const int64_t *__restrict__ nptr, const int *__restrict__ ncell,
                                                    int64_t nn, const double *__restrict__ geo, int64_t N,
                                                    const double *__restrict__ G, double *__restrict__ nodal)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    double s = 0, w = 0;
    for (int64_t q = nptr[v]; q < nptr[v + 1]; ++q) {
        const int c = ncell[q];
        const double a = geo[6 * N + c];
        s += a * G[c];
        w += a;
    }
    nodal[v] = w > 0 ? s / w : 0.0;
}

// out (3 x N point-major) += -1/2 grad: e1 along the first edge, e2 completing
// the plane of the first three vertices, the plane f = f1 + gx x + gy y through
// (0, 0), (L, 0), (x3, y3).  cn holds the first three node indices of each cell.
// Sharp trailing edge (cell_side != nullptr): a cell on side 1 or 2 of the
// trailing edge takes, at the TE nodes (node_te[v] = k >= 0), the value averaged
// over its own side's cells, sidev[(side - 1) * nte + k].
__global__ __launch_bounds__(256) void k_ugradmu_node(const double *__restrict__ pv, int64_t N,
                                                      const int *__restrict__ cn, const double *__restrict__ nodal,
                                                      double sgn, const int *__restrict__ cell_side,
                                                      const int *__restrict__ node_te, const double *__restrict__ sidev,
                                                      int nte, double *__restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double v[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[k][c] = pv[(3 * k + c) * N + j];
    const int side = cell_side ? cell_side[j] : 0;
    double f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int node = cn[3 * j + k];
        double val = nodal[node];
        if (side > 0) {
            const int t = node_te[node];
            if (t >= 0) val = sidev[(int64_t)(side - 1) * nte + t];
        }
        f[k] = sgn * val;
    }
    const double f1 = f[0], f2 = f[1], f3 = f[2];
    double e1[3], w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { e1[c] = v[1][c] - v[0][c]; w[c] = v[2][c] - v[0][c]; }
    const double L = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) e1[c] /= L;
    const double x3 = w[0] * e1[0] + w[1] * e1[1] + w[2] * e1[2];
    double e2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e2[c] = w[c] - x3 * e1[c];
    const double y3 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) e2[c] /= y3;
    const double gx = (f2 - f1) / L, gy = (f3 - f1 - gx * x3) / y3;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * j + c] += -0.5 * (gx * e1[c] + gy * e2[c]);
}

// Sharp trailing edge, the side values: for TE node te[k] and side s = 1, 2 the
// area-weighted mean of Gamma over the cells holding it that lie on side s
// (kept at the plain nodal value when none does).  One lane per (node, side).
__global__ void k_te_side_means(const int *__restrict__ te, int nte, const int64_t *__restrict__ nptr,
                                const int *__restrict__ ncell, const int *__restrict__ cell_side,
                                const double *__restrict__ geo, int64_t N, const double *__restrict__ G,
                                const double *__restrict__ nodal, double *__restrict__ sidev)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nte) return;
    const int s = t / nte + 1, k = t % nte, v = te[k];
    double acc = 0, w = 0;
    for (int64_t q = nptr[v]; q < nptr[v + 1]; ++q) {
        const int c = ncell[q];
        if (cell_side[c] != s) continue;
        const double a = geo[6 * N + c];
        acc += a * G[c];
        w += a;
    }
    sidev[t] = w > 0 ? acc / w : nodal[v];
}

// the nodal values left behind hold side 2's values at the TE nodes (the
// reference's sharpTE pass overwrites them with the second side's last)
__global__ void k_te_store_side2(const int *__restrict__ te, int nte, const double *__restrict__ sidev,
                                 double *__restrict__ nodal)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nte) nodal[te[k]] = sidev[nte + k];
}

// trailing-edge override of the node-based values (force_cellTE): cells pi and
// pi + 1, and pj and pj - 1 when pj exists, take the cell-centred values
__global__ void k_ugradmu_force_te(const int64_t *__restrict__ sh, int64_t ns, int64_t N,
                                   const double *__restrict__ cellv, double *__restrict__ nodev)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ns) return;
    const int64_t pi = sh[6 * e], pj = sh[6 * e + 3];
    const int64_t cs[4] = {pi, pi + 1, pj, pj - 1};
    for (int q = 0; q < 4; ++q) {
        if (q >= 2 && pj < 0) break;
        const int64_t c = cs[q];
        if (c < 0 || c >= N) continue;
        for (int d = 0; d < 3; ++d) nodev[3 * c + d] = cellv[3 * c + d];
    }
}

// combined scheme (:756-921): a cell uses the node-based value when one of its
// edges has no neighbour or the angle between its normal and a neighbour's
// exceeds the critical angle, the cell-centred value otherwise; out += choice
__global__ __launch_bounds__(256) void k_ugradmu_combine(const int *__restrict__ nvv, int64_t N,
                                                         const int *__restrict__ nbr, const double *__restrict__ geo,
                                                         double cosc, const double *__restrict__ cellv,
                                                         const double *__restrict__ nodev, double *__restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int nv = nvv[j];
    const double nx = geo[j], ny = geo[N + j], nz = geo[2 * N + j];
    bool use_node = false;
    for (int k = 0; k < nv; ++k) {
        const int m = nbr[(int64_t)k * N + j];
        if (m < 0) { use_node = true; break; }
        if (nx * geo[m] + ny * geo[N + m] + nz * geo[2 * N + m] < cosc) { use_node = true; break; }
    }
    const double *src = use_node ? nodev : cellv;
    for (int d = 0; d < 3; ++d) out[3 * j + d] += src[3 * j + d];
}

// Element forces (calcfield_F!, src/FLOWPanel_postprocess.jl:1199-1307):
// F_i = -Cp_i q A_i s n_i with s the sign of CPoffset (1 if 0: the normals
// flipped by CPoffset, calc_normals(body; flipbyCPoffset=true)), point-major
// 3 x N; and per workgroup the partial sums of F and of (X_i - Xac) x F_i
// (calcfield_Ftot!, calcfield_Mtot!, :1543-1784), X the control points (geo)
// or the caller's points (SoA).
__global__ __launch_bounds__(256) void k_loads(const double *__restrict__ geo, int64_t N, const double *__restrict__ cp,
                                               double q, double s, const double *__restrict__ X, double xa, double ya,
                                               double za, double *__restrict__ F, double *__restrict__ part)
{
    __shared__ double red[4][6];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i < N) {
        const double c = -cp[i] * q * geo[6 * N + i] * s;
        v[0] = c * geo[i];
        v[1] = c * geo[N + i];
        v[2] = c * geo[2 * N + i];
        if (F) { F[3 * i] = v[0]; F[3 * i + 1] = v[1]; F[3 * i + 2] = v[2]; }
        const double x = (X ? X[i] : geo[3 * N + i]) - xa, y = (X ? X[N + i] : geo[4 * N + i]) - ya,
                     z = (X ? X[2 * N + i] : geo[5 * N + i]) - za;
        v[3] = y * v[2] - z * v[1];
        v[4] = z * v[0] - x * v[2];
        v[5] = x * v[1] - y * v[0];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double a = v[k];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
        if (lane == 0) red[wave][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < 6)
        part[6 * (int64_t)blockIdx.x + threadIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out6 = sum of the nblk partial 6-vectors (one workgroup)
__global__ __launch_bounds__(256) void k_loads_reduce(const double *__restrict__ part, int64_t nblk,
                                                      double *__restrict__ out6)
{
    __shared__ double red[4][6];
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] += part[6 * b + k];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double a = v[k];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
        if (lane == 0) red[wave][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < 6) out6[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                             red[3][threadIdx.x];
}

// Sectional loads of a structured grid (calcfield_sectionalforce!, :1411-1542):
// station j sums F over the chordwise cells i + nchord j and averages the
// span coordinate spandir . X of their points; one wave per station.
__global__ __launch_bounds__(64) void k_sectional(const double *__restrict__ F, const double *__restrict__ X, int64_t N,
                                                  int64_t nchord, int64_t nspan, double sx, double sy, double sz,
                                                  double *__restrict__ outF, double *__restrict__ pos)
{
    const int64_t j = blockIdx.x;
    if (j >= nspan) return;
    double a[4] = {0, 0, 0, 0};
    for (int64_t i = threadIdx.x; i < nchord; i += 64) {
        const int64_t c = i + nchord * j;
        a[0] += F[3 * c];
        a[1] += F[3 * c + 1];
        a[2] += F[3 * c + 2];
        a[3] += sx * X[c] + sy * X[N + c] + sz * X[2 * N + c];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o);
    if (threadIdx.x == 0) {
        outF[3 * j] = a[0];
        outF[3 * j + 1] = a[1];
        outF[3 * j + 2] = a[2];
        pos[j] = a[3] / (double)nchord;
    }
}

}  // namespace

int fp_launch_loads(fp_ctx *ctx, const fp_body *b, const double *cp, double q, const double *X, const double *Xac,
                    double *F, double *part, double *out6)
{
    const int64_t N = b->N;
    if (N == 0) return FP_OK;
    const double s = b->cp_offset < 0 ? -1.0 : 1.0;
    const int64_t nblk = (N + 255) / 256;
    ProfScope ps(ctx, "loads");
    hipLaunchKernelGGL(k_loads, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, b->d_geo, N, cp, q, s, X, Xac[0],
                       Xac[1], Xac[2], F, part);
    hipLaunchKernelGGL(k_loads_reduce, dim3(1), dim3(256), 0, ctx->stream, (const double *)part, nblk, out6);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

size_t fp_loads_partial_bytes(int64_t N) { return (size_t)6 * ((N + 255) / 256 + 1) * sizeof(double); }

int fp_launch_sectional(fp_ctx *ctx, const double *F, const double *X, int64_t N, int64_t nchord, int64_t nspan,
                        const double *spandir, double *outF, double *pos)
{
    if (nspan <= 0 || nchord <= 0) return FP_OK;
    ProfScope ps(ctx, "loads");
    hipLaunchKernelGGL(k_sectional, dim3((unsigned)nspan), dim3(64), 0, ctx->stream, F, X, N, nchord, nspan, spandir[0],
                       spandir[1], spandir[2], outF, pos);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

int fp_launch_ugradmu_node(fp_ctx *ctx, const fp_body *b, const int64_t *d_nptr, const int *d_ncell, int64_t nn,
                           const int *d_cn, double *d_nodal, double *out, const TeSides *te)
{
    const int64_t N = b->N;
    if (N == 0) return FP_OK;
    const double sgn = b->cp_offset < 0 ? -1.0 : 1.0;
    ProfScope ps(ctx, "ugradmu_node");
    if (nn > 0)
        hipLaunchKernelGGL(k_nodal_mean, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, ctx->stream, d_nptr,
                           d_ncell, nn, b->d_geo, N, b->d_strength, d_nodal);
    const bool sharp = te && te->nte > 0;
    if (sharp)
        hipLaunchKernelGGL(k_te_side_means, dim3((unsigned)((2 * te->nte + 63) / 64)), dim3(64), 0, ctx->stream,
                           te->te_nodes, te->nte, d_nptr, d_ncell, te->cell_side, b->d_geo, N, b->d_strength, d_nodal,
                           te->sidev);
    hipLaunchKernelGGL(k_ugradmu_node, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, b->d_pv, N, d_cn,
                       d_nodal, sgn, sharp ? te->cell_side : (const int *)nullptr,
                       sharp ? te->node_te : (const int *)nullptr, sharp ? te->sidev : (const double *)nullptr,
                       sharp ? te->nte : 0, out);
    if (sharp)
        hipLaunchKernelGGL(k_te_store_side2, dim3((unsigned)((te->nte + 63) / 64)), dim3(64), 0, ctx->stream,
                           te->te_nodes, te->nte, (const double *)te->sidev, d_nodal);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

int fp_launch_ugradmu_combine(fp_ctx *ctx, const fp_body *b, const int *d_nbr, const int64_t *d_sh, int force_cell_te,
                              double cos_crit, double *cellv, double *nodev, double *out)
{
    const int64_t N = b->N;
    if (N == 0) return FP_OK;
    ProfScope ps(ctx, "ugradmu_combine");
    if (force_cell_te && b->ns > 0)
        hipLaunchKernelGGL(k_ugradmu_force_te, dim3((unsigned)((b->ns + 63) / 64)), dim3(64), 0, ctx->stream, d_sh,
                           b->ns, N, cellv, nodev);
    hipLaunchKernelGGL(k_ugradmu_combine, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, b->d_nv, N,
                       d_nbr, b->d_geo, cos_crit, cellv, nodev, out);
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}

int fp_launch_ugradmu(fp_ctx *ctx, const fp_body *b, const int *d_nbr, const int64_t *d_sh, int te, double *out)
{
    const int64_t N = b->N;
    if (N == 0) return FP_OK;
    const double sgn = b->cp_offset == 0 ? 1.0 : (b->cp_offset > 0 ? 1.0 : -1.0);
    const dim3 grid((unsigned)((N + 255) / 256));
    ProfScope ps(ctx, "ugradmu_cell");
    if (b->nvp == 4) {
        hipLaunchKernelGGL(k_ugradmu_cell<4>, grid, dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, N, b->d_geo, d_nbr,
                           b->d_strength, sgn, out);
        if (te && b->ns > 0)
            hipLaunchKernelGGL(k_ugradmu_te<4>, dim3(1), dim3(64), 0, ctx->stream, b->d_pv, b->d_nv, N, b->d_geo,
                               d_nbr, b->d_strength, sgn, d_sh, b->ns, out);
    } else {
        hipLaunchKernelGGL(k_ugradmu_cell<8>, grid, dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, N, b->d_geo, d_nbr,
                           b->d_strength, sgn, out);
        if (te && b->ns > 0)
            hipLaunchKernelGGL(k_ugradmu_te<8>, dim3(1), dim3(64), 0, ctx->stream, b->d_pv, b->d_nv, N, b->d_geo,
                               d_nbr, b->d_strength, sgn, d_sh, b->ns, out);
    }
    FP_HIP(ctx, hipGetLastError());
    return FP_OK;
}
