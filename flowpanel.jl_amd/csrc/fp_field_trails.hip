// fp_field_trails.hip -- velocity of a vortex-ring body at many targets as a sum
// over the mesh's edges, walked as trails (gfx950).
//
// U(X) = sum_j Gamma_j sum_{edges a->b of panel j} seg(a, b; X) (the
// reference's Uind!, src/FLOWPanel_liftingbody.jl:576-661, with the ring of
// src/FLOWPanel_elements.jl:442-467) regroups into one term per mesh edge,
//   U(X) = sum_e gamma_e seg(a_e, b_e; X),  gamma_e = sum of +-Gamma_j of the
// panels that hold e (+ when the panel runs a_e -> b_e), as seg(b, a) = -seg(a, b).
// The host walks the edges as trails (a path cover: every edge once, each step
// starting at the node where the last one ended), so a lane evaluates the node
// terms r = X - p and 1/(|r| + offset) once per trail step and keeps the
// previous node's in registers: about one node term and one segment per edge,
// and a quad mesh has about two edges per panel, against four of each per panel
// ring by ring.  A small kernel fills the stream (node, edge vector, gamma) from
// the current strengths before each evaluation; the wake horseshoes follow as
// before.  n . U (the matrix-free residual) uses the same stream; the potential
// and source bodies keep the panel-wise kernel.
#include "fp_internal.h"
#include "fp_elements.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace {
constexpr int TR_TILE = 256;    // stream entries per LDS tile
constexpr int TR_COMP = 7;      // x, y, z, ex, ey, ez, gamma
}  // namespace

// Host: trails over the unique edges of the mesh (0-based CSR cells, nn nodes).
// Per stream entry m: ref[m] = panel * 8 + vertex of the entry's node; the edge
// from entry m-1's node to entry m's node has contributors cpan[cptr[m] ..
// cptr[m+1]) = (panel << 1) | reversed; an entry without contributors starts a
// trail.  Returns the number of entries.
static int64_t build_trails(const int64_t *cell_ptr, const int64_t *cell_idx, int64_t N, int64_t nn,
                            std::vector<int32_t> &ref, std::vector<int32_t> &cptr, std::vector<int32_t> &cpan)
{
    auto nv = [&](int64_t j) { return (int)(cell_ptr[j + 1] - cell_ptr[j]); };
    auto cell = [&](int64_t j, int k) { return cell_idx[cell_ptr[j] + k]; };
    std::vector<int32_t> noderef(nn, -1);
    std::unordered_map<uint64_t, int64_t> emap;
    std::vector<int64_t> ea, eb;                     // edge endpoints (first traversal)
    std::vector<std::vector<std::pair<int64_t, int>>> contrib;  // (panel, runs a -> b ? 0 : 1)
    for (int64_t j = 0; j < N; ++j)
        for (int k = 0; k < nv(j); ++k) {
            const int64_t u = cell(j, k), w = cell(j, (k + 1) % nv(j));
            if (noderef[u] < 0) noderef[u] = (int32_t)(j * 8 + k);
            if (u == w) continue;
            const uint64_t key = u < w ? ((uint64_t)u << 32) | (uint64_t)w : ((uint64_t)w << 32) | (uint64_t)u;
            auto it = emap.find(key);
            int64_t e;
            if (it == emap.end()) {
                e = (int64_t)ea.size();
                emap.emplace(key, e);
                ea.push_back(u);
                eb.push_back(w);
                contrib.emplace_back();
            } else {
                e = it->second;
            }
            contrib[e].emplace_back(j, ea[e] == u ? 0 : 1);
        }
    const int64_t ne = (int64_t)ea.size();
    // node -> incident edges
    std::vector<int64_t> aptr(nn + 1, 0), adj(2 * ne);
    for (int64_t e = 0; e < ne; ++e) { ++aptr[ea[e] + 1]; ++aptr[eb[e] + 1]; }
    for (int64_t v = 0; v < nn; ++v) aptr[v + 1] += aptr[v];
    std::vector<int64_t> fill(aptr.begin(), aptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e) { adj[fill[ea[e]]++] = e; adj[fill[eb[e]]++] = e; }
    std::vector<char> used(ne, 0);
    std::vector<int64_t> next(aptr.begin(), aptr.end() - 1), left(nn);
    for (int64_t v = 0; v < nn; ++v) left[v] = aptr[v + 1] - aptr[v];
    ref.clear();
    cptr.assign(1, 0);
    cpan.clear();
    auto walk = [&](int64_t s) {
        ref.push_back(noderef[s]);
        cptr.push_back((int32_t)cpan.size());
        int64_t cur = s;
        for (;;) {
            while (next[cur] < aptr[cur + 1] && used[adj[next[cur]]]) ++next[cur];
            if (next[cur] == aptr[cur + 1]) break;
            const int64_t e = adj[next[cur]];
            used[e] = 1;
            const int64_t other = ea[e] == cur ? eb[e] : ea[e];
            --left[ea[e]];
            --left[eb[e]];
            // traversed cur -> other; a contributor running ea -> eb keeps its sign when ea == cur
            for (const auto &c : contrib[e]) {
                const int rev = (ea[e] == cur) ? c.second : 1 - c.second;
                cpan.push_back((int32_t)((c.first << 1) | rev));
            }
            ref.push_back(noderef[other]);
            cptr.push_back((int32_t)cpan.size());
            cur = other;
        }
    };
    // trails from odd-degree nodes first (each such node ends or starts one), then the rest
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t v = 0; v < nn; ++v)
            while (left[v] > 0 && (pass == 1 || (left[v] & 1))) walk(v);
    return (int64_t)ref.size();
}

int fp_body_build_trails(fp_ctx *ctx, fp_body *b)
{
    if (b->element != FP_ELEM_VORTEXRING || b->N == 0) return FP_OK;
    std::vector<int32_t> ref, cptr, cpan;
    const int64_t S = build_trails(b->cell_ptr.data(), b->cell_idx.data(), b->N, (int64_t)b->nodes.size() / 3, ref,
                                   cptr, cpan);
    if (S == 0) return FP_OK;
    const size_t ib = (S + (S + 1) + std::max<size_t>(1, cpan.size())) * sizeof(int32_t);
    FP_HIP(ctx, hipMalloc((void **)&b->d_tr_tab, ib));
    FP_HIP(ctx, hipMalloc((void **)&b->d_tr, (size_t)TR_COMP * S * sizeof(double)));
    int32_t *t = b->d_tr_tab;
    FP_HIP(ctx, hipMemcpy(t, ref.data(), S * sizeof(int32_t), hipMemcpyHostToDevice));
    FP_HIP(ctx, hipMemcpy(t + S, cptr.data(), (S + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!cpan.empty())
        FP_HIP(ctx, hipMemcpy(t + 2 * S + 1, cpan.data(), cpan.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    b->tr_S = S;
    return FP_OK;
}

extern "C" int64_t fp_field_trails(const int64_t *cell_ptr, const int64_t *cell_idx, int64_t ncells, int64_t nnodes,
                                   int index_base, int32_t *ref, int32_t *cptr, int32_t *cpan, int64_t cap_entries,
                                   int64_t cap_contrib, int64_t *ncontrib)
{
    if (!cell_ptr || !cell_idx || ncells <= 0 || nnodes <= 0) return -FP_ERR_ARG;
    std::vector<int64_t> cp(ncells + 1), ci((size_t)(cell_ptr[ncells] - cell_ptr[0]));
    for (int64_t c = 0; c <= ncells; ++c) cp[c] = cell_ptr[c] - cell_ptr[0];
    for (int64_t c = 0; c < ncells; ++c)
        if (cp[c + 1] - cp[c] < 3 || cp[c + 1] - cp[c] > FP_MAX_NV) return -FP_ERR_UNSUPPORTED;
    for (size_t k = 0; k < ci.size(); ++k) {
        ci[k] = cell_idx[cell_ptr[0] + (int64_t)k] - index_base;
        if (ci[k] < 0 || ci[k] >= nnodes) return -FP_ERR_ARG;
    }
    std::vector<int32_t> r, p, c;
    const int64_t S = build_trails(cp.data(), ci.data(), ncells, nnodes, r, p, c);
    if (ncontrib) *ncontrib = (int64_t)c.size();
    if (ref && cptr && cpan && cap_entries >= S && cap_contrib >= (int64_t)c.size()) {
        std::memcpy(ref, r.data(), S * sizeof(int32_t));
        std::memcpy(cptr, p.data(), (S + 1) * sizeof(int32_t));
        std::memcpy(cpan, c.data(), c.size() * sizeof(int32_t));
    }
    return S;
}

// Stream entries from the current strengths: node, edge vector from the previous
// entry's node (FP64 coordinates), gamma (0 at a trail start).
__global__ void k_trail_prep(const int32_t *__restrict__ ref, const int32_t *__restrict__ cptr,
                             const int32_t *__restrict__ cpan, int64_t S, const double *__restrict__ pv, int64_t N,
                             const double *__restrict__ strength, double *__restrict__ tr)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= S) return;
    const int32_t r = ref[m];
    const int64_t j = r >> 3;
    const int k = r & 7;
    const double x = pv[(3 * k + 0) * N + j], y = pv[(3 * k + 1) * N + j], z = pv[(3 * k + 2) * N + j];
    double ex = 0.0, ey = 0.0, ez = 0.0, g = 0.0;
    const int c0 = cptr[m], c1 = cptr[m + 1];
    if (c1 > c0) {
        const int32_t rp = ref[m - 1];
        const int64_t jp = rp >> 3;
        const int kp = rp & 7;
        ex = x - pv[(3 * kp + 0) * N + jp];
        ey = y - pv[(3 * kp + 1) * N + jp];
        ez = z - pv[(3 * kp + 2) * N + jp];
        for (int c = c0; c < c1; ++c) {
            const int32_t q = cpan[c];
            const double s = strength[q >> 1];
            g += (q & 1) ? -s : s;
        }
    }
    tr[0 * S + m] = x; tr[1 * S + m] = y; tr[2 * S + m] = z;
    tr[3 * S + m] = ex; tr[4 * S + m] = ey; tr[5 * S + m] = ez;
    tr[6 * S + m] = g;
}

// Sum of the per-split partial results in a fixed order.
__global__ void k_trail_reduce(const double *__restrict__ partial, int splits, int64_t M, int comps,
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

// One lane per target; the stream [p0, p1) in LDS tiles, read by broadcast.
template <int MODE>
__global__ __launch_bounds__(256) void k_field_trail(const double *__restrict__ tr, int64_t S, TargetView tv,
                                                     HorseView hv, const double *__restrict__ strength, double off,
                                                     double cut, int64_t pchunk, double *__restrict__ out,
                                                     double *__restrict__ partial)
{
    __shared__ double st[TR_COMP][TR_TILE];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < tv.M;
    const double X = ok ? tv.x[i] : 0.0, Y = ok ? tv.y[i] : 0.0, Z = ok ? tv.z[i] : 0.0;
    const double off2 = off * off, cut2 = cut * cut;
    const int64_t p0 = (int64_t)blockIdx.y * pchunk;
    const int64_t p1 = min(S, p0 + pchunk);
    double ux = 0, uy = 0, uz = 0;
    // the node before the chunk (its entry's node terms start the first segment)
    double rpx = 0, rpy = 0, rpz = 0, ivp = 0;
    if (p0 > 0 && p0 < S) {
        rpx = X - tr[0 * S + p0 - 1];
        rpy = Y - tr[1 * S + p0 - 1];
        rpz = Z - tr[2 * S + p0 - 1];
        ivp = fp_recip(fp_sqrt(rpx * rpx + rpy * rpy + rpz * rpz) + off);
    }
    for (int64_t t0 = p0; t0 < p1; t0 += TR_TILE) {
        const int cnt = (int)min((int64_t)TR_TILE, p1 - t0);
        __syncthreads();
        for (int t = threadIdx.x; t < TR_COMP * TR_TILE; t += blockDim.x) {
            const int c = t / TR_TILE, q = t % TR_TILE;
            if (q < cnt) st[c][q] = tr[(int64_t)c * S + t0 + q];
        }
        __syncthreads();
        // straight-line body (a select instead of branches, so that the unrolled entries
        // interleave): a trail start (gamma = 0) and a segment below the cutoff add zero
#pragma unroll 2
        for (int q = 0; q < cnt; ++q) {
            const double rx = X - st[0][q], ry = Y - st[1][q], rz = Z - st[2][q];
            const double iv = fp_recip(fp_sqrt_nb(rx * rx + ry * ry + rz * rz) + off);
            const double g = st[6][q];
            const double cx = rpy * rz - rpz * ry;
            const double cy = rpz * rx - rpx * rz;
            const double cz = rpx * ry - rpy * rx;
            const double c2 = cx * cx + cy * cy + cz * cz;
            const double proj = st[3][q] * (rpx * ivp - rx * iv) + st[4][q] * (rpy * ivp - ry * iv) +
                                st[5][q] * (rpz * ivp - rz * iv);
            const double a = g * proj * fp_recip(4.0 * FP_PI * (c2 + off2));
            const double aux = (g != 0.0 && c2 > cut2) ? a : 0.0;
            ux += aux * cx; uy += aux * cy; uz += aux * cz;
            rpx = rx; rpy = ry; rpz = rz; ivp = iv;
        }
    }
    // wake horseshoes (split 0 only), strengths of their panels
    if (blockIdx.y == 0 && hv.nh > 0) {
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
            if (ok)
                for (int hh = 0; hh < cnt; ++hh)
                    horseshoe_u<double>(sh[hh], X, Y, Z, sh[hh][12], off, off2, cut2, ux, uy, uz);
        }
    }
    if (!ok) return;
    const int64_t M = tv.M;
    if (partial) {
        double *pp = partial + (int64_t)blockIdx.y * (MODE == 0 ? 3 : 1) * M;
        if (MODE == 0) { pp[i] = ux; pp[M + i] = uy; pp[2 * M + i] = uz; }
        else pp[i] = ux * tv.nx[i] + uy * tv.ny[i] + uz * tv.nz[i];
    } else {
        if (MODE == 0) { out[3 * i] += ux; out[3 * i + 1] += uy; out[3 * i + 2] += uz; }
        else out[i] += ux * tv.nx[i] + uy * tv.ny[i] + uz * tv.nz[i];
    }
}

// Velocity (mode 0) or n . velocity (mode 1) of a vortex-ring body through its
// trails; *used = 0 when the body has none or FLOWPANEL_FIELD=panel.
int fp_launch_field_trails(fp_ctx *ctx, const fp_body *b, const double *strength, const TargetView &tv, int mode,
                           double *out, double *partial, int splits, int *used)
{
    *used = 0;
    if (mode > 1 || !b->d_tr || b->tr_S == 0) return FP_OK;
    const char *env = getenv("FLOWPANEL_FIELD");
    if (env && strcmp(env, "panel") == 0) return FP_OK;
    *used = 1;
    const int64_t S = b->tr_S;
    const int32_t *t = b->d_tr_tab;
    hipLaunchKernelGGL(k_trail_prep, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, ctx->stream, t, t + S,
                       t + 2 * S + 1, S, (const double *)b->d_pv, b->N, strength, b->d_tr);
    FP_HIP(ctx, hipGetLastError());
    int64_t pchunk = (S + splits - 1) / splits;
    pchunk = (pchunk + TR_TILE - 1) / TR_TILE * TR_TILE;
    const int sp = (int)((S + pchunk - 1) / pchunk);
    double *part = sp > 1 ? partial : nullptr;
    dim3 grid((unsigned)((tv.M + 255) / 256), (unsigned)sp);
    const HorseView hv = horse_view(b);
    {
        ProfScope ps(ctx, mode == 1 ? "field_matvec_f64" : "field_velocity_f64");
        if (mode == 0)
            hipLaunchKernelGGL(k_field_trail<0>, grid, dim3(256), 0, ctx->stream, (const double *)b->d_tr, S, tv, hv,
                               strength, b->kernel_offset, b->kernel_cutoff, pchunk, out, part);
        else
            hipLaunchKernelGGL(k_field_trail<1>, grid, dim3(256), 0, ctx->stream, (const double *)b->d_tr, S, tv, hv,
                               strength, b->kernel_offset, b->kernel_cutoff, pchunk, out, part);
        FP_HIP(ctx, hipGetLastError());
    }
    if (part) {
        const int comps = mode == 0 ? 3 : 1;
        hipLaunchKernelGGL(k_trail_reduce, dim3((unsigned)((tv.M + 255) / 256)), dim3(256), 0, ctx->stream, part, sp,
                           tv.M, comps, out);
        FP_HIP(ctx, hipGetLastError());
    }
    return FP_OK;
}
