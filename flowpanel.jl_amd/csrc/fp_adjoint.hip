// fp_adjoint.hip -- kernels of the lift/drag adjoint (gfx950, FP64).
//
// Objective (the reference's offset-probe force workflow,
// test/runtests_solvers.jl:146-166 with src/FLOWPanel_postprocess.jl:28-118,
// 1059-1112, 1199-1307, 1615-1652):
//   probes X_i = centroid_i + offset * l_i * n_i,  V_i = Uinf + sum_j G_j u_j(X_i) + wake,
//   Cp_i = 1 - |V_i|^2/|Uinf|^2 (Kutta-averaged),  F = -q sum_i Cp_i A_i n_i,
//   CL = F.Lhat / (q S),  CDi = F.Dhat / (q S).
// Reverse pass, per objective m (CL, CDi):
//   * per panel: the adjoints of Cp, area and normal of the objective, Kutta
//     averaging reversed in reverse order, and the probe-velocity adjoint W1;
//   * pairs (probe target, source panel): one kernel evaluates the reverse of
//     every vortex ring at every probe -- the strength adjoint g_j = sum_i
//     u_j(X_i).W1_i, the vertex adjoints of panel j and the target adjoints;
//   * the transpose solve A^T lambda = g runs on the host side (same solvers);
//   * pairs (control point, source panel) again with weights -lambda_i n_i
//     (the adjoint of the AIC entries n_i . u_j(C_i) scaled by G_j);
//   * wake horseshoes like the rings, geometry reversed per panel from exact
//     directional derivatives, and a gather of the panel-vertex adjoints into
//     the mesh nodes.
// Layout: per-panel arrays are SoA over all bodies (M panels); the two
// objectives are stacked (component c = 3m + d).
#include "fp_internal.h"
#include "fp_adjoint.h"
#include "fp_geometry.h"

#include <algorithm>

#define ADJ_TT 64        // targets per LDS tile of the pair kernel
#define ADJ_BLOCK 256

namespace {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// geometry, control points and probes of one body into the concatenated arrays
template <int NVP>
__global__ __launch_bounds__(256) void k_adj_geometry(const double *__restrict__ pv, const int *__restrict__ nvv,
                                                      int64_t N, double cpoff, int charlen, double poff, int pkind,
                                                      double pscale, int64_t M, int64_t col0, double *__restrict__ nrm,
                                                      double *__restrict__ area, double *__restrict__ cpt,
                                                      double *__restrict__ prb)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double p[NVP][3];
#pragma unroll
    for (int k = 0; k < NVP; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) p[k][d] = pv[(3 * k + d) * N + j];
    const int nv = nvv[j];
    double t[3], n[3], o[3], m[3], a, len;
    panel_geom<double, NVP>(p, nv, charlen, t, n, o, m, a, len);
    const double lp = pscale * panel_len<double, NVP>(p, nv, pkind, a);
    const int64_t i = col0 + j;
    for (int d = 0; d < 3; ++d) {
        nrm[d * M + i] = n[d];
        cpt[d * M + i] = m[d] + cpoff * len * n[d];
        prb[d * M + i] = m[d] + poff * lp * n[d];
    }
    area[i] = a;
}

// V += Uinf; Cp = 1 - |V|^2/|Uinf|^2 (src/FLOWPanel_postprocess.jl:1059-1074).
// Velocities are point-major (V[3i + d]), as the field kernel writes them.
__global__ void k_adj_cp(double *__restrict__ V, int64_t M, double ux, double uy, double uz, double uref,
                         double *__restrict__ Cp)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double vx = V[3 * i] + ux, vy = V[3 * i + 1] + uy, vz = V[3 * i + 2] + uz;
    V[3 * i] = vx; V[3 * i + 1] = vy; V[3 * i + 2] = vz;
    const double u = sqrt(vx * vx + vy * vy + vz * vz) / uref;
    Cp[i] = 1 - u * u;
}

// panels averaged by trailing edge e (src/FLOWPanel_postprocess.jl:1075-1097)
__device__ __forceinline__ int kutta_ids(const int64_t *s, int mode, int64_t *ids)
{
    const int64_t pi = s[0], pj = s[3];
    if (mode == FP_KUTTA_REFERENCE) {
        ids[0] = pi; ids[1] = pi + 1;
        if (pj == -1) return 2;
        ids[2] = pj; ids[3] = pj - 1;
        return 4;
    }
    if (mode == FP_KUTTA_PAIR && pj != -1) { ids[0] = pi; ids[1] = pj; return 2; }
    return 0;
}

// sequential Kutta averaging (forward) of c0, or its reverse on c0 and c1
template <bool REV>
__global__ void k_adj_kutta(const int64_t *__restrict__ sh, int64_t ns, int mode, double *c0, double *c1)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t ids[4];
    if (!REV) {
        for (int64_t e = 0; e < ns; ++e) {
            const int cnt = kutta_ids(sh + 6 * e, mode, ids);
            if (!cnt) continue;
            double s = 0;
            for (int k = 0; k < cnt; ++k) s += c0[ids[k]];
            const double ave = s / cnt;
            for (int k = 0; k < cnt; ++k) c0[ids[k]] = ave;
        }
        return;
    }
    for (int64_t e = ns - 1; e >= 0; --e) {
        const int cnt = kutta_ids(sh + 6 * e, mode, ids);
        if (!cnt) continue;
        for (int pass = 0; pass < 2; ++pass) {
            double *c = pass == 0 ? c0 : c1;
            // adjoint of ave: the (distinct) overwritten entries
            double s = 0;
            for (int k = 0; k < cnt; ++k) {
                bool first = true;
                for (int q = 0; q < k; ++q) first = first && ids[q] != ids[k];
                if (first) s += c[ids[k]];
            }
            for (int k = 0; k < cnt; ++k) c[ids[k]] = 0;
            for (int k = 0; k < cnt; ++k) c[ids[k]] += s / cnt;
        }
    }
}

// terms of the force sum: T[d][i] = Cp_i A_i n_i
__global__ void k_adj_force_terms(const double *__restrict__ Cp, const double *__restrict__ area,
                                  const double *__restrict__ nrm, int64_t M, double *__restrict__ T)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double s = Cp[i] * area[i];
    T[i] = s * nrm[i];
    T[M + i] = s * nrm[M + i];
    T[2 * M + i] = s * nrm[2 * M + i];
}

// adjoints of the objectives J_m = -(1/S) sum_i Cp_i A_i n_i . e_m
__global__ void k_adj_objective(const double *__restrict__ Cp, const double *__restrict__ area,
                                const double *__restrict__ nrm, int64_t M, double e0x, double e0y, double e0z,
                                double e1x, double e1y, double e1z, double inv_s, double *__restrict__ Cpb,
                                double *__restrict__ Ab, double *__restrict__ nb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double nx = nrm[i], ny = nrm[M + i], nz = nrm[2 * M + i];
    const double e[2][3] = {{e0x, e0y, e0z}, {e1x, e1y, e1z}};
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const double ne = nx * e[m][0] + ny * e[m][1] + nz * e[m][2];
        Cpb[m * M + i] = -area[i] * ne * inv_s;
        Ab[m * M + i] = -Cp[i] * ne * inv_s;
#pragma unroll
        for (int d = 0; d < 3; ++d) nb[(3 * m + d) * M + i] = -Cp[i] * area[i] * e[m][d] * inv_s;
    }
}

// probe-velocity adjoints W1 = -2 Cpb V / |Uinf|^2 and the |V|^2-weighted
// Cp adjoints Q (columns 6, 7 of the same buffer) for the freestream adjoint
__global__ void k_adj_vbar(const double *__restrict__ Cpb, const double *__restrict__ V, int64_t M, double inv_u2,
                           double *__restrict__ W)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double vx = V[3 * i], vy = V[3 * i + 1], vz = V[3 * i + 2];
    const double v2 = vx * vx + vy * vy + vz * vz;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const double c = Cpb[m * M + i];
        W[(3 * m) * M + i] = -2 * c * vx * inv_u2;
        W[(3 * m + 1) * M + i] = -2 * c * vy * inv_u2;
        W[(3 * m + 2) * M + i] = -2 * c * vz * inv_u2;
        W[(6 + m) * M + i] = c * v2;
    }
}

// Reverse of every vortex ring of one body at every target of a range:
// lane = panel j, targets staged through LDS.  Per block (panel block bx,
// target split by):
//   xpart[bx][c][t]  = sum over the block's panels of G_j (du_j(X_t)/dX)^T w_t   (c = 3m+d)
//   vpart[by][.][j]  = G_j sum over the split's targets of (du_j/dp_k)^T w_t
//   gpart[by][m][j]  = sum over the split's targets of u_j(X_t) . w_t          (GAMMA)
template <int NVP, bool GAMMA>
__global__ __launch_bounds__(ADJ_BLOCK) void k_adj_pairs(const double *__restrict__ pv, int64_t N,
                                                         const double *__restrict__ gam, double off, double cut,
                                                         const double *__restrict__ tx, int64_t MT,
                                                         const double *__restrict__ W, int64_t tchunk,
                                                         double *__restrict__ xpart, double *__restrict__ vpart,
                                                         double *__restrict__ gpart)
{
    __shared__ double st[ADJ_TT][9];
    __shared__ double sx[ADJ_BLOCK / 64][ADJ_TT][6];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t j = (int64_t)blockIdx.x * ADJ_BLOCK + tid;
    const bool ok = j < N;
    double p[NVP][3];
#pragma unroll
    for (int k = 0; k < NVP; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) p[k][d] = ok ? pv[(3 * k + d) * N + j] : 0.0;
    // lanes past the end carry a degenerate ring (every edge of zero length)
    // and zero strength: they add nothing and take part in the wave sums
    const double gj = ok ? gam[j] : 0.0;
    double pb[2][NVP][3], g[2] = {0, 0};
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int k = 0; k < NVP; ++k) pb[m][k][0] = pb[m][k][1] = pb[m][k][2] = 0;
    const int64_t t0 = (int64_t)blockIdx.y * tchunk;
    const int64_t t1 = min(MT, t0 + tchunk);
    for (int64_t base = t0; base < t1; base += ADJ_TT) {
        const int cnt = (int)min((int64_t)ADJ_TT, t1 - base);
        __syncthreads();
        for (int e = tid; e < cnt * 9; e += ADJ_BLOCK) {
            const int t = e % cnt, c = e / cnt;
            st[t][c] = c < 3 ? tx[c * MT + base + t] : W[(c - 3) * MT + base + t];
        }
        __syncthreads();
        for (int t = 0; t < cnt; ++t) {
            const double w[2][3] = {{st[t][3], st[t][4], st[t][5]}, {st[t][6], st[t][7], st[t][8]}};
            double xb[2][3] = {{0, 0, 0}, {0, 0, 0}};
            fpa_ring_rev<NVP, 2>(p, st[t][0], st[t][1], st[t][2], off, cut, w, g, xb, pb);
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double v = wave_sum(gj * xb[m][d]);
                    if (lane == 0) sx[wave][t][3 * m + d] = v;
                }
        }
        __syncthreads();
        for (int e = tid; e < cnt * 6; e += ADJ_BLOCK) {
            const int t = e % cnt, c = e / cnt;
            double s = 0;
#pragma unroll
            for (int w = 0; w < ADJ_BLOCK / 64; ++w) s += sx[w][t][c];
            xpart[((int64_t)blockIdx.x * 6 + c) * MT + base + t] = s;
        }
    }
    if (!ok) return;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int k = 0; k < NVP; ++k)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                vpart[((int64_t)blockIdx.y * (6 * NVP) + (m * NVP + k) * 3 + d) * N + j] = gj * pb[m][k][d];
        if (GAMMA) gpart[((int64_t)blockIdx.y * 2 + m) * N + j] = g[m];
    }
}

// out[c*ldo + i] += sum_s part[(s*C + c)*n + i]
__global__ void k_adj_sum_parts(const double *__restrict__ part, int S, int C, int64_t n, double *__restrict__ out,
                                int64_t ldo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    double s = 0;
    for (int q = 0; q < S; ++q) s += part[((int64_t)q * C + c) * n + i];
    out[c * ldo + i] += s;
}

__device__ __forceinline__ void load_horse(const double *hh, int nh, int h, fpa_v3 &pa, fpa_v3 &pb, fpa_v3 &da,
                                           fpa_v3 &db)
{
    pa = fpa_mk(hh[0 * nh + h], hh[1 * nh + h], hh[2 * nh + h]);
    pb = fpa_mk(hh[3 * nh + h], hh[4 * nh + h], hh[5 * nh + h]);
    da = fpa_mk(hh[6 * nh + h], hh[7 * nh + h], hh[8 * nh + h]);
    db = fpa_mk(hh[9 * nh + h], hh[10 * nh + h], hh[11 * nh + h]);
}

// target adjoints of the horseshoes of one body: Xb[c][t] += G_h (du_h/dX)^T w_t
__global__ __launch_bounds__(256) void k_adj_wake_x(const double *__restrict__ hh, const int *__restrict__ hcol,
                                                    int nh, const double *__restrict__ gam, double off, double cut,
                                                    const double *__restrict__ tx, int64_t MT,
                                                    const double *__restrict__ W, double *__restrict__ Xb)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MT) return;
    const fpa_v3 X = fpa_mk(tx[t], tx[MT + t], tx[2 * MT + t]);
    fpa_v3 w[2], xb[2];
    for (int m = 0; m < 2; ++m) {
        w[m] = fpa_mk(W[(3 * m) * MT + t], W[(3 * m + 1) * MT + t], W[(3 * m + 2) * MT + t]);
        xb[m] = fpa_mk(0, 0, 0);
    }
    for (int h = 0; h < nh; ++h) {
        fpa_v3 pa, pb, da, db;
        load_horse(hh, nh, h, pa, pb, da, db);
        const double s = gam[hcol[h]];
        for (int m = 0; m < 2; ++m) {
            fpa_v3 a = fpa_mk(0, 0, 0), b = a, c = a, d = a;
            fpa_horse_rev(pa, pb, da, db, s, X, off, cut, w[m], a, b, c, d, xb[m]);
        }
    }
    for (int m = 0; m < 2; ++m) {
        Xb[(3 * m) * MT + t] += xb[m].x;
        Xb[(3 * m + 1) * MT + t] += xb[m].y;
        Xb[(3 * m + 2) * MT + t] += xb[m].z;
    }
}

// horseshoe parameter adjoints (unit strength): one block per horseshoe,
// out[h][26] = pa pb da db adjoints for m = 0 (12), m = 1 (12), and (GAMMA)
// the strength adjoints sum_t u_h(X_t).w_t (2)
template <bool GAMMA>
__global__ __launch_bounds__(256) void k_adj_wake_p(const double *__restrict__ hh, int nh, double off, double cut,
                                                    const double *__restrict__ tx, int64_t MT,
                                                    const double *__restrict__ W, double *__restrict__ out)
{
    __shared__ double red[4][26];
    const int h = blockIdx.x;
    fpa_v3 pa, pb, da, db;
    load_horse(hh, nh, h, pa, pb, da, db);
    double acc[26];
    for (int c = 0; c < 26; ++c) acc[c] = 0;
    for (int64_t t = threadIdx.x; t < MT; t += blockDim.x) {
        const fpa_v3 X = fpa_mk(tx[t], tx[MT + t], tx[2 * MT + t]);
        for (int m = 0; m < 2; ++m) {
            const fpa_v3 w = fpa_mk(W[(3 * m) * MT + t], W[(3 * m + 1) * MT + t], W[(3 * m + 2) * MT + t]);
            fpa_v3 a = fpa_mk(0, 0, 0), b = a, c = a, d = a, x = a;
            fpa_horse_rev(pa, pb, da, db, 1.0, X, off, cut, w, a, b, c, d, x);
            const double v[12] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z};
            for (int q = 0; q < 12; ++q) acc[12 * m + q] += v[q];
            if (GAMMA) acc[24 + m] += dot(fpa_horse(pa, pb, da, db, 1.0, X, off, cut), w);
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int c = 0; c < 26; ++c) {
        const double v = wave_sum(acc[c]);
        if (lane == 0) red[wave][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 26)
        out[(int64_t)h * 26 + threadIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// g[m][col0 + hcol[h]] += strength adjoint of horseshoe h (in order: deterministic)
__global__ void k_adj_wake_g(const double *__restrict__ hp, const int *__restrict__ hcol, int nh, int64_t col0,
                             int64_t M, double *__restrict__ g)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int h = 0; h < nh; ++h)
        for (int m = 0; m < 2; ++m) g[m * M + col0 + hcol[h]] += hp[(int64_t)h * 26 + 24 + m];
}

// control-point weights W2 = -lambda n, normal adjoints of the no-flow-through
// rows (b_i = -Uinf.n_i and A_ij = n_i.u_j(C_i)):  nb += -lambda (Uinf + Uind(C))
// (Uc point-major, Uc[3i + d])
__global__ void k_adj_cpw(const double *__restrict__ lam, const double *__restrict__ nrm,
                          const double *__restrict__ Uc, int64_t M, double ux, double uy, double uz,
                          double *__restrict__ W2, double *__restrict__ nb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double U[3] = {ux, uy, uz};
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const double l = lam[m * M + i];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            W2[(3 * m + d) * M + i] = -l * nrm[d * M + i];
            nb[(3 * m + d) * M + i] -= l * (U[d] + Uc[3 * i + d]);
        }
    }
}

// geometry reverse of one body: exact directional derivatives of (n, C, X,
// area) along each vertex coordinate, contracted with their adjoints, added
// to the panel-vertex adjoints; padded vertex slots fold into vertex 0.
template <int NVP>
__global__ __launch_bounds__(256) void k_adj_geom_rev(const double *__restrict__ pv, const int *__restrict__ nvv,
                                                      int64_t N, double cpoff, int charlen, double poff, int pkind,
                                                      double pscale, int64_t M, int64_t col0,
                                                      const double *__restrict__ nb, const double *__restrict__ Xbp,
                                                      const double *__restrict__ Xbc, const double *__restrict__ Ab,
                                                      double *__restrict__ vbar)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int64_t i = col0 + j;
    double p[NVP][3];
#pragma unroll
    for (int k = 0; k < NVP; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) p[k][d] = pv[(3 * k + d) * N + j];
    const int nv = nvv[j];
    double bn[2][3], bx[2][3], bc[2][3], ba[2];
    for (int m = 0; m < 2; ++m) {
        for (int d = 0; d < 3; ++d) {
            bn[m][d] = nb[(3 * m + d) * M + i];
            bx[m][d] = Xbp[(3 * m + d) * M + i];
            bc[m][d] = Xbc[(3 * m + d) * M + i];
        }
        ba[m] = Ab[m * M + i];
    }
    for (int k = 0; k < NVP; ++k) {
        if (k >= nv) break;
        for (int d = 0; d < 3; ++d) {
            fpg_dual q[NVP][3], t[3], n[3], o[3], mm[3], area, len;
            for (int a = 0; a < NVP; ++a)
                for (int b = 0; b < 3; ++b) q[a][b] = fpg_mk(p[a][b], (a == k && b == d) ? 1.0 : 0.0);
            panel_geom<fpg_dual, NVP>(q, nv, charlen, t, n, o, mm, area, len);
            const fpg_dual lp = pscale * panel_len<fpg_dual, NVP>(q, nv, pkind, area);
            for (int m = 0; m < 2; ++m) {
                double s = ba[m] * area.d;
                for (int c = 0; c < 3; ++c) {
                    const double dC = mm[c].d + cpoff * (len.d * n[c].v + len.v * n[c].d);
                    const double dX = mm[c].d + poff * (lp.d * n[c].v + lp.v * n[c].d);
                    s += bn[m][c] * n[c].d + bc[m][c] * dC + bx[m][c] * dX;
                }
                vbar[((m * NVP + k) * 3 + d) * N + j] += s;
            }
        }
    }
    for (int m = 0; m < 2; ++m)
        for (int k = nv; k < NVP; ++k)
            for (int d = 0; d < 3; ++d) {
                double &src = vbar[((m * NVP + k) * 3 + d) * N + j];
                vbar[((m * NVP + 0) * 3 + d) * N + j] += src;
                src = 0;
            }
}

// node adjoints: nodebar[c][v] = sum over the panel-vertex slots of node v
template <int NVP>
__global__ void k_adj_nodes(const int64_t *__restrict__ ptr, const int64_t *__restrict__ ent, int64_t nn, int64_t N,
                            const double *__restrict__ vbar, double *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e) {
        const int64_t k = ent[e] / N, j = ent[e] % N;
        for (int m = 0; m < 2; ++m)
            for (int d = 0; d < 3; ++d) acc[3 * m + d] += vbar[((m * NVP + k) * 3 + d) * N + j];
    }
    for (int c = 0; c < 6; ++c) out[c * nn + v] = acc[c];
}

// LDS-tiled transpose (and optional narrowing) of a row-major FP64 matrix
template <typename TD>
__global__ __launch_bounds__(256) void k_transpose(const double *__restrict__ A, int64_t n, int64_t ld,
                                                   TD *__restrict__ B)
{
    __shared__ double tile[64][65];
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int64_t i = r0 + r, jj = c0 + tx;
        tile[r][tx] = (i < n && jj < n) ? A[i * ld + jj] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int64_t i = c0 + r, jj = r0 + tx;  // B(i, jj) = A(jj, i)
        if (i < n && jj < n) B[i * ld + jj] = (TD)tile[tx][r];
    }
}

int launch_check(fp_ctx *ctx) { FP_HIP(ctx, hipGetLastError()); return FP_OK; }

unsigned blocks(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ============================================================================
// launchers

int fp_adj_geometry(fp_ctx *ctx, const fp_body *b, double poff, int pkind, double pscale, int64_t M, int64_t col0,
                    double *nrm, double *area, double *cpt, double *prb)
{
    if (b->N == 0) return FP_OK;
    if (b->nvp == 4)
        hipLaunchKernelGGL(k_adj_geometry<4>, dim3(blocks(b->N)), dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, b->N,
                           b->cp_offset, b->charlen, poff, pkind, pscale, M, col0, nrm, area, cpt, prb);
    else
        hipLaunchKernelGGL(k_adj_geometry<8>, dim3(blocks(b->N)), dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, b->N,
                           b->cp_offset, b->charlen, poff, pkind, pscale, M, col0, nrm, area, cpt, prb);
    return launch_check(ctx);
}

int fp_adj_cp(fp_ctx *ctx, double *V, int64_t M, const double *U, double *Cp)
{
    const double uref = sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
    hipLaunchKernelGGL(k_adj_cp, dim3(blocks(M)), dim3(256), 0, ctx->stream, V, M, U[0], U[1], U[2], uref, Cp);
    return launch_check(ctx);
}

int fp_adj_kutta(fp_ctx *ctx, const int64_t *d_sh, int64_t ns, int mode, double *c0, double *c1, int reverse)
{
    if (ns == 0 || mode == FP_KUTTA_NONE) return FP_OK;
    if (reverse) hipLaunchKernelGGL(k_adj_kutta<true>, dim3(1), dim3(64), 0, ctx->stream, d_sh, ns, mode, c0, c1);
    else hipLaunchKernelGGL(k_adj_kutta<false>, dim3(1), dim3(64), 0, ctx->stream, d_sh, ns, mode, c0, c1);
    return launch_check(ctx);
}

int fp_adj_force_terms(fp_ctx *ctx, const double *Cp, const double *area, const double *nrm, int64_t M, double *T)
{
    hipLaunchKernelGGL(k_adj_force_terms, dim3(blocks(M)), dim3(256), 0, ctx->stream, Cp, area, nrm, M, T);
    return launch_check(ctx);
}

int fp_adj_objective(fp_ctx *ctx, const double *Cp, const double *area, const double *nrm, int64_t M,
                     const double *e0, const double *e1, double inv_s, double *Cpb, double *Ab, double *nb)
{
    hipLaunchKernelGGL(k_adj_objective, dim3(blocks(M)), dim3(256), 0, ctx->stream, Cp, area, nrm, M, e0[0], e0[1],
                       e0[2], e1[0], e1[1], e1[2], inv_s, Cpb, Ab, nb);
    return launch_check(ctx);
}

int fp_adj_vbar(fp_ctx *ctx, const double *Cpb, const double *V, int64_t M, double inv_u2, double *W)
{
    hipLaunchKernelGGL(k_adj_vbar, dim3(blocks(M)), dim3(256), 0, ctx->stream, Cpb, V, M, inv_u2, W);
    return launch_check(ctx);
}

// split the targets so that the grid has about 4 blocks per CU
static void pair_grid(fp_ctx *ctx, int64_t N, int64_t MT, int &px, int &splits, int64_t &tchunk)
{
    px = (int)((N + ADJ_BLOCK - 1) / ADJ_BLOCK);
    const int want = std::max(1, (4 * ctx->num_cu + px - 1) / px);
    const int64_t tiles = (MT + ADJ_TT - 1) / ADJ_TT;
    splits = (int)std::max<int64_t>(1, std::min<int64_t>(want, tiles));
    tchunk = ((tiles + splits - 1) / splits) * ADJ_TT;
    splits = (int)((MT + tchunk - 1) / tchunk);
}

size_t fp_adj_pairs_work_bytes(fp_ctx *ctx, int64_t N, int nvp, int64_t MT)
{
    int px, splits;
    int64_t tchunk;
    pair_grid(ctx, N, MT, px, splits, tchunk);
    return sizeof(double) * ((size_t)px * 6 * MT + (size_t)splits * 6 * nvp * N + (size_t)splits * 2 * N);
}

int fp_adj_pairs(fp_ctx *ctx, const fp_body *b, const double *tx, int64_t MT, const double *W, int gamma,
                 double *Xb, double *vbar, double *g, int64_t ldg, void *work, size_t work_bytes)
{
    const int64_t N = b->N;
    if (N == 0 || MT == 0) return FP_OK;
    int px, splits;
    int64_t tchunk;
    pair_grid(ctx, N, MT, px, splits, tchunk);
    if (work_bytes < fp_adj_pairs_work_bytes(ctx, N, b->nvp, MT))
        return fp_fail(ctx, FP_ERR_ARG, "internal: adjoint pair work buffer too small");
    double *xpart = (double *)work;
    double *vpart = xpart + (size_t)px * 6 * MT;
    double *gpart = vpart + (size_t)splits * 6 * b->nvp * N;
    {
        ProfScope ps(ctx, gamma ? "adjoint_pairs_probe_f64" : "adjoint_pairs_cp_f64");
        dim3 grid((unsigned)px, (unsigned)splits);
        const double off = b->kernel_offset, cut = b->kernel_cutoff;
        if (b->nvp == 4) {
            if (gamma) hipLaunchKernelGGL((k_adj_pairs<4, true>), grid, dim3(ADJ_BLOCK), 0, ctx->stream, b->d_pv, N,
                                          b->d_strength, off, cut, tx, MT, W, tchunk, xpart, vpart, gpart);
            else hipLaunchKernelGGL((k_adj_pairs<4, false>), grid, dim3(ADJ_BLOCK), 0, ctx->stream, b->d_pv, N,
                                    b->d_strength, off, cut, tx, MT, W, tchunk, xpart, vpart, gpart);
        } else {
            if (gamma) hipLaunchKernelGGL((k_adj_pairs<8, true>), grid, dim3(ADJ_BLOCK), 0, ctx->stream, b->d_pv, N,
                                          b->d_strength, off, cut, tx, MT, W, tchunk, xpart, vpart, gpart);
            else hipLaunchKernelGGL((k_adj_pairs<8, false>), grid, dim3(ADJ_BLOCK), 0, ctx->stream, b->d_pv, N,
                                    b->d_strength, off, cut, tx, MT, W, tchunk, xpart, vpart, gpart);
        }
        int st = launch_check(ctx);
        if (st) return st;
    }
    hipLaunchKernelGGL(k_adj_sum_parts, dim3(blocks(MT), 6), dim3(256), 0, ctx->stream, (const double *)xpart, px, 6,
                       MT, Xb, MT);
    hipLaunchKernelGGL(k_adj_sum_parts, dim3(blocks(N), 6 * b->nvp), dim3(256), 0, ctx->stream,
                       (const double *)vpart, splits, 6 * b->nvp, N, vbar, N);
    if (gamma)
        hipLaunchKernelGGL(k_adj_sum_parts, dim3(blocks(N), 2), dim3(256), 0, ctx->stream, (const double *)gpart,
                           splits, 2, N, g, ldg);
    return launch_check(ctx);
}

int fp_adj_wake(fp_ctx *ctx, const fp_body *b, const double *tx, int64_t MT, const double *W, int gamma, double *Xb,
                double *hpart, double *g, int64_t col0)
{
    if (b->nh == 0 || MT == 0) return FP_OK;
    ProfScope ps(ctx, "adjoint_wake_f64");
    const double off = b->kernel_offset, cut = b->kernel_cutoff;
    hipLaunchKernelGGL(k_adj_wake_x, dim3(blocks(MT)), dim3(256), 0, ctx->stream, b->d_hh, b->d_hcol, b->nh,
                       b->d_strength, off, cut, tx, MT, W, Xb);
    if (gamma) {
        hipLaunchKernelGGL(k_adj_wake_p<true>, dim3(b->nh), dim3(256), 0, ctx->stream, b->d_hh, b->nh, off, cut, tx,
                           MT, W, hpart);
        hipLaunchKernelGGL(k_adj_wake_g, dim3(1), dim3(64), 0, ctx->stream, (const double *)hpart, b->d_hcol, b->nh,
                           col0, MT, g);
    } else {
        hipLaunchKernelGGL(k_adj_wake_p<false>, dim3(b->nh), dim3(256), 0, ctx->stream, b->d_hh, b->nh, off, cut, tx,
                           MT, W, hpart);
    }
    return launch_check(ctx);
}

int fp_adj_cpw(fp_ctx *ctx, const double *lam, const double *nrm, const double *Uc, int64_t M, const double *U,
               double *W2, double *nb)
{
    hipLaunchKernelGGL(k_adj_cpw, dim3(blocks(M)), dim3(256), 0, ctx->stream, lam, nrm, Uc, M, U[0], U[1], U[2], W2,
                       nb);
    return launch_check(ctx);
}

int fp_adj_geom_rev(fp_ctx *ctx, const fp_body *b, double poff, int pkind, double pscale, int64_t M, int64_t col0,
                    const double *nb, const double *Xbp, const double *Xbc, const double *Ab, double *vbar)
{
    if (b->N == 0) return FP_OK;
    if (b->nvp == 4)
        hipLaunchKernelGGL(k_adj_geom_rev<4>, dim3(blocks(b->N)), dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, b->N,
                           b->cp_offset, b->charlen, poff, pkind, pscale, M, col0, nb, Xbp, Xbc, Ab, vbar);
    else
        hipLaunchKernelGGL(k_adj_geom_rev<8>, dim3(blocks(b->N)), dim3(256), 0, ctx->stream, b->d_pv, b->d_nv, b->N,
                           b->cp_offset, b->charlen, poff, pkind, pscale, M, col0, nb, Xbp, Xbc, Ab, vbar);
    return launch_check(ctx);
}

int fp_adj_nodes(fp_ctx *ctx, const fp_body *b, const int64_t *ptr, const int64_t *ent, int64_t nn,
                 const double *vbar, double *out)
{
    if (nn == 0) return FP_OK;
    if (b->nvp == 4)
        hipLaunchKernelGGL(k_adj_nodes<4>, dim3(blocks(nn)), dim3(256), 0, ctx->stream, ptr, ent, nn, b->N, vbar, out);
    else
        hipLaunchKernelGGL(k_adj_nodes<8>, dim3(blocks(nn)), dim3(256), 0, ctx->stream, ptr, ent, nn, b->N, vbar, out);
    return launch_check(ctx);
}

int fp_mat_transpose(fp_ctx *ctx, const double *A, int64_t n, int64_t ld, void *B, int dst_prec)
{
    if (n <= 0) return FP_OK;
    ProfScope ps(ctx, "transpose");
    dim3 grid((unsigned)((n + 63) / 64), (unsigned)((n + 63) / 64));
    if (dst_prec == FP_F64)
        hipLaunchKernelGGL(k_transpose<double>, grid, dim3(256), 0, ctx->stream, A, n, ld, (double *)B);
    else
        hipLaunchKernelGGL(k_transpose<float>, grid, dim3(256), 0, ctx->stream, A, n, ld, (float *)B);
    return launch_check(ctx);
}
