// fp_elements.h -- device element kernels (FP32/FP64 templates).
//
// Each function restates one element of src/FLOWPanel_elements.jl for a single
// target held in registers.  T is the arithmetic type (float or double).
#pragma once

#include <hip/hip_runtime.h>

template <typename T>
__device__ __forceinline__ T fp_sign(T z) { return z > T(0) ? T(1) : (z < T(0) ? T(-1) : z); }

// 1/x: the hardware reciprocal refined by two Newton steps, r <- r + r (1 - x r)
// (a few ulp at most, where the IEEE division sequence is correctly rounded but
// about twice the instructions).  0, infinities and NaN keep the raw reciprocal.
__device__ __forceinline__ double fp_recip(double x)
{
    const double r0 = __builtin_amdgcn_rcp(x);
    double r = fma(r0, fma(-x, r0, 1.0), r0);
    r = fma(r, fma(-x, r, 1.0), r);
    return __builtin_isfinite(r) ? r : r0;
}
__device__ __forceinline__ float fp_recip(float x) { return 1.0f / x; }

// sqrt(x) for x >= 0 as x * rsqrt(x), the hardware reciprocal square root
// refined by two Newton steps y <- y (3 - x y^2) / 2 (a few ulp at most, against
// the longer correctly rounded sequence); x below 2^-1000 (including 0) and
// non-finite x take the library sqrt.
__device__ __forceinline__ double fp_sqrt(double x)
{
    if (!(x > 0x1p-1000 && x < 0x1p1000)) return sqrt(x);
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return x * y;
}
__device__ __forceinline__ float fp_sqrt(float x) { return sqrtf(x); }

// sqrt(x) for 0 <= x < 2^1000 without a branch, for distances to which a
// kernel offset is added: the hardware reciprocal square root of
// max(x, 2^-1000) refined by two Newton steps, times x (a few ulp for
// x > 2^-1000; 0 for x = 0; below 2^-1000 a value under 2^-500).
__device__ __forceinline__ double fp_sqrt_nb(double x)
{
    const double xs = fmax(x, 0x1p-1000);
    double y = __builtin_amdgcn_rsq(xs);
    const double h = 0.5 * xs;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return x * y;
}

// Regularised Biot-Savart segment a -> b (src/FLOWPanel_elements.jl:345-428),
// velocity per unit strength added to u.  e = b - a.
template <typename T>
__device__ __forceinline__ void seg_u(T ax, T ay, T az, T ex, T ey, T ez, T bx, T by, T bz, T X, T Y, T Z,
                                      T s, T off, T off2, T cut2, T &ux, T &uy, T &uz)
{
    const T rix = X - ax, riy = Y - ay, riz = Z - az;
    const T rjx = X - bx, rjy = Y - by, rjz = Z - bz;
    const T cx = riy * rjz - riz * rjy;
    const T cy = riz * rjx - rix * rjz;
    const T cz = rix * rjy - riy * rjx;
    const T c2 = cx * cx + cy * cy + cz * cz;
    const T ii = fp_recip(fp_sqrt(rix * rix + riy * riy + riz * riz) + off);
    const T ij = fp_recip(fp_sqrt(rjx * rjx + rjy * rjy + rjz * rjz) + off);
    const T proj = ex * (rix * ii - rjx * ij) + ey * (riy * ii - rjy * ij) + ez * (riz * ii - rjz * ij);
    if (c2 > cut2) {
        const T aux = s * proj * fp_recip(T(4 * FP_PI) * (c2 + off2));
        ux += aux * cx; uy += aux * cy; uz += aux * cz;
    }
}

// Vortex ring = constant doublet (src/FLOWPanel_elements.jl:442-467): sum of
// the NVP segments (k, k+1 mod NVP).  Node distances are shared by the two
// edges that meet at a node.  p* vertices, e* edge vectors p[k+1]-p[k].
template <typename T, int NVP>
__device__ __forceinline__ void ring_u(const T *px, const T *py, const T *pz, const T *ex, const T *ey,
                                       const T *ez, T X, T Y, T Z, T off, T off2, T cut2, T &ux, T &uy,
                                       T &uz)
{
    T rx[NVP], ry[NVP], rz[NVP], iv[NVP];
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        rx[k] = X - px[k]; ry[k] = Y - py[k]; rz[k] = Z - pz[k];
        iv[k] = fp_recip(fp_sqrt(rx[k] * rx[k] + ry[k] * ry[k] + rz[k] * rz[k]) + off);
    }
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        const int q = (k + 1) % NVP;
        const T cx = ry[k] * rz[q] - rz[k] * ry[q];
        const T cy = rz[k] * rx[q] - rx[k] * rz[q];
        const T cz = rx[k] * ry[q] - ry[k] * rx[q];
        const T c2 = cx * cx + cy * cy + cz * cz;
        const T proj = ex[k] * (rx[k] * iv[k] - rx[q] * iv[q]) + ey[k] * (ry[k] * iv[k] - ry[q] * iv[q]) +
                       ez[k] * (rz[k] * iv[k] - rz[q] * iv[q]);
        if (c2 > cut2) {
            const T aux = proj * fp_recip(T(4 * FP_PI) * (c2 + off2));
            ux += aux * cx; uy += aux * cy; uz += aux * cz;
        }
    }
}

// Semi-infinite vortex from p along unit d (src/FLOWPanel_elements.jl:480-628).
template <typename T>
__device__ __forceinline__ void semiinf_u(T px, T py, T pz, T dx, T dy, T dz, T s, T X, T Y, T Z, T off,
                                          T off2, T cut2, T &ux, T &uy, T &uz)
{
    const T xm = (X - px) * dx + (Y - py) * dy + (Z - pz) * dz;
    const T p0x = px + xm * dx, p0y = py + xm * dy, p0z = pz + xm * dz;
    const T bx = p0x - px, by = p0y - py, bz = p0z - pz;
    if (bx * bx + by * by + bz * bz > off2)
        seg_u<T>(px, py, pz, bx, by, bz, p0x, p0y, p0z, X, Y, Z, s, off, off2, cut2, ux, uy, uz);
    const T hx = X - p0x, hy = Y - p0y, hz = Z - p0z;
    const T h2 = hx * hx + hy * hy + hz * hz;
    if (h2 > cut2) {
        const T aux = s / (T(4 * FP_PI) * (h2 + off2));
        ux += aux * (dy * hz - dz * hy);
        uy += aux * (dz * hx - dx * hz);
        uz += aux * (dx * hy - dy * hx);
    }
}

// Horseshoe (src/FLOWPanel_elements.jl:692-730): infinity -> pa along da with
// -s, bound pa -> pb with s, pb -> infinity along db with s.
template <typename T>
__device__ __forceinline__ void horseshoe_u(const T *h, T X, T Y, T Z, T s, T off, T off2, T cut2, T &ux, T &uy,
                                            T &uz)
{
    // h = pa(3) pb(3) da(3) db(3)
    semiinf_u<T>(h[0], h[1], h[2], h[6], h[7], h[8], -s, X, Y, Z, off, off2, cut2, ux, uy, uz);
    seg_u<T>(h[0], h[1], h[2], h[3] - h[0], h[4] - h[1], h[5] - h[2], h[3], h[4], h[5], X, Y, Z, s, off, off2,
             cut2, ux, uy, uz);
    semiinf_u<T>(h[3], h[4], h[5], h[9], h[10], h[11], s, X, Y, Z, off, off2, cut2, ux, uy, uz);
}

// Hess & Smith planar-polygon sums for a target in panel-local coordinates
// (src/FLOWPanel_elements.jl:79-121 velocity, :199-236 source potential,
// :798-838 doublet potential).  Table layout (stride 1 here):
//   t[0..11] frame, t[12+k] xl, t[12+NVP+k] yl, t[12+2NVP+k] zl,
//   t[12+3NVP+k] d, t[12+4NVP+k] S, t[12+5NVP+k] C.
// Returns V1, V2 (source only), V3 (after sign/nR0 handling, before 1/4pi)
// and the source potential sum phis (before 1/4pi).
template <typename T, int NVP, bool SOURCE>
__device__ __forceinline__ void hs_sums(const T *xl, const T *yl, const T *zl, const T *dd, const T *SS,
                                        const T *CC, int nv, T x, T y, T z, T off, T &V1, T &V2, T &V3,
                                        T &phis)
{
    T r[NVP];
#pragma unroll
    for (int k = 0; k < NVP; ++k)
        r[k] = sqrt((x - xl[k]) * (x - xl[k]) + (y - yl[k]) * (y - yl[k]) + (z - zl[k]) * (z - zl[k]));
    const T az = fabs(z);
    T dtheta = T(2 * FP_PI);
    int nR0 = 0;
    T v1 = 0, v2 = 0, v3 = 0, ps = 0;
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        if (k < nv) {
            const int q = (k + 1) % NVP;
            const T S = SS[k], C = CC[k];
            const T si = (xl[k] - x) * C + (yl[k] - y) * S;
            const T sj = (xl[q] - x) * C + (yl[q] - y) * S;
            const T R = (x - xl[k]) * S - (y - yl[k]) * C;
            const T J = atan2(R * az * (r[k] * sj - r[q] * si), r[k] * r[q] * R * R + z * z * sj * si);
            if (SOURCE) {
                const T Q = log((r[k] + r[q] + dd[k]) / (r[k] + r[q] - dd[k] + off));
                v1 -= S * Q;
                v2 += C * Q;
                ps -= R * Q + az * J;
            }
            v3 -= J;
            dtheta *= (R >= T(0)) ? T(1) : T(0);
            nR0 += (R == T(0));
        }
    }
    phis = ps + az * dtheta;
    v3 += dtheta;
    v3 *= fp_sign(z);
    v3 *= (nR0 > 1) ? T(0) : T(1);
    V1 = v1; V2 = v2; V3 = v3;
}

// Solid angle of the triangle (a, b, c) in the frame (ex, ey, ez, origin o)
// divided by 4 pi: the constant-doublet potential of a triangle
// (src/FLOWPanel_elements.jl:1016-1068 and :1193-1238).
template <typename T>
__device__ __forceinline__ T tri_doublet_phi(const T *P /*9: a b c*/, const T *ex, const T *ey, const T *ez,
                                             const T *o, T X, T Y, T Z)
{
    T xl[3], yl[3], zl[3], dd[3], SS[3], CC[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T dx = P[3 * k] - o[0], dy = P[3 * k + 1] - o[1], dz = P[3 * k + 2] - o[2];
        xl[k] = ex[0] * dx + ex[1] * dy + ex[2] * dz;
        yl[k] = ey[0] * dx + ey[1] * dy + ey[2] * dz;
        zl[k] = ez[0] * dx + ez[1] * dy + ez[2] * dz;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int q = (k + 1) % 3;
        dd[k] = sqrt((xl[q] - xl[k]) * (xl[q] - xl[k]) + (yl[q] - yl[k]) * (yl[q] - yl[k]) +
                     (zl[q] - zl[k]) * (zl[q] - zl[k]));
        SS[k] = (yl[q] - yl[k]) / dd[k];
        CC[k] = (xl[q] - xl[k]) / dd[k];
    }
    const T dx = X - o[0], dy = Y - o[1], dz = Z - o[2];
    const T x = ex[0] * dx + ex[1] * dy + ex[2] * dz;
    const T y = ey[0] * dx + ey[1] * dy + ey[2] * dz;
    const T z = ez[0] * dx + ez[1] * dy + ez[2] * dz;
    T V1, V2, V3, ps;
    hs_sums<T, 3, false>(xl, yl, zl, dd, SS, CC, 3, x, y, z, T(0), V1, V2, V3, ps);
    return V3 / T(4 * FP_PI);
}

// Semi-infinite doublet potential (src/FLOWPanel_elements.jl:959-1092 planar
// part along da, :1110-1244 extra triangle (pda, pb, pdb) when da != db).
// h = pa(3) pb(3) da(3) db(3).  Returns phi per unit strength.
template <typename T>
__device__ __forceinline__ T semiinf_doublet_phi(const T *h, T X, T Y, T Z)
{
    const T *pi = h, *pj = h + 3, *d = h + 6, *db = h + 9;
    const T eps2 = T(2) * (sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-07));
    const T pd = (pj[0] - pi[0]) * d[0] + (pj[1] - pi[1]) * d[1] + (pj[2] - pi[2]) * d[2];
    T pa[3] = {pi[0] + pd * d[0], pi[1] + pd * d[1], pi[2] + pd * d[2]};
    T ey[3] = {pj[0] - pa[0], pj[1] - pa[1], pj[2] - pa[2]};
    const T ny = sqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
    ey[0] /= ny; ey[1] /= ny; ey[2] /= ny;
    const T ex[3] = {d[0], d[1], d[2]};
    const T ez[3] = {ex[1] * ey[2] - ex[2] * ey[1], ex[2] * ey[0] - ex[0] * ey[2], ex[0] * ey[1] - ex[1] * ey[0]};
    T val = 0;
    if (fabs(pd) > eps2) {
        const T P[9] = {pa[0], pa[1], pa[2], pi[0], pi[1], pi[2], pj[0], pj[1], pj[2]};
        val += tri_doublet_phi<T>(P, ex, ey, ez, pi, X, Y, Z);
    }
    // semi-infinite part, frame origin pi
    auto loc = [&](const T *e, T ax, T ay, T az) { return e[0] * (ax - pi[0]) + e[1] * (ay - pi[1]) + e[2] * (az - pi[2]); };
    const T x = loc(ex, X, Y, Z), y = loc(ey, X, Y, Z), z = loc(ez, X, Y, Z);
    const T xa = loc(ex, pa[0], pa[1], pa[2]), ya = loc(ey, pa[0], pa[1], pa[2]);
    const T yb = loc(ey, pj[0], pj[1], pj[2]);
    const T dxa = x - xa;
    T s = atan((yb - y) / z) + atan((yb - y) * dxa / (z * sqrt(dxa * dxa + (yb - y) * (yb - y) + z * z)));
    s -= atan((ya - y) / z) + atan((ya - y) * dxa / (z * sqrt(dxa * dxa + (ya - y) * (ya - y) + z * z)));
    val += s / T(4 * FP_PI);
    // non-planar correction triangle
    if (fabs(d[0] - db[0]) > eps2 || fabs(d[1] - db[1]) > eps2 || fabs(d[2] - db[2]) > eps2) {
        const T L = T(1e12);
        const T *pb = pj;
        const T P[9] = {pb[0] + L * d[0], pb[1] + L * d[1], pb[2] + L * d[2], pb[0], pb[1], pb[2],
                        pb[0] + L * db[0], pb[1] + L * db[1], pb[2] + L * db[2]};
        T zz[3] = {d[1] * db[2] - d[2] * db[1], d[2] * db[0] - d[0] * db[2], d[0] * db[1] - d[1] * db[0]};
        const T nz = sqrt(zz[0] * zz[0] + zz[1] * zz[1] + zz[2] * zz[2]);
        zz[0] /= nz; zz[1] /= nz; zz[2] /= nz;
        const T yy[3] = {zz[1] * d[2] - zz[2] * d[1], zz[2] * d[0] - zz[0] * d[2], zz[0] * d[1] - zz[1] * d[0]};
        val += tri_doublet_phi<T>(P, ex, yy, zz, pb, X, Y, Z);
    }
    return val;
}
