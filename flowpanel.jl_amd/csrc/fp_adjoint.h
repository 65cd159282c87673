// fp_adjoint.h -- reverse-mode (adjoint) forms of the vortex elements.
//
// Each function takes the forward inputs and the adjoint ub of the velocity
// contribution u (3-vector) and ADDS the adjoints of its inputs.  Branch
// decisions of the forward (cutoff / offset tests) are taken as fixed, as in
// any algorithmic derivative.  The forward formulas are those of
// src/FLOWPanel_elements.jl:345-428 (bound vortex), :442-467 (vortex ring),
// :480-628 (semi-infinite), :692-730 (horseshoe).  The functions are
// host+device so that a host build can check every reverse function against
// finite differences of its forward.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FPA_HD __host__ __device__ __forceinline__
#else
#define FPA_HD inline
#endif
#include <math.h>

#define FPA_4PI 12.566370614359172953850573533118

struct fpa_v3 {
    double x, y, z;
};
FPA_HD fpa_v3 fpa_mk(double x, double y, double z) { fpa_v3 r = {x, y, z}; return r; }
FPA_HD fpa_v3 operator+(fpa_v3 a, fpa_v3 b) { return fpa_mk(a.x + b.x, a.y + b.y, a.z + b.z); }
FPA_HD fpa_v3 operator-(fpa_v3 a, fpa_v3 b) { return fpa_mk(a.x - b.x, a.y - b.y, a.z - b.z); }
FPA_HD fpa_v3 operator*(double s, fpa_v3 a) { return fpa_mk(s * a.x, s * a.y, s * a.z); }
FPA_HD double dot(fpa_v3 a, fpa_v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FPA_HD fpa_v3 cross(fpa_v3 a, fpa_v3 b)
{
    return fpa_mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
FPA_HD void acc(fpa_v3 &a, fpa_v3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }

// forward: regularised segment a -> b at X, per unit strength (same as seg_u)
FPA_HD fpa_v3 fpa_seg(fpa_v3 a, fpa_v3 b, fpa_v3 X, double off, double cut)
{
    const fpa_v3 ri = X - a, rj = X - b, e = b - a;
    const fpa_v3 c = cross(ri, rj);
    const double c2 = dot(c, c);
    const double ni = sqrt(dot(ri, ri)) + off, nj = sqrt(dot(rj, rj)) + off;
    const double proj = dot(e, (1.0 / ni) * ri - (1.0 / nj) * rj);
    if (!(c2 > cut * cut)) return fpa_mk(0, 0, 0);
    return (proj / (FPA_4PI * (c2 + off * off))) * c;
}

// reverse of the segment: ub = adjoint of u; adds abar, bbar, Xbar
FPA_HD void fpa_seg_rev(fpa_v3 a, fpa_v3 b, fpa_v3 X, double off, double cut, fpa_v3 ub, fpa_v3 &abar,
                        fpa_v3 &bbar, fpa_v3 &Xbar)
{
    const fpa_v3 ri = X - a, rj = X - b, e = b - a;
    const fpa_v3 c = cross(ri, rj);
    const double c2 = dot(c, c);
    if (!(c2 > cut * cut)) return;
    const double li = sqrt(dot(ri, ri)), lj = sqrt(dot(rj, rj));
    const double ni = li + off, nj = lj + off;
    const fpa_v3 vi = (1.0 / ni) * ri, vj = (1.0 / nj) * rj;
    const double proj = dot(e, vi - vj);
    const double den = c2 + off * off;
    const double K = 1.0 / (FPA_4PI * den);
    // u = proj K c
    fpa_v3 cb = (proj * K) * ub;
    const double sb = dot(ub, c);      // adjoint of proj*K
    const double projb = sb * K;
    const double Kb = sb * proj;
    const double c2b = -Kb * K / den;  // dK/dc2 = -K/den
    acc(cb, (2.0 * c2b) * c);
    // c = ri x rj
    fpa_v3 rib = cross(rj, cb);
    fpa_v3 rjb = cross(cb, ri);
    // proj = e . (vi - vj)
    const fpa_v3 eb = projb * (vi - vj);
    const fpa_v3 vib = projb * e, vjb = (-projb) * e;
    // v = r/(|r|+off): dv = dr/n - r (r.dr)/(|r| n^2)
    if (li > 0) acc(rib, (1.0 / ni) * vib - (dot(ri, vib) / (li * ni * ni)) * ri);
    if (lj > 0) acc(rjb, (1.0 / nj) * vjb - (dot(rj, vjb) / (lj * nj * nj)) * rj);
    // ri = X - a, rj = X - b, e = b - a
    acc(Xbar, rib + rjb);
    acc(abar, (-1.0) * (rib + eb));
    acc(bbar, eb - rjb);
}

// forward: semi-infinite vortex from p along unit d, strength s
FPA_HD fpa_v3 fpa_semi(fpa_v3 p, fpa_v3 d, double s, fpa_v3 X, double off, double cut)
{
    const double xm = dot(X - p, d);
    const fpa_v3 p0 = p + xm * d;
    const fpa_v3 bv = p0 - p;
    fpa_v3 u = fpa_mk(0, 0, 0);
    if (dot(bv, bv) > off * off) u = s * fpa_seg(p, p0, X, off, cut);
    const fpa_v3 h = X - p0;
    const double h2 = dot(h, h);
    if (h2 > cut * cut) acc(u, (s / (FPA_4PI * (h2 + off * off))) * cross(d, h));
    return u;
}

// reverse of the semi-infinite vortex: adds pbar, dbar, Xbar
FPA_HD void fpa_semi_rev(fpa_v3 p, fpa_v3 d, double s, fpa_v3 X, double off, double cut, fpa_v3 ub, fpa_v3 &pbar,
                         fpa_v3 &dbar, fpa_v3 &Xbar)
{
    const double xm = dot(X - p, d);
    const fpa_v3 p0 = p + xm * d;
    const fpa_v3 bv = p0 - p;
    fpa_v3 p0b = fpa_mk(0, 0, 0);
    if (dot(bv, bv) > off * off) fpa_seg_rev(p, p0, X, off, cut, s * ub, pbar, p0b, Xbar);
    const fpa_v3 h = X - p0;
    const double h2 = dot(h, h);
    if (h2 > cut * cut) {
        const double den = h2 + off * off;
        const double K = s / (FPA_4PI * den);
        // u = K (d x h)
        acc(dbar, K * cross(h, ub));
        fpa_v3 hb = K * cross(ub, d);
        const double Kb = dot(ub, cross(d, h));
        const double h2b = -Kb * K / den;
        acc(hb, (2.0 * h2b) * h);
        acc(Xbar, hb);
        acc(p0b, (-1.0) * hb);
    }
    // p0 = p + xm d, xm = (X - p) . d
    acc(pbar, p0b);
    const double xmb = dot(p0b, d);
    acc(dbar, xm * p0b);
    acc(Xbar, xmb * d);
    acc(pbar, (-xmb) * d);
    acc(dbar, xmb * (X - p));
}

// horseshoe (infinity -> pa along da with -s, pa -> pb with s, pb -> infinity along db with s)
FPA_HD fpa_v3 fpa_horse(fpa_v3 pa, fpa_v3 pb, fpa_v3 da, fpa_v3 db, double s, fpa_v3 X, double off, double cut)
{
    fpa_v3 u = fpa_semi(pa, da, -s, X, off, cut);
    acc(u, s * fpa_seg(pa, pb, X, off, cut));
    acc(u, fpa_semi(pb, db, s, X, off, cut));
    return u;
}

FPA_HD void fpa_horse_rev(fpa_v3 pa, fpa_v3 pb, fpa_v3 da, fpa_v3 db, double s, fpa_v3 X, double off, double cut,
                          fpa_v3 ub, fpa_v3 &pab, fpa_v3 &pbb, fpa_v3 &dab, fpa_v3 &dbb, fpa_v3 &Xbar)
{
    fpa_semi_rev(pa, da, -s, X, off, cut, ub, pab, dab, Xbar);
    fpa_seg_rev(pa, pb, X, off, cut, s * ub, pab, pbb, Xbar);
    fpa_semi_rev(pb, db, s, X, off, cut, ub, pbb, dbb, Xbar);
}

// Vortex ring of NVP vertices (edges k -> k+1 mod NVP), unit strength.
template <int NVP>
FPA_HD fpa_v3 fpa_ring(const double (&p)[NVP][3], fpa_v3 X, double off, double cut)
{
    fpa_v3 u = fpa_mk(0, 0, 0);
    for (int k = 0; k < NVP; ++k) {
        const int q = (k + 1) % NVP;
        acc(u, fpa_seg(fpa_mk(p[k][0], p[k][1], p[k][2]), fpa_mk(p[q][0], p[q][1], p[q][2]), X, off, cut));
    }
    return u;
}

// Reverse of the unit-strength vortex ring at one target X for K cotangents
// w[m]:  g[m] += u(X).w[m],  Xb[m] += (du/dX)^T w[m],  pb[m][k] += (du/dp_k)^T w[m].
// Node terms are shared by the two edges meeting at a node, and the
// normalisation r/(|r|+off) is reversed per edge as
//   T_k(v) = v/(|r_k|+off) - r_k (r_k.v) / (|r_k| (|r_k|+off)^2),
// so that no per-node adjoint has to be held between edges.
template <int NVP, int K>
FPA_HD void fpa_ring_rev(const double (&p)[NVP][3], double X, double Y, double Z, double off, double cut,
                         const double (&w)[K][3], double (&g)[K], double (&Xb)[K][3], double (&pb)[K][NVP][3])
{
    double r[NVP][3], iv[NVP], sk[NVP];
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        r[k][0] = X - p[k][0]; r[k][1] = Y - p[k][1]; r[k][2] = Z - p[k][2];
        const double l = sqrt(r[k][0] * r[k][0] + r[k][1] * r[k][1] + r[k][2] * r[k][2]);
        iv[k] = 1.0 / (l + off);
        sk[k] = l > 0 ? iv[k] * iv[k] / l : 0.0;
    }
    const double off2 = off * off, cut2 = cut * cut;
#pragma unroll
    for (int k = 0; k < NVP; ++k) {
        const int q = (k + 1) % NVP;
        const double cx = r[k][1] * r[q][2] - r[k][2] * r[q][1];
        const double cy = r[k][2] * r[q][0] - r[k][0] * r[q][2];
        const double cz = r[k][0] * r[q][1] - r[k][1] * r[q][0];
        const double c2 = cx * cx + cy * cy + cz * cz;
        if (!(c2 > cut2)) continue;
        const double ex = p[q][0] - p[k][0], ey = p[q][1] - p[k][1], ez = p[q][2] - p[k][2];
        const double dvx = r[k][0] * iv[k] - r[q][0] * iv[q];
        const double dvy = r[k][1] * iv[k] - r[q][1] * iv[q];
        const double dvz = r[k][2] * iv[k] - r[q][2] * iv[q];
        const double proj = ex * dvx + ey * dvy + ez * dvz;
        const double den = c2 + off2;
        const double Kf = 1.0 / (FPA_4PI * den);
        const double A = proj * Kf;
        const double rke = (r[k][0] * ex + r[k][1] * ey + r[k][2] * ez) * sk[k];
        const double rqe = (r[q][0] * ex + r[q][1] * ey + r[q][2] * ez) * sk[q];
        // T_k(e) - T_q(e), T_k(e), T_q(e)
        const double Ekx = iv[k] * ex - rke * r[k][0], Eky = iv[k] * ey - rke * r[k][1], Ekz = iv[k] * ez - rke * r[k][2];
        const double Eqx = iv[q] * ex - rqe * r[q][0], Eqy = iv[q] * ey - rqe * r[q][1], Eqz = iv[q] * ez - rqe * r[q][2];
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const double sb = w[m][0] * cx + w[m][1] * cy + w[m][2] * cz;
            g[m] += A * sb;
            const double projb = sb * Kf;
            const double c2b2 = -2.0 * sb * A / den;
            const double cbx = A * w[m][0] + c2b2 * cx;
            const double cby = A * w[m][1] + c2b2 * cy;
            const double cbz = A * w[m][2] + c2b2 * cz;
            // rb_k = r_q x cb + projb E_k,  rb_q = cb x r_k - projb E_q
            const double rkx = r[q][1] * cbz - r[q][2] * cby + projb * Ekx;
            const double rky = r[q][2] * cbx - r[q][0] * cbz + projb * Eky;
            const double rkz = r[q][0] * cby - r[q][1] * cbx + projb * Ekz;
            const double rqx = cby * r[k][2] - cbz * r[k][1] - projb * Eqx;
            const double rqy = cbz * r[k][0] - cbx * r[k][2] - projb * Eqy;
            const double rqz = cbx * r[k][1] - cby * r[k][0] - projb * Eqz;
            Xb[m][0] += rkx + rqx; Xb[m][1] += rky + rqy; Xb[m][2] += rkz + rqz;
            // e = p_q - p_k with adjoint projb (v_k - v_q)
            const double ebx = projb * dvx, eby = projb * dvy, ebz = projb * dvz;
            pb[m][k][0] -= rkx + ebx; pb[m][k][1] -= rky + eby; pb[m][k][2] -= rkz + ebz;
            pb[m][q][0] += ebx - rqx; pb[m][q][1] += eby - rqy; pb[m][q][2] += ebz - rqz;
        }
    }
}
