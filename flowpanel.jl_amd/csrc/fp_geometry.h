// fp_geometry.h -- panel geometry as one function of the vertices.
//
// panel_geom() restates src/FLOWPanel_abstractbody.jl:635-711 (control point:
// vertex mean + in-plane centroid correction), :738-776 (normal from the first
// three vertices), :905-927 (area, fan triangulation from the first vertex) and
// the characteristic lengths of :573-633.  It is templated on the scalar so the
// same code gives the geometry (S = double) and its directional derivatives
// (S = fpg_dual, value + one tangent): the vertex adjoint of the geometry is the
// transpose of the 3*nv directional derivatives, evaluated exactly.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FPG_HD __host__ __device__ __forceinline__
#else
#define FPG_HD inline
#endif
#include <math.h>

// characteristic-length kinds (same values as FP_CHARLEN_* of the C ABI)
#define FPG_CHARLEN_UNITARY 0
#define FPG_CHARLEN_BBOX 1
#define FPG_CHARLEN_MAXDIST 2
#define FPG_CHARLEN_SQRTAREA 3

// forward-mode dual number: value and one directional derivative
struct fpg_dual {
    double v, d;
};
FPG_HD fpg_dual fpg_mk(double v, double d) { fpg_dual r = {v, d}; return r; }
FPG_HD fpg_dual operator+(fpg_dual a, fpg_dual b) { return fpg_mk(a.v + b.v, a.d + b.d); }
FPG_HD fpg_dual operator-(fpg_dual a, fpg_dual b) { return fpg_mk(a.v - b.v, a.d - b.d); }
FPG_HD fpg_dual operator-(fpg_dual a) { return fpg_mk(-a.v, -a.d); }
FPG_HD fpg_dual operator*(fpg_dual a, fpg_dual b) { return fpg_mk(a.v * b.v, a.d * b.v + a.v * b.d); }
FPG_HD fpg_dual operator*(double s, fpg_dual a) { return fpg_mk(s * a.v, s * a.d); }
FPG_HD fpg_dual operator/(fpg_dual a, fpg_dual b)
{
    const double q = a.v / b.v;
    return fpg_mk(q, (a.d - q * b.d) / b.v);
}
FPG_HD fpg_dual operator/(fpg_dual a, double s) { return fpg_mk(a.v / s, a.d / s); }
FPG_HD fpg_dual &operator+=(fpg_dual &a, fpg_dual b) { a.v += b.v; a.d += b.d; return a; }
FPG_HD fpg_dual sqrt(fpg_dual a)
{
    const double s = sqrt(a.v);
    return fpg_mk(s, s > 0 ? a.d / (2 * s) : 0.0);
}
FPG_HD fpg_dual fmin(fpg_dual a, fpg_dual b) { return b.v < a.v ? b : a; }
FPG_HD fpg_dual fmax(fpg_dual a, fpg_dual b) { return b.v > a.v ? b : a; }
FPG_HD double fpg_val(double a) { return a; }
FPG_HD double fpg_val(fpg_dual a) { return a.v; }

template <typename S>
FPG_HD S fpg_const(double c) { return S(c); }
template <>
FPG_HD fpg_dual fpg_const<fpg_dual>(double c) { return fpg_mk(c, 0); }
template <typename S>
FPG_HD S fpg_zero() { return fpg_const<S>(0.0); }

// characteristic length of kind `charlen` (src/FLOWPanel_abstractbody.jl:573-633)
template <typename S, int NVP>
FPG_HD S panel_len(const S (&p)[NVP][3], int nv, int charlen, S area)
{
    if (charlen == FPG_CHARLEN_BBOX) {
        S lo[3] = {p[0][0], p[0][1], p[0][2]}, hi[3] = {p[0][0], p[0][1], p[0][2]};
        for (int k = 0; k < NVP; ++k)
            if (k < nv)
                for (int d = 0; d < 3; ++d) { lo[d] = fmin(lo[d], p[k][d]); hi[d] = fmax(hi[d], p[k][d]); }
        return sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                    (hi[2] - lo[2]) * (hi[2] - lo[2]));
    }
    if (charlen == FPG_CHARLEN_MAXDIST) {
        S len = fpg_zero<S>();
        for (int k = 0; k < NVP; ++k)
            if (k < nv) {
                const S dx = p[0][0] - p[k][0], dy = p[0][1] - p[k][1], dz = p[0][2] - p[k][2];
                len = fmax(len, sqrt(dx * dx + dy * dy + dz * dz));
            }
        return len;
    }
    if (charlen == FPG_CHARLEN_SQRTAREA) return sqrt(area);
    return fpg_const<S>(1.0);
}

// Geometry of one panel with nv <= NVP vertices p[0..nv-1]:
//   t, n, o  panel frame (t along edge 0->1, n = (p1-p0)x(p2-p0)/|.|, o = n x t)
//   m        control point before the normal offset (centroid)
//   area     polygon area (fan from vertex 0)
//   len      characteristic length of kind `charlen`
template <typename S, int NVP>
FPG_HD void panel_geom(const S (&p)[NVP][3], int nv, int charlen, S (&t)[3], S (&n)[3], S (&o)[3], S (&m)[3],
                       S &area, S &len)
{
    S a[3], b[3];
    for (int d = 0; d < 3; ++d) { a[d] = p[1][d] - p[0][d]; b[d] = p[2][d] - p[0][d]; }
    const S na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int d = 0; d < 3; ++d) t[d] = a[d] / na;
    S c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const S nc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int d = 0; d < 3; ++d) n[d] = c[d] / nc;
    o[0] = n[1] * t[2] - n[2] * t[1];
    o[1] = n[2] * t[0] - n[0] * t[2];
    o[2] = n[0] * t[1] - n[1] * t[0];
    for (int d = 0; d < 3; ++d) m[d] = fpg_zero<S>();
    for (int k = 0; k < NVP; ++k)
        if (k < nv)
            for (int d = 0; d < 3; ++d) m[d] += p[k][d];
    const double dn = (double)nv;
    for (int d = 0; d < 3; ++d) m[d] = m[d] / dn;
    S xs = fpg_zero<S>(), ys = fpg_zero<S>();
    for (int k = 0; k < NVP; ++k)
        if (k < nv) {
            const S dx = p[k][0] - m[0], dy = p[k][1] - m[1], dz = p[k][2] - m[2];
            xs += dx * t[0] + dy * t[1] + dz * t[2];
            ys += dx * o[0] + dy * o[1] + dz * o[2];
        }
    xs = xs / dn;
    ys = ys / dn;
    for (int d = 0; d < 3; ++d) m[d] += xs * t[d] + ys * o[d];
    S ac[3] = {fpg_zero<S>(), fpg_zero<S>(), fpg_zero<S>()};
    for (int k = 1; k < NVP - 1; ++k)
        if (k + 1 < nv) {
            const S u0 = p[k][0] - p[0][0], u1 = p[k][1] - p[0][1], u2 = p[k][2] - p[0][2];
            const S w0 = p[k + 1][0] - p[0][0], w1 = p[k + 1][1] - p[0][1], w2 = p[k + 1][2] - p[0][2];
            ac[0] += u1 * w2 - u2 * w1;
            ac[1] += u2 * w0 - u0 * w2;
            ac[2] += u0 * w1 - u1 * w0;
        }
    area = 0.5 * sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
    len = panel_len<S, NVP>(p, nv, charlen, area);
}
