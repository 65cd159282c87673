/*
 * fp_oracle.c -- CPU FP64 restatement of the FLOWPanel.jl panel-method hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle that the HIP
 * library is checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Every routine restates the reference's published algorithm (Julia source,
 * cited file:line) in plain C with FP64 arithmetic.  Indices are 0-based here;
 * the Python driver (oracle/oracle.py) converts from the 1-based convention of
 * the reference.  Arrays of 3-vectors are column-major 3 x n (x,y,z contiguous
 * per point), exactly the reference's `nodes`/`targets`/`normals` layout.
 *
 * Pins (what checks each routine; see tests/test_oracle_*.py):
 *   panel frame          orthonormal, right-handed (test/runtests_grid.jl:122-132)
 *   U_constant_source    U = grad(phi_constant_source) (test/runtests_elements.jl:38-92),
 *                        div U = 0 (:156-215), dphi/dz(+-0) = +-sigma/2 (:254-282),
 *                        far field -> point source, brute-force quadrature
 *   phi_constant_doublet phi(+-0) = -+mu/2, far field -> point doublet (:284-312)
 *   U_vortexring         U = grad(phi_constant_doublet) (:23-26), div U = 0
 *   semi-infinite parts  U = grad(phi_semiinfinite_doublet), div U = 0,
 *                        phi(+-0) = +-mu/2 (test/runtests_semiinfiniteelements.jl),
 *                        long-finite-ring limit
 *   assembly / Uind      column j == U_vortexring of panel j dotted with normals;
 *                        sphere closed form Cp = 1 - 9/4 sin^2(theta)
 *   Cp / Kutta / F / LDS hand-computed small cases
 *   areas                exact for planar polygons; non-planar polygons use a
 *                        fan triangulation from vertex 1 (the un-vendored
 *                        GeometricTools._get_area is not available):
 *                        parity unpinned for non-planar polygons.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define ORC_PI 3.14159265358979323846

/* status codes shared with the C-ABI (include/flowpanel_hip.h) */
#define ORC_OK 0
#define ORC_ERR_NONUNIT_DIRECTION 3
#define ORC_ERR_BAD_DIRECTIONS 4

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double nrm(v3 a) { return sqrt(dot(a, a)); }
static inline v3 at(const double *p, int64_t k) { return mk(p[3 * k], p[3 * k + 1], p[3 * k + 2]); }
static inline double sgn(double z) { return z > 0 ? 1.0 : (z < 0 ? -1.0 : z); }

/* ------------------------------------------------------------------------- */
/* Panel frame (docs/src/elements/paneldefinition.md:14-39; GeometricTools
 * _calc_t/_calc_n/_calc_o): t = (p2-p1)/|p2-p1|,
 * n = (p2-p1)x(p3-p1)/|.|, o = n x t.                                        */
void orc_panel_frame(const double *nodes, const int64_t *panel,
                     double *t, double *o, double *n)
{
    v3 p1 = at(nodes, panel[0]), p2 = at(nodes, panel[1]), p3 = at(nodes, panel[2]);
    v3 a = sub(p2, p1), b = sub(p3, p1);
    v3 tt = scl(a, 1.0 / nrm(a));
    v3 c = cross(a, b);
    v3 nn = scl(c, 1.0 / nrm(c));
    v3 oo = cross(nn, tt);
    t[0] = tt.x; t[1] = tt.y; t[2] = tt.z;
    o[0] = oo.x; o[1] = oo.y; o[2] = oo.z;
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
}

/* Characteristic lengths (src/FLOWPanel_abstractbody.jl:573-633).
 * kind 0 = unitary, 1 = bbox diagonal, 2 = max distance from node 1,
 * 3 = sqrt(area).                                                            */
static double panel_area(const double *nodes, const int64_t *panel, int nv)
{
    /* fan triangulation from the first vertex */
    v3 p1 = at(nodes, panel[0]);
    v3 acc = mk(0, 0, 0);
    for (int k = 1; k + 1 < nv; ++k) {
        v3 a = sub(at(nodes, panel[k]), p1), b = sub(at(nodes, panel[k + 1]), p1);
        acc = add(acc, cross(a, b));
    }
    return 0.5 * nrm(acc);
}

static double charlen(const double *nodes, const int64_t *panel, int nv, int kind)
{
    if (kind == 0) return 1.0;
    if (kind == 1) {
        v3 lo = at(nodes, panel[0]), hi = lo;
        for (int k = 0; k < nv; ++k) {
            v3 p = at(nodes, panel[k]);
            lo.x = fmin(lo.x, p.x); lo.y = fmin(lo.y, p.y); lo.z = fmin(lo.z, p.z);
            hi.x = fmax(hi.x, p.x); hi.y = fmax(hi.y, p.y); hi.z = fmax(hi.z, p.z);
        }
        return nrm(sub(hi, lo));
    }
    if (kind == 2) {
        double l = 0;
        v3 p0 = at(nodes, panel[0]);
        for (int k = 0; k < nv; ++k) {
            double d = nrm(sub(p0, at(nodes, panel[k])));
            if (d > l) l = d;
        }
        return l;
    }
    return sqrt(panel_area(nodes, panel, nv));
}

/* Normals, control points, areas
 * (src/FLOWPanel_abstractbody.jl:635-694 control points, :738-755 normals,
 *  :905-927 areas).  The control point is the vertex mean, plus the in-plane
 *  "centroid" correction of :664-684 (which sums to zero), plus
 *  cpoffset * charlen * n.                                                   */
void orc_geometry(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                  int64_t ncells, double cpoffset, int charlen_kind,
                  double *normals, double *cps, double *areas)
{
    for (int64_t c = 0; c < ncells; ++c) {
        const int64_t *panel = cell_idx + cell_ptr[c];
        int nv = (int)(cell_ptr[c + 1] - cell_ptr[c]);
        double t[3], o[3], n[3];
        orc_panel_frame(nodes, panel, t, o, n);
        v3 tt = mk(t[0], t[1], t[2]), oo = mk(o[0], o[1], o[2]), nn = mk(n[0], n[1], n[2]);
        v3 m = mk(0, 0, 0);
        for (int k = 0; k < nv; ++k) m = add(m, at(nodes, panel[k]));
        m = scl(m, 1.0 / nv);
        double x = 0, y = 0;
        for (int k = 0; k < nv; ++k) {
            v3 d = sub(at(nodes, panel[k]), m);
            x += dot(d, tt);
            y += dot(d, oo);
        }
        x /= nv; y /= nv;
        m = add(m, add(scl(tt, x), scl(oo, y)));
        double l = charlen(nodes, panel, nv, charlen_kind);
        m = add(m, scl(nn, cpoffset * l));
        if (normals) { normals[3 * c] = n[0]; normals[3 * c + 1] = n[1]; normals[3 * c + 2] = n[2]; }
        if (cps) { cps[3 * c] = m.x; cps[3 * c + 1] = m.y; cps[3 * c + 2] = m.z; }
        if (areas) areas[c] = panel_area(nodes, panel, nv);
    }
}

/* ------------------------------------------------------------------------- */
/* Accumulate a velocity into out[ti] (dot with normal) or out[:,ti].         */
static inline void put(double *out, const double *dot_with, int64_t ti, v3 u)
{
    if (dot_with) out[ti] += dot(u, at(dot_with, ti));
    else { out[3 * ti] += u.x; out[3 * ti + 1] += u.y; out[3 * ti + 2] += u.z; }
}

/* Regularised Biot-Savart segment pa->pb (src/FLOWPanel_elements.jl:345-428):
 * u = G rij.(ri/(|ri|+off) - rj/(|rj|+off)) (ri x rj) / (4 pi (|ri x rj|^2 + off^2)),
 * zero when |ri x rj|^2 <= cutoff^2.                                         */
static inline v3 segment(v3 pa, v3 pb, double strength, v3 X, double cutoff, double offset, int *hit)
{
    v3 rij = sub(pb, pa);
    v3 ri = sub(X, pa), rj = sub(X, pb);
    v3 c = cross(ri, rj);
    double c2 = dot(c, c);
    double nri = nrm(ri) + offset, nrj = nrm(rj) + offset;
    double proj = dot(rij, sub(scl(ri, 1.0 / nri), scl(rj, 1.0 / nrj)));
    *hit = c2 > cutoff * cutoff;
    if (!*hit) return mk(0, 0, 0);
    double aux = strength * proj / (4 * ORC_PI * (c2 + offset * offset));
    return scl(c, aux);
}

void orc_U_boundvortex(const double *pa, const double *pb, double strength,
                       const double *targets, int64_t nt, double *out, const double *dot_with,
                       double cutoff, double offset)
{
    v3 a = mk(pa[0], pa[1], pa[2]), b = mk(pb[0], pb[1], pb[2]);
    for (int64_t ti = 0; ti < nt; ++ti) {
        int hit;
        v3 u = segment(a, b, strength, at(targets, ti), cutoff, offset, &hit);
        if (hit) put(out, dot_with, ti, u);
    }
}

/* Vortex ring = constant doublet (src/FLOWPanel_elements.jl:442-467, :740):
 * sum of segments (k, k+1 mod nv).                                           */
void orc_U_vortexring(const double *nodes, const int64_t *panel, int nv, double strength,
                      const double *targets, int64_t nt, double *out, const double *dot_with,
                      double cutoff, double offset)
{
    for (int k = 0; k < nv; ++k) {
        double pa[3], pb[3];
        v3 a = at(nodes, panel[k]), b = at(nodes, panel[(k + 1) % nv]);
        pa[0] = a.x; pa[1] = a.y; pa[2] = a.z; pb[0] = b.x; pb[1] = b.y; pb[2] = b.z;
        orc_U_boundvortex(pa, pb, strength, targets, nt, out, dot_with, cutoff, offset);
    }
}

static int nonunit(v3 d) { return fabs(dot(d, d) - 1.0) > 2 * DBL_EPSILON; }

/* Semi-infinite vortex from p along unit d (src/FLOWPanel_elements.jl:480-628):
 * bound part p -> p0 (foot point of the target on the line) when
 * |p0-p|^2 > off^2, plus G (d x h)/(4 pi (|h|^2 + off^2)) with h = X - p0 when
 * |h|^2 > cutoff^2.                                                          */
int orc_U_semiinfinite_vortex(const double *p, const double *d, double strength,
                              const double *targets, int64_t nt, double *out, const double *dot_with,
                              double cutoff, double offset)
{
    v3 P = mk(p[0], p[1], p[2]), D = mk(d[0], d[1], d[2]);
    if (nonunit(D)) return ORC_ERR_NONUNIT_DIRECTION;
    for (int64_t ti = 0; ti < nt; ++ti) {
        v3 X = at(targets, ti);
        double s = dot(sub(X, P), D);
        v3 P0 = add(P, scl(D, s));
        v3 dp = sub(P0, P);
        if (dot(dp, dp) > offset * offset) {
            int hit;
            v3 u = segment(P, P0, strength, X, cutoff, offset, &hit);
            if (hit) put(out, dot_with, ti, u);
        }
        v3 h = sub(X, P0);
        double h2 = dot(h, h);
        if (h2 > cutoff * cutoff) {
            double aux = strength / (4 * ORC_PI * (h2 + offset * offset));
            put(out, dot_with, ti, scl(cross(D, h), aux));
        }
    }
    return ORC_OK;
}

/* Horseshoe (src/FLOWPanel_elements.jl:692-730): from infinity to pa along da
 * (strength -G), bound pa->pb (G), pb to infinity along db (G).              */
int orc_U_semiinfinite_horseshoe(const double *pa, const double *pb, const double *da, const double *db,
                                 double strength, const double *targets, int64_t nt, double *out,
                                 const double *dot_with, double cutoff, double offset)
{
    int st = orc_U_semiinfinite_vortex(pa, da, -strength, targets, nt, out, dot_with, cutoff, offset);
    if (st) return st;
    orc_U_boundvortex(pa, pb, strength, targets, nt, out, dot_with, cutoff, offset);
    return orc_U_semiinfinite_vortex(pb, db, strength, targets, nt, out, dot_with, cutoff, offset);
}

/* ------------------------------------------------------------------------- */
/* Hess & Smith planar polygon.  Local frame (src/FLOWPanel_elements.jl:54-56):
 * x-axis = o, y-axis = t, z-axis = -n, origin = first vertex.                */
typedef struct { v3 ex, ey, ez, O; } hsframe;

static hsframe hs_frame(const double *nodes, const int64_t *panel)
{
    double t[3], o[3], n[3];
    orc_panel_frame(nodes, panel, t, o, n);
    hsframe f;
    f.ex = mk(o[0], o[1], o[2]);
    f.ey = mk(t[0], t[1], t[2]);
    f.ez = mk(-n[0], -n[1], -n[2]);
    f.O = at(nodes, panel[0]);
    return f;
}
static inline v3 to_local(const hsframe *f, v3 X)
{
    v3 d = sub(X, f->O);
    return mk(dot(f->ex, d), dot(f->ey, d), dot(f->ez, d));
}

/* Per-target Hess-Smith sums.  Returns V (local components, before 1/(4 pi))
 * and the potential sum for the source; doublet solid angle in V.z.
 * src/FLOWPanel_elements.jl:79-121 (velocity), :199-236 (potential).         */
typedef struct { double V1, V2, V3, phis; } hsout;

static hsout hs_sums(const double *nodes, const int64_t *panel, int nv, const hsframe *f,
                     v3 x, double offset, int want_q)
{
    hsout r = {0, 0, 0, 0};
    double dtheta = 2 * ORC_PI;
    int nR0 = 0;
    double phisum = 0;
    for (int k = 0; k < nv; ++k) {
        v3 pi = to_local(f, at(nodes, panel[k]));
        v3 pj = to_local(f, at(nodes, panel[(k + 1) % nv]));
        double dij = nrm(sub(pj, pi));
        double ri = nrm(sub(x, pi)), rj = nrm(sub(x, pj));
        double S = (pj.y - pi.y) / dij, C = (pj.x - pi.x) / dij;
        double si = (pi.x - x.x) * C + (pi.y - x.y) * S;
        double sj = (pj.x - x.x) * C + (pj.y - x.y) * S;
        double R = (x.x - pi.x) * S - (x.y - pi.y) * C;
        double az = fabs(x.z);
        double J = atan2(R * az * (ri * sj - rj * si), ri * rj * R * R + x.z * x.z * sj * si);
        if (want_q) {
            double Q = log((ri + rj + dij) / (ri + rj - dij + offset));
            r.V1 -= S * Q;
            r.V2 += C * Q;
            phisum -= (R * Q + az * J);
        }
        r.V3 -= J;
        dtheta *= (R >= 0) ? 1.0 : 0.0;
        nR0 += (R == 0);
    }
    r.phis = phisum + fabs(x.z) * dtheta;
    r.V3 += dtheta;
    r.V3 *= sgn(x.z);
    r.V3 *= (nR0 > 1) ? 0.0 : 1.0;
    return r;
}

/* U_constant_source (src/FLOWPanel_elements.jl:29-141) */
void orc_U_constant_source(const double *nodes, const int64_t *panel, int nv, double strength,
                           const double *targets, int64_t nt, double *out, const double *dot_with,
                           double offset)
{
    hsframe f = hs_frame(nodes, panel);
    const double k4 = 1.0 / (4 * ORC_PI);
    for (int64_t ti = 0; ti < nt; ++ti) {
        v3 x = to_local(&f, at(targets, ti));
        hsout s = hs_sums(nodes, panel, nv, &f, x, offset, 1);
        double V1 = s.V1 * k4, V2 = s.V2 * k4, V3 = s.V3 * k4;
        v3 u = add(add(scl(f.ex, V1), scl(f.ey, V2)), scl(f.ez, V3));
        put(out, dot_with, ti, scl(u, strength));
    }
}

/* phi_constant_source (src/FLOWPanel_elements.jl:152-241) */
void orc_phi_constant_source(const double *nodes, const int64_t *panel, int nv, double strength,
                             const double *targets, int64_t nt, double *out, double offset)
{
    hsframe f = hs_frame(nodes, panel);
    for (int64_t ti = 0; ti < nt; ++ti) {
        v3 x = to_local(&f, at(targets, ti));
        hsout s = hs_sums(nodes, panel, nv, &f, x, offset, 1);
        out[ti] += strength / (4 * ORC_PI) * s.phis;
    }
}

/* phi_constant_doublet (src/FLOWPanel_elements.jl:751-847) */
void orc_phi_constant_doublet(const double *nodes, const int64_t *panel, int nv, double strength,
                              const double *targets, int64_t nt, double *out)
{
    hsframe f = hs_frame(nodes, panel);
    for (int64_t ti = 0; ti < nt; ++ti) {
        v3 x = to_local(&f, at(targets, ti));
        hsout s = hs_sums(nodes, panel, nv, &f, x, 0.0, 0);
        out[ti] += strength * s.V3 / (4 * ORC_PI);
    }
}

/* Solid angle of a 3-vertex polygon given explicitly (used by the
 * semi-infinite doublet, src/FLOWPanel_elements.jl:1016-1068).               */
static double tri_solid(const v3 *P, v3 ex, v3 ey, v3 ez, v3 O, v3 X)
{
    int64_t idx[3] = {0, 1, 2};
    double pts[9];
    for (int k = 0; k < 3; ++k) { pts[3 * k] = P[k].x; pts[3 * k + 1] = P[k].y; pts[3 * k + 2] = P[k].z; }
    hsframe f; f.ex = ex; f.ey = ey; f.ez = ez; f.O = O;
    v3 x = to_local(&f, X);
    hsout s = hs_sums(pts, idx, 3, &f, x, 0.0, 0);
    return s.V3 / (4 * ORC_PI);
}

/* phi_semiinfinite_doublet, planar (src/FLOWPanel_elements.jl:959-1092) */
int orc_phi_semiinfinite_doublet(const double *pi_, const double *pj_, const double *d, double strength,
                                 const double *targets, int64_t nt, double *out)
{
    v3 D = mk(d[0], d[1], d[2]);
    if (nonunit(D)) return ORC_ERR_NONUNIT_DIRECTION;
    v3 Pi = mk(pi_[0], pi_[1], pi_[2]), Pj = mk(pj_[0], pj_[1], pj_[2]);
    double pd = dot(sub(Pj, Pi), D);
    v3 Pa = add(Pi, scl(D, pd));
    v3 ex = D;
    v3 yv = sub(Pj, Pa);
    v3 ey = scl(yv, 1.0 / nrm(yv));
    v3 ez = cross(ex, ey);
    v3 O = Pi;
    hsframe f; f.ex = ex; f.ey = ey; f.ez = ez; f.O = O;
    v3 a = to_local(&f, Pa), b = to_local(&f, Pj);
    v3 tri[3] = {Pa, Pi, Pj};
    for (int64_t ti = 0; ti < nt; ++ti) {
        v3 X = at(targets, ti);
        if (fabs(pd) > 2 * DBL_EPSILON)
            out[ti] += strength * tri_solid(tri, ex, ey, ez, O, X);
        v3 x = to_local(&f, X);
        double dx = x.x - a.x;
        double val = atan((b.y - x.y) / x.z)
                   + atan((b.y - x.y) * dx / (x.z * sqrt(dx * dx + (b.y - x.y) * (b.y - x.y) + x.z * x.z)));
        val -= atan((a.y - x.y) / x.z)
             + atan((a.y - x.y) * dx / (x.z * sqrt(dx * dx + (a.y - x.y) * (a.y - x.y) + x.z * x.z)));
        out[ti] += strength / (4 * ORC_PI) * val;
    }
    return ORC_OK;
}

/* phi_semiinfinite_doublet, non-planar (src/FLOWPanel_elements.jl:1110-1244):
 * planar part along da plus a large triangle (pb, pb+L db, pb+L da).         */
int orc_phi_semiinfinite_doublet2(const double *pa_, const double *pb_, const double *da, const double *db,
                                  double strength, const double *targets, int64_t nt, double *out,
                                  double infinitelength)
{
    v3 Da = mk(da[0], da[1], da[2]), Db = mk(db[0], db[1], db[2]);
    if (nonunit(Da) || nonunit(Db)) return ORC_ERR_NONUNIT_DIRECTION;
    if (dot(Da, Db) <= 0) return ORC_ERR_BAD_DIRECTIONS;
    int st = orc_phi_semiinfinite_doublet(pa_, pb_, da, strength, targets, nt, out);
    if (st) return st;
    if (fabs(da[0] - db[0]) > 2 * DBL_EPSILON || fabs(da[1] - db[1]) > 2 * DBL_EPSILON ||
        fabs(da[2] - db[2]) > 2 * DBL_EPSILON) {
        v3 Pb = mk(pb_[0], pb_[1], pb_[2]);
        v3 Pda = add(Pb, scl(Da, infinitelength)), Pdb = add(Pb, scl(Db, infinitelength));
        v3 ex = Da;
        v3 zv = cross(Da, Db);
        v3 ez = scl(zv, 1.0 / nrm(zv));
        v3 ey = cross(ez, ex);
        v3 tri[3] = {Pda, Pb, Pdb};
        for (int64_t ti = 0; ti < nt; ++ti)
            out[ti] += strength * tri_solid(tri, ex, ey, ez, Pb, at(targets, ti));
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Shedding: 6 x ns, 0-based (pi, nia, nib, pj, nja, njb), pj = -1 for none.
 * Trailing-edge node pair of one side (src/FLOWPanel_liftingbody.jl:519-566). */
static void te_nodes(const int64_t *cell_ptr, const int64_t *cell_idx, int64_t p, int64_t na, int64_t nb,
                     int64_t *te)
{
    te[0] = cell_idx[cell_ptr[p] + na];
    te[1] = cell_idx[cell_ptr[p] + nb];
}

/* AIC of a vortex-ring body with rigid wake (src/FLOWPanel_liftingbody.jl:455-572):
 * G[:, col0 + j] = n . U_vortexring(panel j, 1) at every CP, then each
 * shedding edge adds a unit horseshoe to column pi (da, db) and to column pj
 * (db, da).  G is column-major with leading dimension ldg; M targets.        */
int orc_assemble_vortexring(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                            int64_t ncells, const int64_t *shedding, int64_t ns,
                            const double *Das, const double *Dbs,
                            const double *cps, const double *normals, int64_t M,
                            double offset, double cutoff, double *G, int64_t ldg)
{
    for (int64_t e = 0; e < ns; ++e) {
        if (nonunit(at(Das, e)) || nonunit(at(Dbs, e))) return ORC_ERR_NONUNIT_DIRECTION;
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t j = 0; j < ncells; ++j) {
        double *col = G + j * ldg;
        for (int64_t i = 0; i < M; ++i) col[i] = 0.0;
        orc_U_vortexring(nodes, cell_idx + cell_ptr[j], (int)(cell_ptr[j + 1] - cell_ptr[j]), 1.0,
                         cps, M, col, normals, cutoff, offset);
    }
    for (int64_t e = 0; e < ns; ++e) {
        const int64_t *s = shedding + 6 * e;
        int64_t te[2];
        double pa[3], pb[3];
        te_nodes(cell_ptr, cell_idx, s[0], s[1], s[2], te);
        memcpy(pa, nodes + 3 * te[0], sizeof pa); memcpy(pb, nodes + 3 * te[1], sizeof pb);
        orc_U_semiinfinite_horseshoe(pa, pb, Das + 3 * e, Dbs + 3 * e, 1.0, cps, M, G + s[0] * ldg,
                                     normals, cutoff, offset);
        if (s[3] != -1) {
            te_nodes(cell_ptr, cell_idx, s[3], s[4], s[5], te);
            memcpy(pa, nodes + 3 * te[0], sizeof pa); memcpy(pb, nodes + 3 * te[1], sizeof pb);
            orc_U_semiinfinite_horseshoe(pa, pb, Dbs + 3 * e, Das + 3 * e, 1.0, cps, M, G + s[3] * ldg,
                                         normals, cutoff, offset);
        }
    }
    return ORC_OK;
}

/* AIC of a constant-source body (src/FLOWPanel_nonliftingbody.jl:141-186). */
void orc_assemble_source(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                         int64_t ncells, const double *cps, const double *normals, int64_t M,
                         double offset, double *G, int64_t ldg)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t j = 0; j < ncells; ++j) {
        double *col = G + j * ldg;
        for (int64_t i = 0; i < M; ++i) col[i] = 0.0;
        orc_U_constant_source(nodes, cell_idx + cell_ptr[j], (int)(cell_ptr[j + 1] - cell_ptr[j]), 1.0,
                              cps, M, col, normals, offset);
    }
}

/* Induced velocity of a solved vortex-ring body with wake
 * (src/FLOWPanel_liftingbody.jl:576-661; wake strengths (G_pi, G_pj) per
 * _get_wakestrength_mu :1078-1082).  out (3 x nt) is accumulated.
 * Targets are split over threads; each thread sums every panel in order.     */
int orc_uind_vortexring(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                        int64_t ncells, const int64_t *shedding, int64_t ns,
                        const double *Das, const double *Dbs, const double *strength,
                        const double *targets, int64_t nt, double *out, double offset, double cutoff)
{
    for (int64_t e = 0; e < ns; ++e) {
        if (nonunit(at(Das, e)) || nonunit(at(Dbs, e))) return ORC_ERR_NONUNIT_DIRECTION;
    }
#pragma omp parallel for schedule(static)
    for (int64_t ti = 0; ti < nt; ++ti) {
        const double *X = targets + 3 * ti;
        double *o = out + 3 * ti;
        for (int64_t j = 0; j < ncells; ++j)
            orc_U_vortexring(nodes, cell_idx + cell_ptr[j], (int)(cell_ptr[j + 1] - cell_ptr[j]),
                             strength[j], X, 1, o, NULL, cutoff, offset);
        for (int64_t e = 0; e < ns; ++e) {
            const int64_t *s = shedding + 6 * e;
            int64_t te[2];
            double pa[3], pb[3];
            te_nodes(cell_ptr, cell_idx, s[0], s[1], s[2], te);
            memcpy(pa, nodes + 3 * te[0], sizeof pa); memcpy(pb, nodes + 3 * te[1], sizeof pb);
            orc_U_semiinfinite_horseshoe(pa, pb, Das + 3 * e, Dbs + 3 * e, strength[s[0]], X, 1, o,
                                         NULL, cutoff, offset);
            if (s[3] != -1) {
                te_nodes(cell_ptr, cell_idx, s[3], s[4], s[5], te);
                memcpy(pa, nodes + 3 * te[0], sizeof pa); memcpy(pb, nodes + 3 * te[1], sizeof pb);
                orc_U_semiinfinite_horseshoe(pa, pb, Dbs + 3 * e, Das + 3 * e, strength[s[3]], X, 1, o,
                                             NULL, cutoff, offset);
            }
        }
    }
    return ORC_OK;
}

/* Induced velocity of a solved constant-source body
 * (src/FLOWPanel_nonliftingbody.jl:188-218).                                 */
void orc_uind_source(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx, int64_t ncells,
                     const double *strength, const double *targets, int64_t nt, double *out, double offset)
{
#pragma omp parallel for schedule(static)
    for (int64_t ti = 0; ti < nt; ++ti)
        for (int64_t j = 0; j < ncells; ++j)
            orc_U_constant_source(nodes, cell_idx + cell_ptr[j], (int)(cell_ptr[j + 1] - cell_ptr[j]),
                                  strength[j], targets + 3 * ti, 1, out + 3 * ti, NULL, offset);
}

/* Potential of a solved vortex-ring body (src/FLOWPanel_liftingbody.jl:664-742). */
int orc_phi_vortexring(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                       int64_t ncells, const int64_t *shedding, int64_t ns,
                       const double *Das, const double *Dbs, const double *strength,
                       const double *targets, int64_t nt, double *out)
{
    int status = ORC_OK;
#pragma omp parallel for schedule(static)
    for (int64_t ti = 0; ti < nt; ++ti) {
        const double *X = targets + 3 * ti;
        for (int64_t j = 0; j < ncells; ++j)
            orc_phi_constant_doublet(nodes, cell_idx + cell_ptr[j], (int)(cell_ptr[j + 1] - cell_ptr[j]),
                                     strength[j], X, 1, out + ti);
        for (int64_t e = 0; e < ns; ++e) {
            const int64_t *s = shedding + 6 * e;
            int64_t te[2];
            te_nodes(cell_ptr, cell_idx, s[0], s[1], s[2], te);
            int st = orc_phi_semiinfinite_doublet2(nodes + 3 * te[0], nodes + 3 * te[1], Das + 3 * e,
                                                   Dbs + 3 * e, strength[s[0]], X, 1, out + ti, 1e12);
            if (st) status = st;
            if (s[3] != -1) {
                te_nodes(cell_ptr, cell_idx, s[3], s[4], s[5], te);
                st = orc_phi_semiinfinite_doublet2(nodes + 3 * te[0], nodes + 3 * te[1], Dbs + 3 * e,
                                                   Das + 3 * e, strength[s[3]], X, 1, out + ti, 1e12);
                if (st) status = st;
            }
        }
    }
    return status;
}

/* Potential of a solved constant-source body (src/FLOWPanel_nonliftingbody.jl:221-249). */
void orc_phi_source(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx, int64_t ncells,
                    const double *strength, const double *targets, int64_t nt, double *out, double offset)
{
#pragma omp parallel for schedule(static)
    for (int64_t ti = 0; ti < nt; ++ti)
        for (int64_t j = 0; j < ncells; ++j)
            orc_phi_constant_source(nodes, cell_idx + cell_ptr[j], (int)(cell_ptr[j + 1] - cell_ptr[j]),
                                    strength[j], targets + 3 * ti, 1, out + ti, offset);
}

/* ------------------------------------------------------------------------- */
/* Pressure coefficient with Kutta averaging (src/FLOWPanel_postprocess.jl:1059-1112).
 * out is accumulated; shedding may be NULL for non-lifting bodies.           */
void orc_calcfield_cp(const double *Us, int64_t n, double uref, const int64_t *shedding, int64_t ns,
                      int kutta, double *out)
{
    for (int64_t i = 0; i < n; ++i) {
        double u = nrm(at(Us, i)) / uref;
        out[i] += 1 - u * u;
    }
    if (!kutta || !shedding) return;
    for (int64_t e = 0; e < ns; ++e) {
        const int64_t *s = shedding + 6 * e;
        int64_t pi = s[0], pj = s[3];
        if (pj != -1) {
            double ave = (out[pi] + out[pi + 1] + out[pj] + out[pj - 1]) / 4;
            out[pi] = ave; out[pi + 1] = ave; out[pj] = ave; out[pj - 1] = ave;
        } else {
            double ave = (out[pi] + out[pi + 1]) / 2;
            out[pi] = ave; out[pi + 1] = ave;
        }
    }
}

/* Force per element with Kutta override (src/FLOWPanel_postprocess.jl:1199-1307). */
void orc_calcfield_F(const double *areas, const double *normals, const double *cps_, int64_t n,
                     double uinf, double rho, const int64_t *shedding, int64_t ns, int kutta, double *out)
{
    for (int64_t i = 0; i < n; ++i) {
        double val = -0.5 * rho * uinf * uinf * cps_[i] * areas[i];
        out[3 * i] += val * normals[3 * i];
        out[3 * i + 1] += val * normals[3 * i + 1];
        out[3 * i + 2] += val * normals[3 * i + 2];
    }
    if (!kutta || !shedding) return;
    double q = 0.5 * rho * uinf * uinf;
    for (int64_t e = 0; e < ns; ++e) {
        const int64_t *s = shedding + 6 * e;
        int64_t pi = s[0], pj = s[3];
        int64_t ids[4];
        int cnt;
        double ave;
        if (pj != -1) {
            ave = (cps_[pi] + cps_[pi + 1] + cps_[pj] + cps_[pj - 1]) / 4;
            ids[0] = pi; ids[1] = pi + 1; ids[2] = pj; ids[3] = pj - 1; cnt = 4;
        } else {
            ave = (cps_[pi] + cps_[pi + 1]) / 2;
            ids[0] = pi; ids[1] = pi + 1; cnt = 2;
        }
        for (int k = 0; k < cnt; ++k) {
            int64_t c = ids[k];
            for (int d = 0; d < 3; ++d) out[3 * c + d] = -ave * q * areas[c] * normals[3 * c + d];
        }
    }
}

/* Surface velocity from the gradient of the doublet strength, cell-centred
 * Green-Gauss scheme (src/FLOWPanel_postprocess.jl:139-244; RigidWakeBody
 * trailing-edge treatment :245-497).  For every edge k -> k+1 of cell c with a
 * neighbour m across it (neighbor[c*maxnv + k], -1 if none):
 *   r = p_{k+1} - p_k, d = (r x n_c)/|r x n_c|,
 *   f = |(p_k - cp_m) x r| / (|(p_k - cp_c) x r| + |(p_k - cp_m) x r|),
 *   out_c -= 0.5 sgn d (f G_c + (1-f) G_m) |r| / A_c,   sgn = sign(CPoffset) (1 if 0).
 * With te != 0 every neighbour term of the shedding panels pi, pj is added
 * back, once per appearance in the shedding list.  Polygons use all nv edges
 * (the reference's grids are triangles, for which this is its formula). */
static void ugm_edge(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx, const int64_t *neighbor,
                     int maxnv, const double *G, const double *normals, const double *cps, const double *areas,
                     double sgn, int64_t c, int k, double *acc)
{
    const int64_t *p = cell_idx + cell_ptr[c];
    const int nv = (int)(cell_ptr[c + 1] - cell_ptr[c]);
    const int64_t m = neighbor[c * maxnv + k];
    if (m < 0) return;
    const double *a = nodes + 3 * p[k], *b = nodes + 3 * p[(k + 1) % nv];
    const double *n = normals + 3 * c;
    double r[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double d[3] = {r[1] * n[2] - r[2] * n[1], r[2] * n[0] - r[0] * n[2], r[0] * n[1] - r[1] * n[0]};
    double dm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double vm[3], vn[3], cm[3], cn[3];
    for (int i = 0; i < 3; ++i) { vm[i] = a[i] - cps[3 * c + i]; vn[i] = a[i] - cps[3 * m + i]; }
    cm[0] = vm[1] * r[2] - vm[2] * r[1]; cm[1] = vm[2] * r[0] - vm[0] * r[2]; cm[2] = vm[0] * r[1] - vm[1] * r[0];
    cn[0] = vn[1] * r[2] - vn[2] * r[1]; cn[1] = vn[2] * r[0] - vn[0] * r[2]; cn[2] = vn[0] * r[1] - vn[1] * r[0];
    double dmain = sqrt(cm[0] * cm[0] + cm[1] * cm[1] + cm[2] * cm[2]);
    double dnear = sqrt(cn[0] * cn[0] + cn[1] * cn[1] + cn[2] * cn[2]);
    double f = dnear / (dmain + dnear);
    double face = f * G[c] + (1.0 - f) * G[m];
    double mag = face * sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) / areas[c];
    for (int i = 0; i < 3; ++i) acc[i] += 0.5 * sgn * d[i] / dm * mag;
}

void orc_ugradmu_cell(const double *nodes, const int64_t *cell_ptr, const int64_t *cell_idx, int64_t ncells,
                      const int64_t *neighbor, int maxnv, const double *G, const double *normals, const double *cps,
                      const double *areas, double cpoffset, const int64_t *shedding, int64_t ns, int te,
                      double *out)
{
    const double sgn = cpoffset == 0 ? 1.0 : (cpoffset > 0 ? 1.0 : -1.0);
    for (int64_t c = 0; c < ncells; ++c) {
        const int nv = (int)(cell_ptr[c + 1] - cell_ptr[c]);
        double acc[3] = {0, 0, 0};
        for (int k = 0; k < nv; ++k)
            ugm_edge(nodes, cell_ptr, cell_idx, neighbor, maxnv, G, normals, cps, areas, sgn, c, k, acc);
        for (int i = 0; i < 3; ++i) out[3 * c + i] -= acc[i];
    }
    if (!te || !shedding) return;
    for (int64_t e = 0; e < ns; ++e) {
        for (int side = 0; side < 2; ++side) {
            const int64_t c = shedding[6 * e + 3 * side];
            if (c < 0) continue;
            const int nv = (int)(cell_ptr[c + 1] - cell_ptr[c]);
            double acc[3] = {0, 0, 0};
            for (int k = 0; k < nv; ++k)
                ugm_edge(nodes, cell_ptr, cell_idx, neighbor, maxnv, G, normals, cps, areas, sgn, c, k, acc);
            for (int i = 0; i < 3; ++i) out[3 * c + i] += acc[i];
        }
    }
}
