// fp_internal.h -- shared declarations of the gfx950 panel-method library.
//
// Device data layout (all structure-of-arrays, FP64 master copies):
//   panel vertex table  pv[(3*k + d) * N + j]   vertex k (0..NVP-1), coordinate d, panel j.
//                       Panels with fewer than NVP vertices repeat their first
//                       vertex; that closing edge has zero length and adds
//                       exactly nothing to a vortex ring.
//   panel frame table   hs[c * N + j]: Hess-Smith frame ex(3) ey(3) ez(3) O(3),
//                       local vertex coords (3*NVP), edge d/S/C (3*NVP).
//   geometry            geo[c * N + j]: nx ny nz cx cy cz area.
//   horseshoes          hh[c * nh + h]: pa(3) pb(3) da(3) db(3), hcol[h] = panel.
//   dense matrices      row-major, A[i * ld + j] = G(i, j) (row i = control point,
//                       column j = panel): row interchanges of the LU are then
//                       contiguous and the AIC store of a wavefront (64 panels of
//                       one target row) is one coalesced segment.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <string>
#include <vector>

#include "../../include/flowpanel_hip.h"

#define FP_PI 3.14159265358979323846
#define FP_MAX_NV 8
#define FP_HS_STRIDE(nvp) (12 + 6 * (nvp))
#define FP_GEO_STRIDE 7

struct fp_prof_rec {
    std::string name;
    hipEvent_t a, b;
};

struct fp_pool_blk {
    void *p;
    size_t bytes;
};
// device buffers of at least FP_POOL_MIN bytes are kept by the context when the
// call that used them ends, for the next request of the same size (repeated
// solves of one body re-use their AIC storage without hipMalloc/hipFree); at
// most FP_POOL_MAX bytes are kept, oldest released first
constexpr size_t FP_POOL_MIN = (size_t)16 << 20;
constexpr size_t FP_POOL_MAX = (size_t)64 << 30;

struct fp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cu = 256;
    // second stream for set-up work that runs beside the main stream (created on
    // first use, destroyed with the context)
    hipStream_t side = nullptr;
    // third stream: look-ahead Gauss-Jordan chunks of the block-Jacobi inverses
    hipStream_t side2 = nullptr;
    // fourth stream: block-Jacobi inverses formed beside the assembly
    hipStream_t side3 = nullptr;
    // fork/join events of the look-ahead Gauss-Jordan passes, kept for the
    // context's life (destroying an event that is still pending holds the host)
    hipEvent_t gj_fork = nullptr, gj_join = nullptr;
    // the same for the coarse-space inverse's look-ahead (on the fourth stream)
    hipEvent_t cgj_fork = nullptr, cgj_join = nullptr;
    // cached scratch (grown on demand, freed with the context)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // per-kernel device timing (fp_ctx_profile)
    bool prof = false;
    std::string prof_only;  // when set, only scopes of this name record events
    std::vector<fp_prof_rec> prof_pending;
    std::vector<hipEvent_t> prof_pool;  // recycled events
    std::map<std::string, std::pair<int64_t, double>> prof_acc;
    // released large device buffers (fp_pool_put / fp_pool_trim)
    std::vector<fp_pool_blk> pool;
    size_t pool_bytes = 0;
    ~fp_ctx();
};
void fp_pool_put(fp_ctx *ctx, void *p, size_t bytes);
void fp_pool_trim(fp_ctx *ctx);

// Brackets the launches of one kernel with events on the context stream when
// profiling is on (fp_ctx_profile); the elapsed times are folded in on query.
struct ProfScope {
    fp_ctx *c;
    int idx = -1;
    ProfScope(fp_ctx *ctx, const char *name) : c(ctx)
    {
        if (!name || !c->prof || (!c->prof_only.empty() && c->prof_only != name)) return;
        fp_prof_rec r;
        r.name = name;
        if (!take(r.a)) return;
        if (!take(r.b)) { c->prof_pool.push_back(r.a); return; }
        (void)hipEventRecord(r.a, c->stream);
        c->prof_pending.push_back(r);
        idx = (int)c->prof_pending.size() - 1;
    }
    ~ProfScope()
    {
        if (idx >= 0) (void)hipEventRecord(c->prof_pending[idx].b, c->stream);
    }
    bool take(hipEvent_t &e)
    {
        if (!c->prof_pool.empty()) { e = c->prof_pool.back(); c->prof_pool.pop_back(); return true; }
        return hipEventCreate(&e) == hipSuccess;
    }
};

struct fp_body {
    fp_ctx *ctx = nullptr;
    int element = FP_ELEM_VORTEXRING;
    int64_t N = 0;
    int nvp = 4;
    double kernel_offset = 1e-8, kernel_cutoff = 1e-14, cp_offset = 1e-14;
    int charlen = FP_CHARLEN_UNITARY;
    // host copies (0-based)
    std::vector<double> nodes;
    std::vector<int64_t> cell_ptr, cell_idx;
    std::vector<int64_t> shedding;  // 6 per edge, edge-major
    int64_t ns = 0;
    std::vector<double> das, dbs;   // 3 per edge
    // device
    double *d_pv = nullptr;
    int *d_nv = nullptr;
    double *d_hs = nullptr;
    double *d_geo = nullptr;
    double *d_uinf = nullptr;       // 3 x N SoA
    double *d_hh = nullptr;         // horseshoes
    int *d_hcol = nullptr;
    int nh = 0;
    double *d_strength = nullptr;
    bool has_uinf = false, has_wake = false, solved = false;
    uint64_t version = 0;  // bumped whenever device geometry or wake buffers change
    int *d_nbr = nullptr;         // cell neighbours across each edge (built on first use)
    int64_t *d_sh = nullptr;      // shedding table on the device (built on first use)
    int32_t *d_eb = nullptr;      // shared node/edge assembly blocks (triangles and quads)
    int64_t eb_nblocks = 0, eb_nodes = 0, eb_edges = 0;
    int32_t *d_pb = nullptr;      // the same blocks with their edges paired per lane (edge-pair assembly)
    int64_t pb_nblocks = 0;
    int32_t *d_tr_tab = nullptr;  // edge trails of the mesh (node refs, contributor CSR)
    double *d_tr = nullptr;       // trail stream filled from the strengths before each field evaluation
    int64_t tr_S = 0;
};

struct fp_mat {
    fp_ctx *ctx = nullptr;
    int64_t n = 0, ld = 0;
    int precision = FP_F64;         // storage of A
    void *A = nullptr;              // working copy (unfactorised until `factored`)
    double *A64 = nullptr;          // FP64 original kept for residuals (uploaded matrices)
    int *ipiv = nullptr;            // device, 0-based
    int factored = -1;              // -1 none, else precision of the factors in `F`
    void *F = nullptr;              // factors (row-major, same ld)
    std::vector<fp_body *> bodies;  // bodies of an assembled AIC (matrix-free FP64 residual)
    // block-Jacobi preconditioner
    void *BJ = nullptr;
    int *bj_piv = nullptr;
    int bj_block = 0, bj_nblocks = 0;
};

// ---- error helpers --------------------------------------------------------
int fp_fail(fp_ctx *ctx, int code, const std::string &msg);
int fp_hip_fail(fp_ctx *ctx, hipError_t e, const char *what);
#define FP_HIP(ctx, call)                                                  \
    do {                                                                   \
        hipError_t _e = (call);                                            \
        if (_e != hipSuccess) return fp_hip_fail((ctx), _e, #call);        \
    } while (0)

// ---- views passed to kernels ---------------------------------------------
struct PanelView {
    const double *pv;
    const int *nv;
    const double *hs;
    int64_t N;
    int nvp;
    int element;
    double off, cut;
};
struct TargetView {
    const double *x, *y, *z;     // points
    const double *nx, *ny, *nz;  // normals (may be null when not dotting)
    int64_t M;
};
struct HorseView {
    const double *hh;
    const int *col;
    int nh;
};

static inline PanelView panel_view(const fp_body *b)
{
    PanelView v;
    v.pv = b->d_pv; v.nv = b->d_nv; v.hs = b->d_hs; v.N = b->N; v.nvp = b->nvp;
    v.element = b->element; v.off = b->kernel_offset; v.cut = b->kernel_cutoff;
    return v;
}
static inline HorseView horse_view(const fp_body *b)
{
    HorseView h; h.hh = b->d_hh; h.col = b->d_hcol; h.nh = b->nh;
    return h;
}

// one body of a batched assembly: its panels, its control points and its own
// matrix (row-major, ld); items [tile0, tile0 + nbx * nby) of the work queue
struct AssemblyItem {
    PanelView pan;
    TargetView tv;
    void *G;
    int64_t ld;
    int64_t tile0;
    int64_t nbx;
};

// ---- launchers (fp_influence.hip) ---------------------------------------
// geometry of one body (normals, CPs, areas, Hess-Smith frames)
int fp_launch_geometry(fp_ctx *ctx, fp_body *b);
// AIC column block of one body into a row-major matrix (T = float|double); with
// G32 set, the same entries rounded to FP32 go to G32 (same ld) as well
int fp_gemm_sub_batched_f32(fp_ctx *ctx, int64_t m, int64_t n, int64_t k, const float *A, int64_t lda, int64_t sA,
                            const float *B, int64_t ldb, int64_t sB, float *C, int64_t ldc, int64_t sC, int batch);
int fp_body_build_edge_blocks(fp_ctx *ctx, fp_body *b);
// The diagonal blocks (block x block, identity-padded) of a vortex-ring body's FP32 AIC with its wake
// columns, compact (block b at X + b * block^2), for the block-Jacobi inverses.
int fp_launch_assemble_diag(fp_ctx *ctx, const fp_body *b, const TargetView &tv, int block, float *X);
int fp_body_build_trails(fp_ctx *ctx, fp_body *b);
int fp_launch_field_trails(fp_ctx *ctx, const fp_body *b, const double *strength, const TargetView &tv, int mode,
                           double *out, double *partial, int splits, int *used);
template <typename T>
int fp_launch_assemble_edges(fp_ctx *ctx, const fp_body *b, const TargetView &tv, T *G, int64_t ld, int64_t col0,
                             float *G32, int *used);
template <typename T>
int fp_launch_assemble(fp_ctx *ctx, const fp_body *b, const TargetView &tv, T *G, int64_t ld, int64_t col0,
                       float *G32 = nullptr);
// wake horseshoes of one body only (added into its columns)
template <typename T>
int fp_launch_assemble_wake(fp_ctx *ctx, const fp_body *b, const TargetView &tv, T *G, int64_t ld, int64_t col0);
// tiles of one body in the assembly queue (256 panels x tm targets)
int fp_assembly_queue_tiles(int64_t N, int64_t M, int tm, int *nbx, int64_t *ntiles);
// persistent queue assembly of the vortex rings of several bodies (no wake);
// `counter` is 8 bytes of device memory, reset by the launch itself
int fp_launch_assemble_queue(fp_ctx *ctx, const AssemblyItem *d_items, int nitems, int64_t total, int nvp, int tm,
                             int prec, int *counter, int blocks);
// Field sums of one body over targets.  mode 0: out3 (3 x M point-major, out3[3i + d]) += U;
// mode 1: out1[i] += n_i . U (strengths `x`);  mode 2: out1 += phi.
int fp_launch_field(fp_ctx *ctx, const fp_body *b, const double *strength, const TargetView &tv, int mode,
                    double *out, double *partial, size_t partial_bytes);
size_t fp_field_partial_bytes(int64_t M, int64_t N);
// element forces, integrated force and moment, sectional loads (fp_post.hip)
int fp_launch_loads(fp_ctx *ctx, const fp_body *b, const double *cp, double q, const double *X, const double *Xac,
                    double *F, double *part, double *out6);
size_t fp_loads_partial_bytes(int64_t N);
int fp_launch_sectional(fp_ctx *ctx, const double *F, const double *X, int64_t N, int64_t nchord, int64_t nspan,
                        const double *spandir, double *outF, double *pos);
// FP32 LU of a randomized matrix without pivoting (fp_lu_nopiv.hip)
int64_t fp_rbt_size(int64_t n);
size_t fp_rbt_weights_bytes(int64_t n);
int fp_rbt_weights(fp_ctx *ctx, int64_t n, uint64_t seed, double *d_W);
template <typename Tin>
int fp_rbt_matrix(fp_ctx *ctx, const Tin *A, int64_t n, int64_t lda, float *Ar, int64_t ldr, const double *d_W);
int fp_rbt_vec(fp_ctx *ctx, const double *in, int64_t ldin, double *out, int64_t ldout, int64_t n, int64_t k,
               const double *d_W, int transpose_u);
size_t fp_getrf_nopiv_work_bytes();
int fp_getrf_nopiv_f32(fp_ctx *ctx, float *A, int64_t n, int64_t ld, int *info_dev, void *work, float *YL = nullptr,
                       float *YU = nullptr);
int fp_trtri_nopiv_f32(fp_ctx *ctx, const float *A, int64_t n, int64_t ld, float *YL, float *YU, float *T);
// One body of the batched FP32 no-pivot LU (a device array of these, one per body):
// A (n x n, ld, row-major), the LU's work (fp_getrf_nopiv_work_bytes), its info word.
struct fp_lu_desc {
    float *A;
    int64_t ld, n;
    float *work;
    int *info;
};
// LU without pivoting of every body's matrix, one launch per step for all bodies
// (n_host: the bodies' orders, as in the device table)
int fp_getrf_nopiv_f32_batched(fp_ctx *ctx, const fp_lu_desc *d_desc, const int64_t *n_host, int nb);
// the batched LU's trailing updates of outer block [kk, kk + ob) (fp_dense.hip)
int fp_gemm_sub_lub(fp_ctx *ctx, const fp_lu_desc *d_desc, int nb, int64_t kk, int ob, int64_t max_right);
// E^-1 in place (FP64, n x n) through the randomized no-pivot FP32 LU and the explicit
// inverses of its factors; allocates nothing, never synchronises the host
size_t fp_inverse_rbt_work_bytes(int64_t n);
int fp_inverse_rbt_f32(fp_ctx *ctx, double *E, int64_t n, int64_t lde, void *work, size_t work_bytes, int *info_dev,
                       uint64_t seed);
// workspace of fp_coarse_setup for nc coarse vectors
size_t fp_coarse_work_bytes(int64_t nc);
// the coarse apply's vector workspace (tc) for n rows and nc coarse vectors
size_t fp_coarse_tc_bytes(int64_t n, int nc);
// device address of fp_uind_noalloc's scratch array, queried once per GPU (fp_noalloc.hip)
void fp_noalloc_init(fp_ctx *ctx);

// ---- dense linear algebra (fp_dense.hip) --------------------------------
// P A = L U in place (row-major); perm[i] = original row now at row i.
template <typename T>
int fp_getrf(fp_ctx *ctx, T *A, int64_t n, int64_t ld, int *perm, int *info_dev, void *work, size_t work_bytes);
size_t fp_getrf_work_bytes(int64_t n, int elem_bytes);
// solve with LU factors: x (FP64, n x k SoA columns, ldx) overwritten.  The
// inverses of the diagonal blocks are computed into `work`; reuse_diag skips
// that when the previous call with the same `work` used the same factors.
template <typename T>
int fp_getrs(fp_ctx *ctx, const T *LU, int64_t n, int64_t ld, const int *ipiv, double *x, int64_t k, int64_t ldx,
             void *work, size_t work_bytes, bool reuse_diag = false);
size_t fp_getrs_work_bytes(int64_t n, int64_t k);
// y = alpha*A*x + beta*y, A row-major (T), x,y FP64
template <typename T>
int fp_gemv(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, const double *x, double *y, double alpha, double beta);
// C -= A B (row-major; MFMA tiles)
template <typename T>
int fp_gemm_sub(fp_ctx *ctx, int64_t m, int64_t n, int64_t k, const T *A, int64_t lda, const T *B, int64_t ldb, T *C,
                int64_t ldc);
// least squares of the watertight formulation: C (= -G^T G) -> normal matrix with the
// prescribed rows/columns replaced by the identity, b[p] = val_p
int fp_lsq_fix(fp_ctx *ctx, double *C, int64_t n, int64_t ld, const int64_t *pidx, int np, double *b,
               const double *pval);
int fp_lsq_rhs(fp_ctx *ctx, const double *G, int64_t n, int64_t ld, const int64_t *pidx, const double *pval, int np,
               double *rhs);
// Y[:, c] = A X[:, c], c < k (X, Y column-major)
template <typename T>
int fp_gemv_multi(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, const double *X, int64_t ldx, int64_t k, double *Y,
                  int64_t ldy);
// out[c] = |R[:, c]|^2 (device)
int fp_colnorm2(fp_ctx *ctx, const double *R, int64_t n, int64_t ld, int64_t k, double *out);
// Block-Jacobi inverses straight from A, block <= 1024: Gauss-Jordan in registers
// up to 256 rows, above that blocked Gauss-Jordan over chunks of at most 256
// rows (scratch of fp_bj_gj_scratch_floats floats), and their application
size_t fp_bj_gj_scratch_floats(int64_t n, int block);
constexpr int FP_COARSE_MAX = 4096;   // largest Gauss-Jordan block (the coarse-space inverse)
template <typename T>
int fp_bj_gj_inverse(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, int block, float *inv, int *info_dev,
                     float *scratch, hipEvent_t copied = nullptr, bool lookahead = true, bool prefilled = false);
int fp_bj_apply_inv(fp_ctx *ctx, const float *inv, int64_t n, int block, const double *x, double *y, int accumulate);
// Galerkin coarse space of chunk indicators: AZ = A Z (n x nc, FP32 rounding of FP64 sums),
// Einv = (Z^T AZ)^-1 from that same AZ (so that Z^T (v - AZ Einv Z^T v) = 0)
template <typename T>
int fp_coarse_setup(fp_ctx *ctx, const T *A, int64_t n, int64_t ld, int cs, float *AZ, double *Einv, double *work,
                    int *info_dev, bool lookahead = false);
// y = Z Einv Z^T v and r2 = v - AZ Einv Z^T v (tc: 2*nc scratch)
int fp_coarse_apply(fp_ctx *ctx, const double *v, int64_t n, int cs, const float *AZ, const double *Einv, double *tc,
                    double *y, double *r2);
// small vector kernels
int fp_vec_copy_convert(fp_ctx *ctx, const double *src, void *dst, int64_t n, int dst_prec);
int fp_mat_convert(fp_ctx *ctx, const double *src, int64_t n, int64_t lds, void *dst, int64_t ldd, int dst_prec,
                   int transpose);
int fp_mat_narrow(fp_ctx *ctx, const double *src, float *dst, int64_t n, int64_t ld);
int fp_rhs_kernel(fp_ctx *ctx, const fp_body *b, double *rhs);   // rhs_i = -Uinf_i . n_i
int fp_rhs_from(fp_ctx *ctx, const fp_body *b, const double *uinf_soa, double *rhs);  // the same for a 3 x N SoA Uinf
int fp_residual_update(fp_ctx *ctx, const double *b, double *r, int64_t n);   // r = b - r
int fp_dot(fp_ctx *ctx, const double *x, const double *y, int64_t n, double *d_out);
int fp_axpy(fp_ctx *ctx, double a, const double *x, double *y, int64_t n);
int fp_fill(fp_ctx *ctx, double *x, int64_t n, double v);   // x[:] = v (a kernel: safe inside graphs)
int fp_fill_f32(fp_ctx *ctx, float *x, int64_t n, float v);

// multi-dot out[c] = sum_i V[c*ldv + i] w[i] (part: 256*nv doubles)
int fp_mdot(fp_ctx *ctx, const double *V, int64_t ldv, int nv, const double *w, int64_t n, double *part,
            double *out);
// fused classical Gram-Schmidt pass: w -= V (V^T w) through per-chunk partials
// (FP_CGS_PARTS chunks; at most FP_CGS_MAXV basis vectors);
// with normpart set, per-256-row sums of squares of the new w (Arnoldi norm)
constexpr int FP_CGS_PARTS = 32, FP_CGS_MAXV = 256;
size_t fp_cgs_work_doubles(int64_t n, int m);
int fp_cgs_pass(fp_ctx *ctx, const double *V, int64_t ldv, int nv, double *w, int64_t n, double *part,
                double *normpart);
// Hessenberg column j from the two passes' partials and the norm; Givens update; v_{j+1} = w / h_{j+1,j}
int fp_arnoldi_step2(fp_ctx *ctx, const double *part1, const double *part2, const double *normpart, int m, int j,
                     double bnorm, double *H, double *cs, double *sn, double *g, double *relh, double *hnv,
                     const double *w, double *vnext, int64_t n);

// ---- surface post-processing (fp_post.hip) --------------------------------
// out (3 x N point-major) += -1/2 grad(mu) of one solved body, cell-centred
// Green-Gauss; d_nbr: neighbour across edge k of cell j at [k*N + j]
int fp_launch_ugradmu(fp_ctx *ctx, const fp_body *b, const int *d_nbr, const int64_t *d_sh, int te, double *out);
// sharp trailing edge of the node-based scheme: side (1, 2 or 0) of every cell,
// the TE node list, node -> TE index (-1 elsewhere), and room for the 2 x nte side values
struct TeSides {
    const int *cell_side = nullptr;
    const int *te_nodes = nullptr;
    const int *node_te = nullptr;
    double *sidev = nullptr;
    int nte = 0;
};
int fp_launch_ugradmu_node(fp_ctx *ctx, const fp_body *b, const int64_t *d_nptr, const int *d_ncell, int64_t nn,
                           const int *d_cn, double *d_nodal, double *out, const TeSides *te = nullptr);
int fp_launch_ugradmu_combine(fp_ctx *ctx, const fp_body *b, const int *d_nbr, const int64_t *d_sh, int force_cell_te,
                              double cos_crit, double *cellv, double *nodev, double *out);

// ---- lift/drag adjoint (fp_adjoint.hip) ----------------------------------
// Per-panel arrays are SoA over all M panels of the bodies (panel i = col0 + j
// of its body); the two objectives (CL, CDi) are stacked, component 3m + d.
// normals, areas, control points and probes of one body
int fp_adj_geometry(fp_ctx *ctx, const fp_body *b, double poff, int pkind, double pscale, int64_t M, int64_t col0,
                    double *nrm, double *area, double *cpt, double *prb);
// V (3 x M point-major, induced velocity) += Uinf; Cp = 1 - |V|^2/|Uinf|^2
int fp_adj_cp(fp_ctx *ctx, double *V, int64_t M, const double *Uinf, double *Cp);
// Kutta averaging of c0 (reverse = 0) or its reverse on c0 and c1
int fp_adj_kutta(fp_ctx *ctx, const int64_t *d_sh, int64_t ns, int mode, double *c0, double *c1, int reverse);
int fp_adj_force_terms(fp_ctx *ctx, const double *Cp, const double *area, const double *nrm, int64_t M, double *T);
int fp_adj_objective(fp_ctx *ctx, const double *Cp, const double *area, const double *nrm, int64_t M,
                     const double *e0, const double *e1, double inv_s, double *Cpb, double *Ab, double *nb);
int fp_adj_vbar(fp_ctx *ctx, const double *Cpb, const double *V, int64_t M, double inv_u2, double *W);
size_t fp_adj_pairs_work_bytes(fp_ctx *ctx, int64_t N, int nvp, int64_t MT);
// reverse of the rings of one body at MT targets with weights W (6 x MT):
// Xb (6 x MT) +=, vbar (6*nvp x N) +=, and (gamma) g[m*ldg + j] += u_j . w
int fp_adj_pairs(fp_ctx *ctx, const fp_body *b, const double *tx, int64_t MT, const double *W, int gamma,
                 double *Xb, double *vbar, double *g, int64_t ldg, void *work, size_t work_bytes);
// the same for the body's wake horseshoes; hpart (nh x 26) receives the
// unit-strength horseshoe parameter adjoints (and strength adjoints)
int fp_adj_wake(fp_ctx *ctx, const fp_body *b, const double *tx, int64_t MT, const double *W, int gamma, double *Xb,
                double *hpart, double *g, int64_t col0);
int fp_adj_cpw(fp_ctx *ctx, const double *lam, const double *nrm, const double *Uc, int64_t M, const double *Uinf,
               double *W2, double *nb);
int fp_adj_geom_rev(fp_ctx *ctx, const fp_body *b, double poff, int pkind, double pscale, int64_t M, int64_t col0,
                    const double *nb, const double *Xbp, const double *Xbc, const double *Ab, double *vbar);
int fp_adj_nodes(fp_ctx *ctx, const fp_body *b, const int64_t *ptr, const int64_t *ent, int64_t nn,
                 const double *vbar, double *out);
// B = A^T (row-major n x n, leading dimension ld for both), FP64 or narrowed to FP32
int fp_mat_transpose(fp_ctx *ctx, const double *A, int64_t n, int64_t ld, void *B, int dst_prec);
// column p_q of B += w_q (w: np x n) -- the border of the least-squares solve
int fp_lsq_border(fp_ctx *ctx, double *B, int64_t n, int64_t ld, const int64_t *pidx, int np, const double *w);

// ---- Krylov (fp_abi.hip) ---------------------------------------------------
