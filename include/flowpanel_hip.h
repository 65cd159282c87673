/*
 * flowpanel_hip.h -- C ABI of the MI355X (gfx950) panel-method library.
 *
 * Drop-in boundary for the hot path of FLOWPanel.jl: body constructors,
 * `solve`, the solver-function protocol, `Uind!`/`phi!`, `calcfield_U`/
 * `calcfield_Cp`.  The reference binds these through Julia multiple dispatch;
 * a Julia host calls this library with `ccall` (one `ccall` per entry point,
 * plain pointers and sizes), the Python host mirror in flowpanel.jl_amd/ calls
 * it through ctypes.
 *
 * Conventions (the reference's):
 *   - 3-vector arrays are column-major 3 x n float64 (x, y, z contiguous per
 *     point): nodes, targets, normals, control points, Uinfs, Das, Dbs.
 *   - Matrices exchanged with the host are column-major (column j = panel j).
 *   - Cells are CSR (cell_ptr[ncells+1], cell_idx) with `index_base` 0 or 1
 *     (1 = Julia).  `shedding` is 6 x ns column-major (pi, nia, nib, pj, nja,
 *     njb): panel index, local vertex indices of its trailing edge, partner
 *     panel (-1 = none, whatever the base) and its local vertex indices.
 *   - Outputs documented as "+=" accumulate into caller memory, as the
 *     reference's kernels do (src/FLOWPanel_postprocess.jl:25-26).
 *   - Host pointers are borrowed for the duration of a call.  Device state is
 *     owned by the handles and freed by the matching destroy call.
 *   - A context owns one HIP stream and is not thread-safe; distinct contexts
 *     may be used from distinct threads.
 *   - Every function returns FP_OK or one of the FP_ERR_* codes below, and
 *     fp_ctx_last_error() describes the last failure.
 */
#ifndef FLOWPANEL_HIP_H
#define FLOWPANEL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FP_OK 0
#define FP_ERR_ARG 1               /* null pointer / bad enum / bad index                     */
#define FP_ERR_SHAPE 2             /* size mismatch: src/FLOWPanel_liftingbody.jl:180-189      */
#define FP_ERR_NONUNIT_DIRECTION 3 /* non-unit wake direction: src/FLOWPanel_elements.jl:495-498 */
#define FP_ERR_BAD_DIRECTIONS 4    /* da.db <= 0: src/FLOWPanel_elements.jl:1130-1133          */
#define FP_ERR_UNSOLVED 5          /* Uind!/phi! before solve: src/FLOWPanel_abstractbody.jl:104-107 */
#define FP_ERR_SINGULAR 6          /* exactly zero pivot in getrf (LAPACK info > 0)            */
#define FP_ERR_NOT_CONVERGED 7     /* refinement / GMRES missed the requested tolerance        */
#define FP_ERR_DEVICE 8            /* HIP runtime error                                        */
#define FP_ERR_NOMEM 9             /* device allocation failed                                 */
#define FP_ERR_UNSUPPORTED 10      /* valid request this library does not implement            */

/* ---- element types (src/FLOWPanel_elements.jl:12-17) ------------------- */
#define FP_ELEM_SOURCE 0     /* ConstantSource                                    */
#define FP_ELEM_VORTEXRING 1 /* VortexRing == ConstantDoublet (for the velocity)   */

/* ---- arithmetic precision of a dense matrix ----------------------------- */
#define FP_F64 0
#define FP_F32 1

/* ---- characteristic length of a panel (src/FLOWPanel_abstractbody.jl:573-633) */
#define FP_CHARLEN_UNITARY 0
#define FP_CHARLEN_BBOX 1
#define FP_CHARLEN_MAXDIST 2
#define FP_CHARLEN_SQRTAREA 3

/* ---- linear solvers (the solver-function protocol,
 *      src/FLOWPanel_linearsolver.jl:24-157) -------------------------------- */
#define FP_SOLVER_LU 0       /* solve_ludiv!/solve_backslash!: FP64 getrf + getrs            */
#define FP_SOLVER_LU32_IR 1  /* FP32 getrf (MFMA) + FP64 iterative refinement                */
#define FP_SOLVER_GMRES_IR 2 /* FP32 block-Jacobi GMRES inside FP64 iterative refinement     */
#define FP_SOLVER_GMRES 3    /* solve_gmres!: FP64 GMRES (memory, atol = rtol = sqrt(eps))   */

/* Kutta averaging of Cp at the trailing edge (src/FLOWPanel_postprocess.jl:1075-1097) */
#define FP_KUTTA_NONE 0
#define FP_KUTTA_REFERENCE 1 /* average (pi, pi+1, pj, pj-1): structured triangulated grids */
#define FP_KUTTA_PAIR 2      /* average (pi, pj): structured quad grids                     */

typedef struct fp_ctx fp_ctx;
typedef struct fp_body fp_body;
typedef struct fp_mat fp_mat;

typedef struct fp_solve_opts {
    int solver;        /* FP_SOLVER_*                                                     */
    double tol;        /* target ||b - A x|| / ||b|| of the FP64 result                   */
    double inner_tol;  /* relative tolerance of each inner GMRES solve (GMRES_IR)         */
    int restart;       /* GMRES memory (Krylov basis size before restart)                 */
    int max_iters;     /* cap on the total number of GMRES iterations                     */
    int max_refine;    /* cap on refinement steps                                          */
    int precond_block; /* block-Jacobi block size, at most 1024 (GMRES_IR, GMRES)         */
    int coarse_block;  /* > 0: add a Galerkin coarse correction on the indicator vectors of
                          consecutive chunks of this many panels (e.g. a fifth of a
                          chordwise strip); at most 4096 chunks                              */
    int residual_mode; /* refinement residual r = b - A x in FP64 from: FP_RESID_STORED (the
                          FP64 AIC kept beside the FP32 copy) or FP_RESID_MATRIX_FREE
                          (influences recomputed in FP64, no FP64 matrix in memory)          */
    int precond_order; /* which panels form the block-Jacobi blocks and coarse chunks:
                          FP_ORDER_INDEX (consecutive indices: structured grids) or
                          FP_ORDER_SPATIAL (consecutive along a Morton curve of the control
                          points: unstructured meshes; bodies only, not fp_solve_mat)          */
} fp_solve_opts;

#define FP_RESID_MATRIX_FREE 0
#define FP_RESID_STORED 1
#define FP_ORDER_INDEX 0
#define FP_ORDER_SPATIAL 1

typedef struct fp_solve_stats {
    int refinements;   /* FP64 refinement steps taken                                      */
    int iterations;    /* GMRES iterations (all inner solves)                              */
    double residual;   /* final ||b - A x|| / ||b|| in FP64                                */
    double ms_assemble;/* device time of AIC assembly                                      */
    double ms_factor;  /* device time of the factorisation / preconditioner set-up         */
    double ms_solve;   /* device time of substitution / Krylov / refinement                */
    int info;          /* getrf info (first zero pivot, 1-based) or 0                      */
} fp_solve_stats;

/* Fill `opts` with the defaults of `solver`. */
void fp_solve_opts_default(fp_solve_opts *opts, int solver);

/* ---- context ------------------------------------------------------------ */
int fp_ctx_create(int device, fp_ctx **out);
void fp_ctx_destroy(fp_ctx *ctx);
const char *fp_ctx_last_error(const fp_ctx *ctx);
int fp_ctx_synchronize(fp_ctx *ctx);
/* Free the device buffers (16 MiB and larger) that the context keeps after a call for re-use by the next
 * call of the same size; the context keeps at most 64 GiB of them and also frees them when an allocation
 * fails or the context is destroyed.  No counterpart in the reference (memory management of this
 * library). */
int fp_ctx_trim(fp_ctx *ctx);

/* ---- bodies ------------------------------------------------------------- */
/* NonLiftingBody{ConstantSource}(grid) (src/FLOWPanel_nonliftingbody.jl:32-72) and
 * RigidWakeBody{VortexRing}(grid, shedding; CPoffset, kerneloffset, kernelcutoff,
 * characteristiclength) (src/FLOWPanel_liftingbody.jl:43-134).  Uploads the
 * geometry and computes normals, control points and areas on the device
 * (src/FLOWPanel_abstractbody.jl:635-927).  Panels may have 3..8 vertices. */
int fp_body_create(fp_ctx *ctx, const double *nodes, int64_t nnodes, const int64_t *cell_ptr,
                   const int64_t *cell_idx, int64_t ncells, int index_base, int element,
                   const int64_t *shedding, int64_t nsheddings, double cp_offset, int charlen_kind,
                   double kernel_offset, double kernel_cutoff, fp_body **out);
void fp_body_destroy(fp_body *body);
int64_t fp_body_ncells(const fp_body *body);
int64_t fp_body_nsheddings(const fp_body *body);
/* calc_normals / calc_controlpoints / calc_areas (src/FLOWPanel_abstractbody.jl:706-927);
 * any output may be NULL. */
int fp_body_geometry(fp_ctx *ctx, const fp_body *body, double *normals, double *cps, double *areas);
/* Freestream at every control point (3 x ncells) and trailing-wake directions
 * Das/Dbs (3 x nsheddings each; NULL for a non-lifting body), as passed to
 * solve(body, Uinfs, Das, Dbs) (src/FLOWPanel_liftingbody.jl:172-189). */
int fp_body_set_freestream(fp_ctx *ctx, fp_body *body, const double *Uinfs, const double *Das,
                           const double *Dbs);
int fp_body_set_strength(fp_ctx *ctx, fp_body *body, const double *strength);
int fp_body_get_strength(fp_ctx *ctx, const fp_body *body, double *strength);

/* solve(body, Uinfs, Das, Dbs; solver) for one body or the components of a
 * MultiBody (src/FLOWPanel_liftingbody.jl:172-216, src/FLOWPanel_nonliftingbody.jl:93-130,
 * src/FLOWPanel_multibody.jl:312-364): AIC assembly, no-flow-through RHS,
 * linear solve; the strengths stay on the device in the bodies.  Uses the
 * freestream stored by fp_body_set_freestream.  `opts` may be NULL (LU). */
int fp_solve_bodies(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const fp_solve_opts *opts,
                    fp_solve_stats *stats);

/* solve(RigidWakeBody{VortexRing, 2}, Uinfs, Das, Dbs; elprescribe) -- the
 * watertight least-squares formulation (src/FLOWPanel_liftingbody.jl:236-316,
 * 357-447; elprescribe as calc_elprescribe, src/FLOWPanel_multibody.jl:565-599):
 * the npres prescribed strengths (pres_idx, `index_base` 0 or 1, values pres_val)
 * move to the right-hand side and G_red^T G_red x = G_red^T b is solved with
 * opts->solver; the bodies receive x with the prescribed values in place. */
int fp_solve_bodies_lsq(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const int64_t *pres_idx,
                        const double *pres_val, int64_t npres, int index_base, const fp_solve_opts *opts,
                        fp_solve_stats *stats);

/* The same for k freestreams at once under the fixed-wake convention: the wake
 * directions stored by fp_body_set_freestream are shared by all k cases, so the
 * AIC is one matrix and the freestreams are k right-hand sides (one assembly, one
 * factorisation or preconditioner; the refinement paths refine all columns
 * together).  Uinfs: 3 x N x k (N = sum of ncells, column-major per freestream);
 * strengths: N x k column-major.  The bodies' stored strengths are not changed.
 * The reference solves one freestream per call (src/FLOWPanel_liftingbody.jl:172-216)
 * and takes several RHS columns in its solver protocol (src/FLOWPanel_linearsolver.jl:149-157). */
int fp_solve_bodies_multi(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const double *Uinfs, int64_t k,
                          const fp_solve_opts *opts, double *strengths, fp_solve_stats *stats);

/* A batch of independent bodies (each its own linear system, no mutual
 * influence), e.g. many wings of uneven panel counts.  fp_batch_create sizes
 * the device buffers, times the batched assembly for each tile shape (targets
 * per LDS tile 32/64/128 x persistent workgroups 2/4/8 per CU) and keeps the
 * fastest, then captures one hipGraph of: a persistent work-queue launch that
 * assembles every body's AIC, and per body (on parallel streams) its wake
 * columns, LU factorisation and refinement sweeps.  fp_batch_solve replays the
 * graph (re-capturing first if a body's geometry or wake changed), finishes any
 * body that still needs refinement, and leaves the strengths in the bodies.
 * opts->solver: FP_SOLVER_LU32_IR (FP32 AIC and LU; FP64 residuals, matrix-free
 * with FP_RESID_MATRIX_FREE or from a stored FP64 AIC) or FP_SOLVER_LU.
 * Each body is solve(body, Uinfs, Das, Dbs) (src/FLOWPanel_liftingbody.jl:172-216);
 * the reference has no batch of bodies. */
typedef struct fp_batch fp_batch;
typedef struct fp_batch_stats {
    double ms_graph;       /* device time of one graph replay (assembly + factorisations + sweeps) */
    double ms_capture;     /* re-capture and instantiation before this replay (0 if none)        */
    double ms_extra;       /* eager refinement after the replay                                  */
    double residual;       /* worst FP64 relative residual over the bodies                       */
    int64_t panels;        /* total panels                                                       */
    int nbodies, sweeps_in_graph, extra_sweeps, graph_nodes, recaptured;
    int tile_targets, queue_workgroups;  /* the autotuned assembly shape                          */
} fp_batch_stats;
int fp_batch_create(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const fp_solve_opts *opts, fp_batch **out);
int fp_batch_solve(fp_ctx *ctx, fp_batch *batch, fp_batch_stats *stats);
void fp_batch_destroy(fp_batch *batch);
/* The k-th autotuning candidate (k = 0..8): tile shape and its measured milliseconds. */
int fp_batch_autotune_get(const fp_batch *batch, int k, int *tile_targets, int *workgroups, double *ms);

/* ---- dense AIC and the solver protocol ---------------------------------- */
/* _G_U!/_G_Uvortexring! (src/FLOWPanel_liftingbody.jl:455-572,
 * src/FLOWPanel_nonliftingbody.jl:141-186, src/FLOWPanel_multibody.jl:502-555):
 * device-resident AIC of the bodies with their stored wake directions. */
int fp_assemble(fp_ctx *ctx, fp_body *const *bodies, int nbodies, int precision, fp_mat **out);
/* Host only, no device: the column blocks the vortex-ring assembly of a triangle/quad body uses (mesh
 * node and edge terms computed once per target and block, shared by the block's panels; at most 256
 * panels, 384 nodes and 640 edges per block).  Returns the
 * number of blocks (negative FP_ERR_* on bad input); when tab is not null and cap_blocks is at least
 * that number, writes 2048 int32 per block: 512 node references (panel * 8 + vertex, -1 = none), 768
 * edges (local node a | b << 16, -1 = none), 256 columns (-1 = none), 256 pairs of packed 16-bit edge
 * references ((edge << 1) | reversed, 0xFFFF = none); terms (2 values, may be null) = the nodes and the
 * edges summed over the blocks.  No counterpart in the reference (a layout of this library, exported
 * so that it can be inspected). */
int64_t fp_assembly_blocks(const double *nodes, int64_t nnodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                           int64_t ncells, int index_base, int32_t *tab, int64_t cap_blocks, int64_t *terms);
/* Host only, no device: the edge-pair tables of the alternative vortex-ring assembly (the same kind of
 * column blocks, at most 504 distinct edges, each block's edges paired at shared nodes so that a lane
 * owns a path n0 - n1 - n2: edge n0 -> n1 in slot lane, n1 -> n2 in slot 256 + lane).  Returns the
 * number of blocks (negative FP_ERR_* on bad input); when tab is not null and cap_blocks is at least
 * that number, writes 1536 int32 per block: 256 columns (-1 = none), 256 pairs of packed 16-bit slot
 * references ((slot << 1) | reversed, 0xFFFF = none), then the node references (panel * 8 + vertex,
 * -1 = none) of n0, n1 and n2, 256 each; terms (2 values, may be null) = the lanes used and the edges,
 * summed over the blocks.  No counterpart in the reference (a layout of this library, exported so that
 * it can be inspected). */
int64_t fp_assembly_pairs(const double *nodes, int64_t nnodes, const int64_t *cell_ptr, const int64_t *cell_idx,
                          int64_t ncells, int index_base, int32_t *tab, int64_t cap_blocks, int64_t *terms);
/* Host only, no device: the edge trails the velocity of a vortex-ring body is summed over (every mesh edge
 * once, gamma_e = the signed strengths of the panels that hold it).  Returns the number S of stream
 * entries (negative FP_ERR_* on bad input) and, in *ncontrib, the number of (panel, sign) contributors;
 * when the arrays are given and large enough (cap_entries >= S, cap_contrib >= *ncontrib) writes
 * ref[S] = panel * 8 + vertex of each entry's node, cptr[S + 1] and cpan[*ncontrib] = (panel << 1) |
 * reversed, the contributors of the edge from entry m - 1's node to entry m's (none: a trail starts).
 * No counterpart in the reference (a layout of this library, exported so that it can be inspected). */
int64_t fp_field_trails(const int64_t *cell_ptr, const int64_t *cell_idx, int64_t ncells, int64_t nnodes,
                        int index_base, int32_t *ref, int32_t *cptr, int32_t *cpan, int64_t cap_entries,
                        int64_t cap_contrib, int64_t *ncontrib);
/* Upload a caller matrix (column-major n x n) for the solver protocol. */
int fp_mat_upload(fp_ctx *ctx, int64_t n, const double *A, fp_mat **out);
int64_t fp_mat_size(const fp_mat *mat);
/* The FP32 block-Jacobi inverses the Krylov solvers apply (blocks of `block` rows along the index order,
 * Gauss-Jordan with partial pivoting in chunks of at most 64 rows, the last block padded with the
 * identity) of a matrix that is not factorised: out (host) = ceil(n / b) row-major b x b blocks,
 * b = min(block, n), block <= 1024.  FP_ERR_SINGULAR on a zero pivot.  No counterpart in the reference
 * (the preconditioner of this library, exported so that it can be inspected). */
int fp_mat_block_inverses(fp_ctx *ctx, const fp_mat *mat, int block, float *out);
/* Column-major n x n copy of the (unfactorised) matrix. */
int fp_mat_download(fp_ctx *ctx, const fp_mat *mat, double *A);
void fp_mat_destroy(fp_mat *mat);
/* solver(y, A, b) for k right-hand sides (b, x column-major n x k):
 * solve_ludiv!/solve_backslash!/solve_gmres! (src/FLOWPanel_linearsolver.jl:24-157).
 * The factorisation is cached in `mat` and reused by later calls. */
int fp_solve_mat(fp_ctx *ctx, fp_mat *mat, const double *b, int64_t k, double *x, const fp_solve_opts *opts,
                 fp_solve_stats *stats);
/* getrf on the device: LU factors (row interchanges applied, L unit lower and U
 * packed column-major like LAPACK) and 1-based ipiv (LAPACK convention).
 * `precision` FP_F64 or FP_F32 (factors returned widened to FP64). */
int fp_mat_getrf(fp_ctx *ctx, fp_mat *mat, int precision, double *LU, int64_t *ipiv);

/* ---- field evaluation and post-processing ------------------------------- */
/* Uind!(body, targets, out) (src/FLOWPanel_abstractbody.jl:101-110,
 * src/FLOWPanel_liftingbody.jl:576-661, src/FLOWPanel_nonliftingbody.jl:188-218):
 * out (3 x nt) += velocity induced by the solved bodies. */
int fp_uind(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const double *targets, int64_t nt,
            double *out);
/* phi!(body, targets, out) (src/FLOWPanel_abstractbody.jl:118-127,
 * src/FLOWPanel_liftingbody.jl:664-742): out (nt) += potential. */
int fp_phi(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const double *targets, int64_t nt,
           double *out);
/* calcfield_U(body, body) at the control points (src/FLOWPanel_postprocess.jl:28-110):
 * out (3 x N, N = sum of ncells) += Uinf + induced velocity. */
int fp_calcfield_U(fp_ctx *ctx, fp_body *const *bodies, int nbodies, double *out);
/* calcfield_Cp!(out, body, Us, Uref) (src/FLOWPanel_postprocess.jl:1059-1112):
 * out (n) += 1 - (|U|/Uref)^2, then Kutta averaging over the shedding edges. */
int fp_calcfield_cp(fp_ctx *ctx, const double *Us, int64_t n, double uref, const int64_t *shedding,
                    int64_t nsheddings, int index_base, int kutta_mode, double *out);

/* calcfield_F!(out, body, areas, normals, Cps, Uinf, rho) (src/FLOWPanel_postprocess.jl:1199-1307) with the
 * integrated loads calcfield_Ftot! (:1543-1590) and calcfield_Mtot! (:1715-1784), on the device:
 * F_i = -Cp_i (rho Uinf^2 / 2) A_i n_i, the normals flipped by the sign of CPoffset (the reference's
 * calc_normals(body; flipbyCPoffset=true)).  Cps: N (all bodies one after the other).  kutta_mode: the
 * trailing-edge cells of each vortex-ring body take the average Cp of their group (FP_KUTTA_*), and their F
 * overwrites what `F` held (the other cells add to it), as the reference does.  Outputs may be NULL:
 * F (3 x N), Ftot (3, +=), Mtot (3, +=) the moment about Xac (NULL: the origin) with the force of each cell
 * applied at X (3 x N; NULL: the bodies' control points). */
int fp_calcfield_F(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const double *Cps, double Uinf, double rho,
                   int kutta_mode, const double *Xac, const double *X, double *F, double *Ftot, double *Mtot);
/* calcfield_sectionalforce!(outf, outpos, body, controlpoints, Fs; spandirection) (src/FLOWPanel_postprocess.jl:
 * 1411-1542) for a structured grid of nchord x nspan cells (cell i + nchord j: chordwise index fastest): the
 * force F (3 x N) summed over each spanwise station j, divided by the station spacing (central differences,
 * one-sided at the ends), and the station position, the mean of spandir . X over its cells (X: 3 x N, NULL
 * the control points).  outF: 3 x nspan, outpos: nspan. */
int fp_calcfield_sectionalforce(fp_ctx *ctx, const fp_body *body, const double *F, const double *X, int64_t nchord,
                                int64_t nspan, const double *spandir, double *outF, double *outpos);

/* calcfield_Ugradmu_cell!(out, body) (src/FLOWPanel_postprocess.jl:139-497):
 * out (3 x N, += ) -= 1/2 grad(mu), the surface velocity due to the gradient of
 * the doublet strength, by the cell-centred Green-Gauss scheme with inverse-
 * distance face values; with te_correction the trailing-edge panels drop their
 * neighbour terms (the RigidWakeBody variant).  Cells are neighbours when they
 * share an edge. */
int fp_calcfield_ugradmu(fp_ctx *ctx, fp_body *const *bodies, int nbodies, int te_correction, double *out);
/* Sharp trailing edge of the node-based scheme (sharpTE, src/FLOWPanel_postprocess.jl:692-743): the cells on
 * either side of the trailing edge take, at the TE nodes, the area-weighted mean of Gamma over their own side's
 * cells only, so the jump of Gamma across the trailing edge is not averaged away.  Side 1 is the shedding
 * panels pi, side 2 their partners pj; FP_SHARPTE_PAIRS adds pi + 1 and pj - 1 (the two triangles of a
 * structured triangulated grid's trailing-edge quad, the reference's structured-grid layout), FP_SHARPTE_CELLS
 * takes pi and pj alone (quad grids, unstructured meshes).  The TE nodes are the shedding edges' nodes. */
#define FP_SHARPTE_OFF 0
#define FP_SHARPTE_PAIRS 1
#define FP_SHARPTE_CELLS 2

/* calcfield_Ugradmu_node!(out, body; sharpTE) (src/FLOWPanel_postprocess.jl:622-748):
 * out (3 x N, +=) -1/2 grad(mu) from nodal values of Gamma, each the area-weighted mean over the cells holding
 * the node, and per cell the in-plane gradient of the plane through its first three vertices (sign flipped
 * with CPoffset < 0).  nodal_out (optional): the nodal values ("gamma_node"), the bodies' nodes one after the
 * other; with a sharp trailing edge they hold side 2's values at the TE nodes, as the reference leaves them. */
int fp_calcfield_ugradmu_node(fp_ctx *ctx, fp_body *const *bodies, int nbodies, int sharp_te, double *out,
                              double *nodal_out);
/* calcfield_Ugradmu!(out, out_cell, out_node, body; sharpTE, force_cellTE, anglecrit)
 * (src/FLOWPanel_postprocess.jl:756-1000):
 * out (3 x N, +=) the cell-centred value (with the trailing-edge treatment) of every cell except those with
 * an edge without neighbour or a neighbour whose normal is more than anglecrit_deg away, which take the
 * node-based value (with sharp_te as in fp_calcfield_ugradmu_node); with force_cell_te the shedding panels
 * pi, pi+1, pj, pj-1 keep cell-centred values.  All edges of a cell are examined (the reference loops over
 * three, every edge of a triangle). */
int fp_calcfield_ugradmu_combined(fp_ctx *ctx, fp_body *const *bodies, int nbodies, int force_cell_te,
                                  double anglecrit_deg, int sharp_te, double *out);

/* ---- shape updates ------------------------------------------------------- */
int64_t fp_body_nnodes(const fp_body *body);
/* Move the mesh nodes (3 x nnodes, same connectivity) as the reference does by
 * editing grid._nodes in place: the device geometry and wake horseshoes are
 * recomputed and the body is marked unsolved. */
int fp_body_set_nodes(fp_ctx *ctx, fp_body *body, const double *nodes);

/* ---- lift/drag coefficients and their adjoint ----------------------------
 * The reference's offset-probe force workflow (test/runtests_solvers.jl:136-166)
 * on vortex-ring bodies, as one function of the mesh nodes and Uinf:
 *   solve with Uinfs = Uinf at every control point and wake directions
 *     Das = Dbs = Uinf/|Uinf| (src/FLOWPanel_liftingbody.jl:172-216);
 *   velocity at the probes centroid + probe_offset * l * n, l = probe_length x
 *     the characteristic length of kind probe_charlen (calcfield_U with offset and
 *     characteristiclength, src/FLOWPanel_postprocess.jl:81-99);
 *   Cp = 1 - |U|^2/|Uinf|^2 with Kutta averaging (:1059-1112);
 *   F_i = -Cp_i q A_i n_i (:1199-1307);
 *   CL = F.Lhat/(q S), CDi = F.Dhat/(q S), Dhat = Uinf/|Uinf|,
 *     Lhat = normalize(Dhat x span_dir) (calcfield_LDS, :1615-1652).
 * The Kutta override of calcfield_F re-averages values that are already
 * averaged, so it is the identity here and is not applied twice. */
typedef struct fp_liftdrag_opts {
    double probe_offset;  /* reference test: 0.02                           */
    double probe_length;  /* reference test: b/ar (with FP_CHARLEN_UNITARY) */
    int probe_charlen;    /* FP_CHARLEN_*                                    */
    int kutta_mode;       /* FP_KUTTA_*                                      */
    double span_dir[3];   /* Shat, reference test: (0, 1, 0)                 */
    double ref_area;      /* S, reference test: b^2/ar                       */
    fp_solve_opts solve;  /* forward solve; the adjoint's transpose solve uses the same solver */
} fp_liftdrag_opts;

typedef struct fp_liftdrag_stats {
    double CL, CDi;
    double ms_forward;          /* freestream, assembly, solve, probes, coefficients      */
    double ms_adjoint;          /* reverse kernels, transpose solve, gathers               */
    double ms_transpose_solve;  /* transposition, set-up and solve of A^T lambda = dJ/dG  */
    int iterations_adjoint;
    fp_solve_stats forward;     /* the forward solve                                       */
} fp_liftdrag_stats;

void fp_liftdrag_opts_default(fp_liftdrag_opts *opts, int solver);
/* CL and CDi of the bodies at Uinf (3); stats may be NULL.  The strengths of the
 * solve stay in the bodies, as after fp_solve_bodies. */
int fp_lift_drag(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const double *Uinf,
                 const fp_liftdrag_opts *opts, double *CL, double *CDi, fp_liftdrag_stats *stats);
/* Forward as fp_lift_drag, then the hand-written reverse pass (transpose solve
 * plus the reverse of the geometry, influence and post-processing kernels):
 * dCL/dnodes and dCDi/dnodes (3 x nnodes per body, bodies concatenated) and
 * dCL/dUinf, dCDi/dUinf (3).  Any gradient output may be NULL.  The ImplicitAD
 * rule of the reference's linear solve (src/FLOWPanel_linearsolver.jl:73,128)
 * is the transpose solve here; the solver is opts->solve.solver with the FP64
 * residual stored (mixed precision: FP32 inner solves, FP64 refinement). */
int fp_lift_drag_adjoint(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const double *Uinf,
                         const fp_liftdrag_opts *opts, double *dCL_dnodes, double *dCDi_dnodes,
                         double *dCL_dUinf, double *dCDi_dUinf, fp_liftdrag_stats *stats);

/* ---- field evaluation that allocates nothing ----------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP stream/graph capture, tight time-stepping loops that
 * evaluate Uind! every step, harnesses that forbid hipMalloc.  The same as
 * fp_uind, but on device memory of the context's GPU: d_targets (3 x nt,
 * column-major, as fp_uind's targets) and d_out (3 x nt, +=).  The work is
 * queued on the context stream and the call returns without synchronising.
 * Target staging and the partial sums of the split launch live in one 64 MiB
 * device array per GPU, set up by fp_ctx_create and held until the process
 * ends; nt is limited to FP_UIND_NOALLOC_MAX_TARGETS (larger: FP_ERR_ARG).
 * Every call on one GPU uses that GPU's array: calls that overlap on one GPU
 * (two contexts or streams at once) race. */
#define FP_UIND_NOALLOC_MAX_TARGETS ((int64_t)1 << 20)
int fp_uind_noalloc(fp_ctx *ctx, fp_body *const *bodies, int nbodies, const double *d_targets, int64_t nt,
                    double *d_out);

/* ---- instrumentation (no counterpart in the reference) ------------------
 * Per-kernel device time measured with HIP events on the context stream, for
 * roofline reporting.  fp_ctx_profile(ctx, 1) clears the counters and starts
 * recording; fp_ctx_profile_get returns the launches and total milliseconds
 * recorded so far for one kernel class (e.g. "gemv_f32", "gemm_f64",
 * "assemble_vr_f32", "field_matvec_f64", "panel_tslu", "bj_apply"). */
int fp_ctx_profile(fp_ctx *ctx, int enable);
/* As fp_ctx_profile(ctx, 1), but only launches of one kernel class record
 * events (two per launch), so the timing adds little host work to a timed run. */
int fp_ctx_profile_only(fp_ctx *ctx, const char *kernel);
int fp_ctx_profile_get(fp_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms);
/* One device-to-device copy of `bytes` through the HIP runtime's copy kernel: a
 * plain copy of known size to calibrate HBM byte counters (FETCH_SIZE/WRITE_SIZE). */
int fp_ctx_copy_calibrate(fp_ctx *ctx, int64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* FLOWPANEL_HIP_H */
