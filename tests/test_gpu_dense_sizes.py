"""Dense kernels at the block boundaries of the device LU and triangular solves.

getrf works in 512-column outer blocks of 32-column sub-panels (the outer
U12 = L11^-1 A12 update through an explicit inverse), getrs in 128-row
blocks with 128-row update slices, the block-Jacobi inverses by Gauss-Jordan
in registers up to 128 rows and by blocked Gauss-Jordan over chunks of at most
128 rows up to 1024.  Sizes straddle each of those boundaries.  Bars: P A = L U to 1e-12 of max|A| (FP64), solutions to
1e-11 relative against numpy (FP64 LU and FP32 LU + FP64 refinement), GMRES-IR
to its 1e-12 target.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 128, 129, 511, 512, 513, 1025])
def test_getrf_block_boundaries(fp, n):
    rng = np.random.default_rng(1000 + n)
    A = rng.standard_normal((n, n))
    LU, ipiv = fp.getrf(A)
    L = np.tril(LU, -1) + np.eye(n)
    Uu = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(ipiv - 1):
        PA[[i, p]] = PA[[p, i]]
    assert np.abs(PA - L @ Uu).max() / np.abs(A).max() < 1e-12
    # (tournament pivoting picks each pivot from the winners of a reduction tree, so
    # |l_ij| <= 1 -- the GEPP bound -- need not hold; P A = L U and the solves are the bar)


@pytest.mark.parametrize("n", [1, 33, 128, 129, 257, 640])
@pytest.mark.parametrize("k", [1, 3])
@pytest.mark.parametrize("solver", ["solve_ludiv", "solve_lu32_ir"])
def test_getrs_block_boundaries(fp, n, k, solver):
    rng = np.random.default_rng(2000 + 7 * n + k)
    A = rng.standard_normal((n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, k))
    X = np.zeros((n, k))
    getattr(fp, solver)(X, A, B)
    assert rel(X, np.linalg.solve(A, B)) < 1e-11, (n, k, solver)


@pytest.mark.parametrize("n,k,solver", [(3001, 1, "solve_ludiv"), (3001, 2, "solve_ludiv"), (2000, 40, "solve_ludiv"),
                                        (3001, 1, "solve_lu32_ir"), (2000, 3, "solve_lu32_ir")])
def test_getrs_sync_free_sweeps(fp, n, k, solver):
    # one launch per triangular sweep: 24 blocks in a chain of ready flags; 16 blocks x 40
    # right-hand sides = 640 workgroups, more than are resident at once (tickets order them)
    rng = np.random.default_rng(3000 + n + k)
    A = rng.standard_normal((n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, k))
    X = np.zeros((n, k))
    getattr(fp, solver)(X, A, B)
    assert rel(X, np.linalg.solve(A, B)) < 1e-11, (n, k, solver)


@pytest.mark.parametrize("block", [1, 64, 128, 129, 250, 256, 257, 300, 500, 512, 513, 777])
def test_block_jacobi_paths(fp, block):
    # blocks <= 128: Gauss-Jordan inverses from A; above: blocked Gauss-Jordan over
    # ceil(block / 128) chunks (777: one block covering the whole matrix)
    rng = np.random.default_rng(3000 + block)
    n = 777
    A = rng.standard_normal((n, n)) * 0.3 / np.sqrt(n) + 3.0 * np.eye(n)
    for s in range(0, n, 50):                       # diagonal blocks for the preconditioner to capture
        e = min(n, s + 50)
        A[s:e, s:e] += rng.standard_normal((e - s, e - s)) * 0.2
    b = rng.standard_normal(n)
    y = np.zeros(n)
    fp.solve_gmres_ir(y, A, b, precond_block=block, tol=1e-12)
    assert np.linalg.norm(A @ y - b) / np.linalg.norm(b) < 1e-11
    assert rel(y, np.linalg.solve(A, b)) < 1e-9


@pytest.mark.parametrize("n,block", [(1001, 128), (1001, 250), (1001, 300), (1001, 500), (1130, 500), (1100, 512),
                                     (777, 257), (1600, 750), (2100, 1000), (1500, 1024), (900, 777)])
def test_block_jacobi_exact_on_block_diagonal(fp, n, block):
    """On a matrix that is block diagonal with the preconditioner's own blocks,
    M^-1 A = I up to the FP32 inverse: every inner solve ends after one
    iteration.  A wrong chunk, panel or offset of the blocked inverse (blocks
    above 256, identity-padded ragged last blocks) would show as many more."""
    rng = np.random.default_rng(4000 + n + block)
    A = np.zeros((n, n))
    for s in range(0, n, block):
        e = min(n, s + block)
        A[s:e, s:e] = rng.standard_normal((e - s, e - s)) / np.sqrt(e - s) + 2.0 * np.eye(e - s)
    b = rng.standard_normal(n)
    y = np.zeros(n)
    fp.solve_gmres_ir(y, A, b, precond_block=block, tol=1e-12)
    st = dict(fp.solvers.last_stats)
    assert rel(y, np.linalg.solve(A, b)) < 1e-10, st
    assert st["iterations"] <= st["refinements"], st


def test_block_jacobi_block_limit(fp):
    A = np.eye(1100) * 2.0
    with pytest.raises(Exception, match="1024"):
        fp.solve_gmres_ir(np.zeros(1100), A, np.ones(1100), precond_block=1025)


def test_buffer_pool_reuse_and_trim(fp):
    """Large work buffers (16 MiB and up) go back to the context after a call and
    are handed to the next call of the same size: repeated solves must give the
    same bits, before and after the context frees them (trim)."""
    rng = np.random.default_rng(5000)
    n = 1600                                         # FP64 matrix copies of 20 MB: pooled sizes
    A = rng.standard_normal((n, n)) / np.sqrt(n) + 3.0 * np.eye(n)
    b = rng.standard_normal(n)
    y1 = fp.solve_gmres_ir(None, A, b, precond_block=400, coarse_block=400, tol=1e-12)
    y2 = fp.solve_gmres_ir(None, A, b, precond_block=400, coarse_block=400, tol=1e-12)
    assert np.array_equal(y1, y2)
    fp.context().trim()
    y3 = fp.solve_gmres_ir(None, A, b, precond_block=400, coarse_block=400, tol=1e-12)
    assert np.array_equal(y1, y3)
    assert rel(y1, np.linalg.solve(A, b)) < 1e-10
    x1 = fp.solve_ludiv(None, A, b)
    x2 = fp.solve_ludiv(None, A, b)
    assert np.array_equal(x1, x2)


@pytest.mark.parametrize("coarse,n", [(8, 1001), (5, 1001), (3, 1001), (1, 1300), (2, 2601)])
def test_coarse_space_above_128_vectors(fp, coarse, n):
    """Coarse spaces of 126, 201, 334, 1,300 and 1,301 chunk vectors: above 128 the
    Galerkin matrix is inverted by the blocked Gauss-Jordan of the block-Jacobi
    path (one block of nc rows, above 1,024 too); the two-level solve must still
    reach its FP64 target and take no more iterations than block-Jacobi alone."""
    rng = np.random.default_rng(6000 + coarse)
    A = rng.standard_normal((n, n)) * 0.3 / np.sqrt(n) + 3.0 * np.eye(n) + 0.5 / n
    b = rng.standard_normal(n)
    y = fp.solve_gmres_ir(None, A, b, precond_block=100, coarse_block=coarse, tol=1e-12)
    it2 = fp.solvers.last_stats["iterations"]
    fp.solve_gmres_ir(None, A, b, precond_block=100, tol=1e-12)
    it1 = fp.solvers.last_stats["iterations"]
    assert rel(y, np.linalg.solve(A, b)) < 1e-10
    assert it2 <= it1, (it2, it1)


@pytest.mark.parametrize("n,coarse", [(2600, 2), (2601, 5), (3000, 1)])
def test_coarse_inverse_lu_matches_gauss_jordan(fp, n, coarse, monkeypatch):
    """Coarse spaces of 1,300, 521 and 3,000 vectors inverted through the randomized
    no-pivot FP32 LU (FLOWPANEL_COARSE_INV=lu: U^T E V, explicit inverses of its
    factors, V (.) U^T back) against the blocked Gauss-Jordan with pivoting, on a
    nonsymmetric matrix with strong coupling between chunks: both inverses are FP32
    approximations of the same E^-1, so the two-level GMRES takes the same number of
    iterations (within one) and reaches the same FP64 solution.  A transposed or
    mis-indexed inverse (a wrong butterfly side) would not converge in as few.
    (The diagonal shift keeps E's eigenvalues away from zero: with 2 I instead of
    8 I the chunk coupling puts them in a disk around 4 of radius ~8, and E is too
    close to singular for any FP32 factorisation.)"""
    rng = np.random.default_rng(8000 + n)
    A = rng.standard_normal((n, n)) / np.sqrt(n) + 8.0 * np.eye(n)
    A += np.kron(rng.uniform(0.0, 0.4, (n // coarse + 1, n // coarse + 1)), np.ones((coarse, coarse)))[:n, :n] / coarse
    b = rng.standard_normal(n)
    monkeypatch.setenv("FLOWPANEL_COARSE_INV", "lu")
    y_lu = fp.solve_gmres_ir(None, A, b, precond_block=100, coarse_block=coarse, tol=1e-12)
    it_lu = fp.solvers.last_stats["iterations"]
    monkeypatch.delenv("FLOWPANEL_COARSE_INV")
    y_gj = fp.solve_gmres_ir(None, A, b, precond_block=100, coarse_block=coarse, tol=1e-12)
    it_gj = fp.solvers.last_stats["iterations"]
    fp.solve_gmres_ir(None, A, b, precond_block=100, tol=1e-12)
    it_bj = fp.solvers.last_stats["iterations"]
    x = np.linalg.solve(A, b)
    assert rel(y_lu, x) < 1e-10 and rel(y_gj, x) < 1e-10
    assert abs(it_lu - it_gj) <= 1, (it_lu, it_gj)
    assert it_lu < it_bj, (it_lu, it_bj)


@pytest.mark.parametrize("n,coarse", [(1001, 7), (1200, 50), (999, 63), (640, 1)])
def test_coarse_chunk_sums_coalesced_match_lane_sums(fp, n, coarse, monkeypatch):
    """The coarse space's chunk sums (A Z and Z^T A Z) read through LDS segments
    against the one-lane-per-chunk kernels (FLOWPANEL_COARSE_LANE=1): the same
    iterations and the same solution on a nonsymmetric matrix with strong
    coupling between chunks, rows of odd and even length (16-byte and 8-byte
    aligned rows of A Z), a last chunk cut short and chunks of one entry."""
    rng = np.random.default_rng(7000 + n)
    A = rng.standard_normal((n, n)) / np.sqrt(n) + 2.0 * np.eye(n)
    A += np.kron(rng.uniform(0.0, 0.4, (n // coarse + 1, n // coarse + 1)), np.ones((coarse, coarse)))[:n, :n] / coarse
    b = rng.standard_normal(n)
    y1 = fp.solve_gmres_ir(None, A, b, precond_block=100, coarse_block=coarse, tol=1e-12)
    it1 = fp.solvers.last_stats["iterations"]
    monkeypatch.setenv("FLOWPANEL_COARSE_LANE", "1")
    y2 = fp.solve_gmres_ir(None, A, b, precond_block=100, coarse_block=coarse, tol=1e-12)
    it2 = fp.solvers.last_stats["iterations"]
    x = np.linalg.solve(A, b)
    print("n %d chunk %d: iterations %d / %d, %.1e / %.1e" % (n, coarse, it1, it2, rel(y1, x), rel(y2, x)))
    assert rel(y1, x) < 1e-10 and rel(y2, x) < 1e-10
    assert abs(it1 - it2) <= 1


@pytest.mark.parametrize("n,block", [(1001, 48), (1001, 100), (1001, 300), (2050, 1000), (1001, 250), (777, 1024)])
def test_block_jacobi_apply_vector_loads(fp, n, block, monkeypatch):
    """The block-Jacobi product with 16-byte loads (block % 4 == 0) against the
    one-float-per-lane product (FLOWPANEL_BJ_APPLY_SCALAR=1): both solves reach
    the FP64 target with the same iterations, with a short last block and a
    block of 1024; block 250 takes the one-float path either way."""
    rng = np.random.default_rng(7100 + block)
    A = rng.standard_normal((n, n)) * 0.4 / np.sqrt(n) + 2.0 * np.eye(n)
    b = rng.standard_normal(n)
    y1 = fp.solve_gmres_ir(None, A, b, precond_block=block, tol=1e-12)
    it1 = fp.solvers.last_stats["iterations"]
    monkeypatch.setenv("FLOWPANEL_BJ_APPLY_SCALAR", "1")
    y2 = fp.solve_gmres_ir(None, A, b, precond_block=block, tol=1e-12)
    it2 = fp.solvers.last_stats["iterations"]
    x = np.linalg.solve(A, b)
    assert rel(y1, x) < 1e-10 and rel(y2, x) < 1e-10
    assert abs(it1 - it2) <= 1, (it1, it2)


def _blocks_matrix(n, block, rng, permute):
    """Weakly coupled diagonal blocks: well-conditioned random blocks, or (permute) blocks
    whose rows are permuted so that an elimination without row exchanges would meet tiny
    pivots (blocks of at most one chunk: the blocked Gauss-Jordan pivots within chunks)."""
    A = rng.standard_normal((n, n)) * 0.01
    for b0 in range(0, n, block):
        m = min(block, n - b0)
        B = rng.standard_normal((m, m)) * 0.5 / np.sqrt(m) + 2.0 * np.eye(m)
        if permute:
            B[np.arange(m), np.arange(m)] *= 1e-3
            B = B[rng.permutation(m)]
        A[b0:b0 + m, b0:b0 + m] = B
    return A


def _check_block_inverses(A, X, block, tol):
    n = A.shape[0]
    b = min(block, n)
    worst = 0.0
    for k in range(X.shape[0]):
        b0, m = k * b, min(b, n - k * b)
        ref = np.linalg.inv(A[b0:b0 + m, b0:b0 + m])
        worst = max(worst, np.abs(X[k, :m, :m] - ref).max() / np.abs(ref).max())
        if m < b:                                        # padding: diag(inverse, I)
            pad = np.eye(b, dtype=np.float32)
            pad[:m, :m] = X[k, :m, :m]
            assert np.array_equal(X[k], pad)
    assert worst < tol, worst
    return worst


@pytest.mark.parametrize("n,block,permute", [(130, 1, True), (130, 7, True), (130, 64, True), (200, 65, False),
                                             (1001, 100, False), (1001, 250, False), (2050, 1000, False),
                                             (1030, 1024, False)])
def test_block_inverses_against_numpy(fp, n, block, permute, monkeypatch):
    """The block-Jacobi inverses (one-wave Gauss-Jordan on chunks of at most 64 rows,
    blocked above 64) against numpy's inverse of each diagonal block, FP32 accuracy:
    blocks that need row exchanges, blocks of many chunks, a short last block padded
    with the identity; the 256-thread register grid (FLOWPANEL_BJ_GJ=grid, chunks of
    128 rows) likewise."""
    rng = np.random.default_rng(7300 + n + block)
    A = _blocks_matrix(n, block, rng, permute)
    w1 = _check_block_inverses(A, fp.solvers.block_inverses(A, block), block, 1e-4)
    monkeypatch.setenv("FLOWPANEL_BJ_GJ", "grid")
    X = fp.solvers.block_inverses(A, block)
    b = min(block, n)
    w2 = max(np.abs(X[k, :min(b, n - k * b), :min(b, n - k * b)] -
                    np.linalg.inv(A[k * b:k * b + min(b, n - k * b), k * b:k * b + min(b, n - k * b)])).max() /
             np.abs(np.linalg.inv(A[k * b:k * b + min(b, n - k * b), k * b:k * b + min(b, n - k * b)])).max()
             for k in range(X.shape[0]))
    print("n %d block %d: worst block error %.1e (grid %.1e)" % (n, block, w1, w2))
    assert w2 < 1e-4


@pytest.mark.parametrize("block", [250, 1000])
def test_block_inverses_of_an_aic(fp, geo, block):
    """Diagonal blocks of a wing's AIC (1,440 panels with its wake) against numpy."""
    g, sh = geo.loft_wing(n_chord=48, n_span=30, span=6.0, sweep_deg=20.0)
    body = fp.RigidWakeBody(g, sh)
    V = fp.direction(5.0)
    fp.solvers.set_freestream(body, np.tile(V[:, None], (1, body.ncells)), np.tile(V[:, None], (1, body.nsheddings)),
                              np.tile(V[:, None], (1, body.nsheddings)))
    A = fp.assemble(body, "f64")
    w = _check_block_inverses(A, fp.solvers.block_inverses(A, block), block, 1e-4)
    print("AIC block %d: worst block error %.1e" % (block, w))


def test_block_inverses_report_a_zero_pivot(fp):
    A = np.eye(300) * 3.0
    A[130:140, 130:140] = 0.0                            # a singular 10 x 10 piece of block 1
    with pytest.raises(fp._native.FPError) as e:
        fp.solvers.block_inverses(A, 100)
    assert "zero pivot in diagonal block 1" in str(e.value)


# ---- FP32 LU of the randomized matrix without pivoting (the LU32 refinement path) ----
# The no-pivot LU works in 64-column steps on U^T A V, n padded to a multiple of 4:
# sizes straddle the pad and the 64-row blocks.

@pytest.mark.parametrize("n", [1, 3, 4, 5, 63, 64, 65, 127, 130, 700, 1031])
def test_randomized_nopivot_lu32_refinement(fp, n):
    rng = np.random.default_rng(4000 + n)
    A = rng.standard_normal((n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    fp.solve_lu32_ir(x, A, b, tol=1e-12)
    assert rel(x, np.linalg.solve(A, b)) < 1e-11, n
    # an accurate FP32 factorisation: a handful of refinement steps
    assert fp.last_stats["refinements"] <= 4, fp.last_stats


def test_randomized_nopivot_lu32_needs_no_pivots(fp):
    # a zero diagonal (plain LU without pivoting breaks down at the first step);
    # the butterflies make every leading block nonsingular
    rng = np.random.default_rng(7)
    n = 300
    P = np.eye(n)[rng.permutation(n)]
    A = 5.0 * P + 0.1 * rng.standard_normal((n, n))
    np.fill_diagonal(A, 0.0)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    fp.solve_lu32_ir(x, A, b, tol=1e-12)
    assert rel(x, np.linalg.solve(A, b)) < 1e-10


def test_lu32_pivoted_and_randomized_agree(fp, geo, monkeypatch):
    g, sh = geo.loft_wing(n_chord=30, n_span=20, span=6.0, sweep_deg=25.0, taper=0.6)
    body = fp.RigidWakeBody(g, sh)
    V = fp.direction(4.2)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings))
    st_r = fp.solve(body, U, D, D, solver=fp.solve_lu32_ir, solver_optargs={"tol": 1e-12})
    g_r = body.strength.copy()
    monkeypatch.setenv("FLOWPANEL_LU32_PIVOTED", "1")
    st_p = fp.solve(body, U, D, D, solver=fp.solve_lu32_ir, solver_optargs={"tol": 1e-12})
    assert rel(body.strength, g_r) < 1e-10, (st_r, st_p)


# With four or more right-hand sides the refinement corrections of the randomized
# path apply explicit inverses of the no-pivot factors (-L^-1, -U^-1 built level by
# level on MFMA): two GEMVs per correction (FLOWPANEL_LU32_INVERSE=1 forces them,
# =0 forces the blocked triangular solves, the reference point here).  Sizes
# straddle the 64-row blocks, the pairwise merges (ragged last pairs) and the pad
# to a multiple of 4.
@pytest.mark.parametrize("n", [5, 64, 130, 300, 1031, 2500])
def test_lu32_explicit_inverses_match_triangular_solves(fp, n, monkeypatch):
    rng = np.random.default_rng(9100 + n)
    A = rng.standard_normal((n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, 3))
    X = np.zeros((n, 3))
    monkeypatch.setenv("FLOWPANEL_LU32_INVERSE", "1")
    fp.solve_lu32_ir(X, A, B, tol=1e-12)
    st_inv = dict(fp.last_stats)
    monkeypatch.setenv("FLOWPANEL_LU32_INVERSE", "0")
    X2 = np.zeros((n, 3))
    fp.solve_lu32_ir(X2, A, B, tol=1e-12)
    st_trs = dict(fp.last_stats)
    ref = np.linalg.solve(A, B)
    assert rel(X, ref) < 1e-11 and rel(X2, ref) < 1e-11, (st_inv, st_trs)
    # the inverses are as good a preconditioner as the substitutions
    assert st_inv["refinements"] <= st_trs["refinements"] + 1, (st_inv, st_trs)


def test_lu32_explicit_inverses_on_a_wing(fp, geo, orc, monkeypatch):
    monkeypatch.setenv("FLOWPANEL_LU32_INVERSE", "1")
    g, sh = geo.loft_wing(n_chord=40, n_span=30, span=6.0, sweep_deg=25.0, taper=0.6)
    body = fp.RigidWakeBody(g, sh)
    V = fp.direction(4.2)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings))
    st = fp.solve(body, U, D, D, solver=fp.solve_lu32_ir, solver_optargs={"tol": 1e-12})
    import types
    ob = types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=g.cell_idx, ncells=body.ncells,
                               element="vortexring", shedding=body.shedding, nsheddings=body.nsheddings,
                               kerneloffset=body.kerneloffset, kernelcutoff=body.kernelcutoff, CPoffset=body.CPoffset,
                               charlen_kind=0)
    Go, normals, _ = orc.assemble([ob], [D], [D])
    xo = orc.solve_lu(Go, orc.rhs(U, normals))
    assert rel(body.strength, xo) < 1e-10, st
