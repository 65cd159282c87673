"""Batched freestreams (fixed-wake convention) and multi-RHS solver protocol on the GPU.

k freestreams share one AIC when the wake directions are fixed (here along +x
for every case): one assembly, one factorisation or preconditioner, k right-hand
sides.  Each column is checked against the oracle's LAPACK solve of the same
system (1e-10 relative, the FP64-path tolerance).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _obody(b):
    import types
    return types.SimpleNamespace(nodes=b.grid.nodes, cell_ptr=b.grid.cell_ptr, cell_idx=b.grid.cell_idx,
                                 ncells=b.ncells, element="vortexring", shedding=b.shedding,
                                 nsheddings=b.nsheddings, kerneloffset=b.kerneloffset,
                                 kernelcutoff=b.kernelcutoff, CPoffset=b.CPoffset, charlen_kind=0)


def _freestreams(fp, body, k, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    alpha = rng.uniform(-5.0, 10.0, k)
    beta = rng.uniform(-4.0, 4.0, k)
    V = np.stack([fp.direction(a, b) for a, b in zip(alpha, beta)], axis=1)   # 3 x k
    return np.repeat(V[:, None, :], body.ncells, axis=1)                     # 3 x N x k


@pytest.mark.parametrize("solver", ["solve_ludiv", "solve_lu32_ir", "solve_gmres_ir"])
def test_freestream_batch_matches_oracle(fp, geo, orc, solver):
    g, sh = geo.loft_wing(n_chord=24, n_span=16, span=6.0, sweep_deg=30.0, taper=0.5)
    body = fp.RigidWakeBody(g, sh)
    k = 12
    U = _freestreams(fp, body, k)
    D = np.tile(np.array([[1.0], [0.0], [0.0]]), (1, body.nsheddings))
    opt = {"precond_block": 48} if solver == "solve_gmres_ir" else {}
    G, st = fp.solve_freestreams(body, U, D, D, solver=getattr(fp, solver), solver_optargs=opt)
    Go, normals, _ = orc.assemble([_obody(body)], [D], [D])
    import scipy.linalg as sla
    lu = sla.lu_factor(Go)
    for c in range(k):
        ref = sla.lu_solve(lu, orc.rhs(U[:, :, c], normals))
        assert rel(G[:, c], ref) < 1e-10, (c, rel(G[:, c], ref))
    # one freestream through the single solve agrees with its column
    fp.solve(body, U[:, :, 5], D, D, solver=getattr(fp, solver), solver_optargs=opt)
    assert rel(body.strength, G[:, 5]) < 1e-10


def test_solver_protocol_multi_rhs_factor_once(fp):
    rng = np.random.default_rng(17)
    n, k = 500, 20
    A = rng.standard_normal((n, n)) + 3 * n ** 0.5 * np.eye(n)
    B = rng.standard_normal((n, k))
    for solver in ("solve_ludiv", "solve_lu32_ir"):
        X = np.zeros((n, k))
        getattr(fp, solver)(X, A, B)
        assert rel(X, np.linalg.solve(A, B)) < 1e-11, solver
        if solver == "solve_lu32_ir":
            assert fp.last_stats["refinements"] >= 2   # FP32 factors: refinement did the work


def _wings(fp, geo, sizes, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for nc, ns in sizes:
        ar = rng.uniform(6, 12)
        g, sh = geo.loft_wing(airfoil="0012", n_chord=nc, n_span=ns, span=ar * 0.75, root_chord=1.0, taper=0.75,
                              sweep_deg=rng.uniform(0, 35))
        out.append(fp.RigidWakeBody(g, sh))
    return out


@pytest.mark.parametrize("solver,resid", [("solve_lu32_ir", 0), ("solve_lu32_ir", 1), ("solve_ludiv", 1)])
def test_body_batch_matches_single_solves(fp, geo, solver, resid):
    bodies = _wings(fp, geo, [(20, 10), (34, 18), (12, 6), (26, 30), (40, 8)])
    refs = []
    for k, b in enumerate(bodies):
        V = fp.direction(2.0 + k, 0.5 * k)
        U = np.tile(V[:, None], (1, b.ncells))
        D = np.tile(V[:, None], (1, b.nsheddings))
        fp.solve(b, U, D, D)                 # FP64 LU reference (also stores the freestream)
        refs.append(b.strength.copy())
    batch = fp.BodyBatch(bodies, solver=getattr(fp, solver), solver_optargs={"residual_mode": resid, "tol": 1e-12})
    assert len(batch.autotune_table()) == 9
    from flowpanel_amd import _native as nat
    for rep in range(3):
        for b in bodies:                     # eager work between replays: clear the device strengths
            z = np.zeros(b.ncells)
            b.ctx.check(nat.lib().fp_body_set_strength(b.ctx.handle, b.handle, nat.dptr(z)))
        st = batch.solve()
        assert st["nbodies"] == 5 and st["residual"] <= 1e-12
        for b, ref in zip(bodies, refs):
            assert rel(b.strength, ref) < 1e-10, (rep, rel(b.strength, ref))
        fp.calcfield_U(bodies[0], bodies[0])  # an unrelated eager kernel between replays
    # a new freestream on one body re-captures the graph and is honoured
    b = bodies[2]
    V = fp.direction(7.0)
    U = np.tile(V[:, None], (1, b.ncells))
    D = np.tile(V[:, None], (1, b.nsheddings))
    fp.solve(b, U, D, D)
    ref = b.strength.copy()
    st = batch.solve()
    assert st["recaptured"] == 1
    assert rel(b.strength, ref) < 1e-10


@pytest.mark.parametrize("lu", ["bodies", "batched"])
def test_body_batch_factorisations_agree(fp, geo, lu, monkeypatch):
    """The batched FP32 no-pivot LU (one launch per step for every body: bodies of 1,152 to
    4,560 unknowns, so the steps end at different outer blocks and the last outer blocks are
    ragged) against each body's own LU launches (FLOWPANEL_BATCH_LU=bodies), through the
    whole mixed solve with three graph replays: both reach the FP64 LU's strengths to 1e-10."""
    if lu == "bodies":
        monkeypatch.setenv("FLOWPANEL_BATCH_LU", "bodies")
    bodies = _wings(fp, geo, [(48, 24), (60, 76), (36, 32), (20, 60)])
    refs = []
    for k, b in enumerate(bodies):
        V = fp.direction(3.0 + k)
        U = np.tile(V[:, None], (1, b.ncells))
        D = np.tile(V[:, None], (1, b.nsheddings))
        fp.solve(b, U, D, D)
        refs.append(b.strength.copy())
    batch = fp.BodyBatch(bodies, solver=fp.solve_lu32_ir, solver_optargs={"residual_mode": 0, "tol": 1e-12})
    for rep in range(3):
        st = batch.solve()
        assert st["residual"] <= 1e-12, st
        for b, ref in zip(bodies, refs):
            assert rel(b.strength, ref) < 1e-10, (lu, rep, rel(b.strength, ref))
