"""Parity of the HIP path against the CPU oracle (FP64 restatement of the reference).

Tolerances: FP64 paths 1e-10 relative (the agreement asked of the FP64 path,
measured against max |.| of the quantity); FP32 assembly 2e-5 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.fixture(scope="module")
def sphere(fp, geo):
    g = geo.sphere(n_meridional=12, n_azimuthal=24)
    return fp.NonLiftingBody(g)


@pytest.fixture(scope="module")
def wing(fp, geo):
    g, sh = geo.loft_wing(n_chord=24, n_span=16, span=6.0, sweep_deg=20.0, taper=0.6,
                          twist_root_deg=1.0, twist_tip_deg=-2.0)
    b = fp.RigidWakeBody(g, sh)
    return b


def _freestream(body, alpha=4.2, beta=0.0, geo=None):
    from flowpanel_amd import direction
    V = direction(alpha, beta)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, max(body.nsheddings, 0)))
    return U, D


def test_geometry_matches_oracle(fp, orc, wing, sphere):
    for b in (wing, sphere):
        n, c, a = fp.calc_normals(b), fp.calc_controlpoints(b), fp.calc_areas(b)
        on, oc, oa = orc.geometry(b.grid.nodes, b.grid.cell_ptr, b.grid.cell_idx, b.CPoffset, 0)
        assert rel(n, on) < 1e-14
        assert rel(c, oc) < 1e-14
        assert rel(a, oa) < 1e-14


def _obody(b):
    import types
    return types.SimpleNamespace(nodes=b.grid.nodes, cell_ptr=b.grid.cell_ptr, cell_idx=b.grid.cell_idx,
                                 ncells=b.ncells, element="vortexring" if b.nsheddings or b.element.code else "source",
                                 shedding=b.shedding, nsheddings=b.nsheddings, kerneloffset=b.kerneloffset,
                                 kernelcutoff=b.kernelcutoff, CPoffset=b.CPoffset, charlen_kind=0)


def test_assembly_source_f64(fp, orc, sphere):
    U, _ = _freestream(sphere)
    fp.solvers.set_freestream(sphere, U)
    G = fp.assemble(sphere, "f64")
    Go, _, _ = orc.assemble(_obody(sphere))
    assert rel(G, Go) < 1e-12


def test_assembly_vortexring_wake_f64(fp, orc, wing):
    U, D = _freestream(wing)
    fp.solvers.set_freestream(wing, U, D, D)
    G = fp.assemble(wing, "f64")
    Go, _, _ = orc.assemble([_obody(wing)], [D], [D])
    assert rel(G, Go) < 1e-12


def test_assembly_vortexring_f32(fp, orc, wing):
    U, D = _freestream(wing)
    fp.solvers.set_freestream(wing, U, D, D)
    G = fp.assemble(wing, "f32")
    Go, _, _ = orc.assemble([_obody(wing)], [D], [D])
    assert rel(G, Go) < 2e-5


@pytest.mark.parametrize("n", [37, 300, 700])
def test_getrf_matches_lapack(fp, n):
    import scipy.linalg as sla
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    LU, ipiv = fp.getrf(A)
    # reconstruct P A = L U with the returned (LAPACK-convention) pivots
    L = np.tril(LU, -1) + np.eye(n)
    Uu = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(ipiv - 1):
        PA[[i, p]] = PA[[p, i]]
    assert np.abs(PA - L @ Uu).max() / np.abs(A).max() < 1e-12
    lu_ref = sla.lu_factor(A)
    x = sla.lu_solve((LU, ipiv - 1), np.ones(n))
    xr = sla.lu_solve(lu_ref, np.ones(n))
    assert rel(x, xr) < 1e-10


@pytest.mark.parametrize("n", [300, 700])
def test_getrf_f32(fp, n):
    rng = np.random.default_rng(n + 1)
    A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    LU, ipiv = fp.getrf(A, "f32")
    L = np.tril(LU, -1) + np.eye(n)
    Uu = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(ipiv - 1):
        PA[[i, p]] = PA[[p, i]]
    assert np.abs(PA - L @ Uu).max() / np.abs(A).max() < 2e-5


@pytest.mark.parametrize("solver", ["solve_ludiv", "solve_lu32_ir", "solve_gmres_ir", "solve_gmres"])
def test_solver_protocol_matrix(fp, solver):
    rng = np.random.default_rng(7)
    n = 400
    A = rng.standard_normal((n, n)) + 3 * n ** 0.5 * np.eye(n)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    kw = {"precond_block": 64} if solver == "solve_gmres" else {}
    getattr(fp, solver)(x, A, b, **kw)
    xr = np.linalg.solve(A, b)
    tol = 1e-7 if solver == "solve_gmres" else 1e-11
    assert rel(x, xr) < tol


@pytest.mark.parametrize("solver", ["solve_ludiv", "solve_lu32_ir", "solve_gmres_ir"])
def test_solve_wing_matches_oracle(fp, orc, wing, solver):
    U, D = _freestream(wing)
    opt = {"precond_block": 48} if solver == "solve_gmres_ir" else {}
    stats = fp.solve(wing, U, D, D, solver=getattr(fp, solver), solver_optargs=opt)
    Go, normals, cps = orc.assemble([_obody(wing)], [D], [D])
    gam = orc.solve_lu(Go, orc.rhs(U, normals))
    assert rel(wing.strength, gam) < TOL64, stats
    assert stats["residual"] < 1e-11


@pytest.mark.parametrize("solver", ["solve_gmres_ir", "solve_gmres"])
def test_two_level_preconditioner_matches_oracle(fp, orc, wing, solver):
    """Block-Jacobi + Galerkin coarse correction on one-strip chunks (24 panels)."""
    U, D = _freestream(wing)
    opt = {"precond_block": 48, "coarse_block": 24, "tol": 1e-12}
    stats = fp.solve(wing, U, D, D, solver=getattr(fp, solver), solver_optargs=opt)
    Go, normals, cps = orc.assemble([_obody(wing)], [D], [D])
    gam = orc.solve_lu(Go, orc.rhs(U, normals))
    assert rel(wing.strength, gam) < TOL64, stats
    plain = fp.solve(wing, U, D, D, solver=getattr(fp, solver), solver_optargs={"precond_block": 48, "tol": 1e-12})
    assert stats["iterations"] <= plain["iterations"]


def test_solve_sphere_and_cp(fp, orc, sphere):
    U, _ = _freestream(sphere, alpha=0.0)
    fp.solve(sphere, U)
    Go, normals, cps = orc.assemble(_obody(sphere))
    sig = orc.solve_lu(Go, orc.rhs(U, normals))
    assert rel(sphere.strength, sig) < TOL64
    Us = fp.calcfield_U(sphere, sphere)
    Uo = U + orc.uind(_obody(sphere), sig, cps)
    assert rel(Us, Uo) < TOL64
    Cp = fp.calcfield_Cp(sphere, 1.0)
    assert rel(Cp, orc.cp(Uo, 1.0)) < TOL64


def test_uind_phi_offbody(fp, orc, wing):
    U, D = _freestream(wing)
    fp.solve(wing, U, D, D)
    rng = np.random.default_rng(3)
    T = np.asfortranarray(rng.uniform([-1, -4, -1], [3, 4, 1], size=(500, 3)).T)
    out = np.zeros((3, 500), order="F")
    fp.Uind(wing, T, out)
    ref = orc.uind(_obody(wing), wing.strength, T, D, D)
    assert rel(out, ref) < TOL64
    ph = np.zeros(500)
    fp.phi(wing, T, ph)
    pr = orc.phi(_obody(wing), wing.strength, T, D, D)
    assert rel(ph, pr) < 1e-9


def test_cp_kutta_modes(fp, orc, geo):
    g, sh = geo.loft_wing(n_chord=12, n_span=4, span=3.0)
    gt, sht = geo.triangulate(g, sh)
    b = fp.RigidWakeBody(gt, sht)
    rng = np.random.default_rng(5)
    Us = rng.standard_normal((3, b.ncells))
    cp = fp.calcfield_Cp(b, 1.3, Us=Us)
    ref = orc.cp(Us, 1.3, sht, kutta=True)
    assert rel(cp, ref) < 1e-14


def test_multibody_matches_oracle(fp, orc, geo):
    g1, s1 = geo.loft_wing(n_chord=12, n_span=6, span=3.0)
    g2 = geo.sphere(n_meridional=6, n_azimuthal=10, center=(0.5, 0.0, -1.5), radius=0.3)
    w = fp.RigidWakeBody(g1, s1)
    s = fp.NonLiftingBody(g2)
    mb = fp.MultiBody([w, s], ["wing", "pod"])
    V = fp.direction(3.0)
    U = np.tile(V[:, None], (1, mb.ncells))
    D = np.tile(V[:, None], (1, mb.nsheddings))
    fp.solve(mb, U, D, D)
    Go, normals, cps = orc.assemble([_obody(w), _obody(s)], [D, None], [D, None])
    x = orc.solve_lu(Go, orc.rhs(U, normals))
    assert rel(mb.strength, x) < TOL64


def test_unsolved_body_raises(fp, geo):
    b = fp.NonLiftingBody(geo.sphere(n_meridional=4, n_azimuthal=6))
    with pytest.raises(fp.FPError) as e:
        fp.Uind(b, np.zeros((3, 2)), np.zeros((3, 2)))
    assert e.value.code == 5


def test_nonunit_wake_direction_raises(fp, wing):
    U, D = _freestream(wing)
    with pytest.raises(fp.FPError) as e:
        fp.solve(wing, U, 1.01 * D, D)
    assert e.value.code == 3
