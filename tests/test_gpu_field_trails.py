"""The edge-trail velocity kernel on the device.

Uind!/calcfield_U of a vortex-ring body (src/FLOWPanel_liftingbody.jl:576-661)
are summed over the mesh's edges walked as trails (gamma_e from the current
strengths, wake horseshoes as before).  Checked against the oracle (1e-10 of
max|U|) and against the panel-wise kernel (FLOWPANEL_FIELD=panel, 1e-12) on a
quad wing with its wake, a triangulated wing, a reduced blended-wing-body
loft, a watertight sphere and a mesh of hexagons (eight-vertex storage), with
few targets (the panel range split over workgroups) and many; and through the
matrix-free FP64 residual (n . U at the control points) of the mixed solver
against the FP64 LU solution.
"""
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _obody(b):
    return types.SimpleNamespace(nodes=b.grid.nodes, cell_ptr=b.grid.cell_ptr, cell_idx=b.grid.cell_idx,
                                 ncells=b.ncells, element="vortexring", shedding=b.shedding,
                                 nsheddings=b.nsheddings, kerneloffset=b.kerneloffset,
                                 kernelcutoff=b.kernelcutoff, CPoffset=b.CPoffset, charlen_kind=0)


def _hexagons(fp, geo):
    g, sh = geo.loft_wing(n_chord=12, n_span=10)
    q = g.cell_idx.reshape(-1, 4)
    n1, n2 = g.dims
    cells = []
    for j in range(n2):
        for i in range(0, n1 - 1, 2):
            a, b = q[i + n1 * j], q[i + 1 + n1 * j]
            cells.append([a[0], b[0], b[1], b[2], b[3], a[3]] if a[1] == b[0] and a[2] == b[3] else a)
    return fp.Grid(g.nodes, cells)


def _body(fp, geo, case):
    if case == "quads_wake":
        return fp.RigidWakeBody(*geo.loft_wing(n_chord=16, n_span=12))
    if case == "triangles_wake":
        return fp.RigidWakeBody(*geo.triangulate(*geo.loft_wing(n_chord=16, n_span=10)))
    if case == "bwb":
        return fp.RigidWakeBody(*geo.bwb(n_chord=50, n_span=24))
    if case == "sphere":
        return fp.RigidWakeBody(geo.sphere(n_meridional=8, n_azimuthal=12), None, formulation=2)
    return fp.RigidWakeBody(_hexagons(fp, geo), None)


@pytest.mark.parametrize("case", ["quads_wake", "triangles_wake", "bwb", "sphere", "hexagons"])
def test_trail_velocity_matches_oracle_and_panelwise(fp, geo, orc, case, monkeypatch):
    body = _body(fp, geo, case)
    V = fp.direction(4.0)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings)) if body.nsheddings else None
    print("case %s: %d panels" % (case, body.ncells), flush=True)
    fp.solve(body, U, D, D) if body.nsheddings else fp.solve(body, U)
    rng = np.random.default_rng(7)
    for nt in (17, 3000):
        X = np.vstack([rng.uniform(-1, 6, nt), rng.uniform(-5, 5, nt), rng.uniform(-1.5, 1.5, nt)])
        Ut = fp.Uind(body, X, np.zeros((3, nt)))
        Uo = orc.uind(_obody(body), body.strength, X, D, D)
        monkeypatch.setenv("FLOWPANEL_FIELD", "panel")
        Up = fp.Uind(body, X, np.zeros((3, nt)))
        monkeypatch.delenv("FLOWPANEL_FIELD")
        print("  nt %d: vs oracle %.1e, vs panel-wise %.1e" % (nt, rel(Ut, Uo), rel(Ut, Up)), flush=True)
        assert rel(Ut, Uo) < 1e-10
        assert rel(Ut, Up) < 1e-12


def test_matrix_free_residual_through_trails(fp, geo, orc):
    g, sh = geo.loft_wing(n_chord=40, n_span=30)
    body = fp.RigidWakeBody(g, sh)
    V = fp.direction(5.0)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings))
    st = fp.solve(body, U, D, D, solver=fp.solve_gmres_ir,
                  solver_optargs={"precond_block": 160, "residual_mode": 0, "tol": 1e-11})
    x = body.strength.copy()
    Go, normals, _ = orc.assemble([_obody(body)], [D], [D])
    xo = orc.solve_lu(Go, orc.rhs(U, normals))
    print("matrix-free residual: %s, rel %.1e" % (st, rel(x, xo)), flush=True)
    assert rel(x, xo) < 1e-9
