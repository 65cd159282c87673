"""Surface post-processing on the GPU against the oracle: the cell-centred
Green-Gauss -1/2 grad(mu) (src/FLOWPanel_postprocess.jl:139-497), 1e-12."""
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("tri", [False, True])
@pytest.mark.parametrize("te", [False, True])
def test_ugradmu_cell_matches_oracle(fp, geo, orc, tri, te):
    g, sh = geo.loft_wing(airfoil="2412", n_chord=20, n_span=10, span=5.0, sweep_deg=20.0)
    if tri:
        g, sh = geo.triangulate(g, sh)
    body = fp.RigidWakeBody(g, sh, CPoffset=-1e-14 if tri else 1e-14)
    V = fp.direction(5.0)
    fp.solve(body, np.tile(V[:, None], (1, body.ncells)), np.tile(V[:, None], (1, body.nsheddings)),
             np.tile(V[:, None], (1, body.nsheddings)))
    out = fp.calcfield_Ugradmu_cell(body, te_correction=te)
    ob = types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=g.cell_idx, ncells=g.ncells,
                               CPoffset=body.CPoffset, charlen_kind=0)
    ref = orc.ugradmu_cell(ob, body.strength, sh, te=te)
    assert rel(out, ref) < 1e-12
    # the field adds to the induced velocity as in the reference workflow
    fp.calcfield_U(body, body)
    U = fp.addfields(body, "Ugradmu", "U")
    assert U.shape == (3, body.ncells)


def test_ugradmu_requires_solve(fp, geo):
    g, sh = geo.loft_wing(n_chord=8, n_span=4)
    body = fp.RigidWakeBody(g, sh)
    with pytest.raises(fp.FPError):
        fp.calcfield_Ugradmu_cell(body)


@pytest.mark.parametrize("tri", [False, True])
def test_ugradmu_node_and_combined_match_oracle(fp, geo, orc, tri):
    """Node-based (src/FLOWPanel_postprocess.jl:622-748) and combined (:756-1000)
    schemes against the oracle's restatement, 1e-12."""
    g, sh = geo.loft_wing(airfoil="2412", n_chord=20, n_span=10, span=5.0, sweep_deg=20.0)
    if tri:
        g, sh = geo.triangulate(g, sh)
    body = fp.RigidWakeBody(g, sh, CPoffset=-1e-14 if tri else 1e-14)
    V = fp.direction(5.0)
    fp.solve(body, np.tile(V[:, None], (1, body.ncells)), np.tile(V[:, None], (1, body.nsheddings)),
             np.tile(V[:, None], (1, body.nsheddings)))
    ob = types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=g.cell_idx, ncells=g.ncells,
                               CPoffset=body.CPoffset, charlen_kind=0)
    out = fp.calcfield_Ugradmu_node(body)
    ref, nodal = orc.ugradmu_node(ob, body.strength)
    assert rel(out, ref) < 1e-12
    assert rel(body.get_field("gamma_node"), nodal) < 1e-13
    for force in (True, False):
        for ang in (30.0, 10.0):
            outc = fp.calcfield_Ugradmu(body, force_cellTE=force, anglecrit=ang)
            refc = orc.ugradmu_combined(ob, body.strength, sh, force_cell_te=force, anglecrit=ang)
            assert rel(outc, refc) < 1e-12, (force, ang)


def _closed_loop(fp, g):
    """The same grid with coincident nodes merged: the section loop closes at the
    trailing edge (upper and lower TE cells share their TE nodes)."""
    X = np.asarray(g.nodes)
    _, first, inv = np.unique(np.round(X.T, 12), axis=0, return_index=True, return_inverse=True)
    rep = first[inv.ravel()]
    return fp.Grid(X, (np.asarray(g.cell_ptr), rep[np.asarray(g.cell_idx)]), dims=g.dims)


@pytest.mark.parametrize("tri", [False, True])
def test_ugradmu_sharp_te_matches_oracle(fp, geo, orc, tri):
    """sharpTE (src/FLOWPanel_postprocess.jl:692-743) of the node-based and combined schemes."""
    g, sh = geo.loft_wing(airfoil="2412", n_chord=20, n_span=10, span=5.0, sweep_deg=20.0)
    if tri:
        g, sh = geo.triangulate(g, sh)
    g = _closed_loop(fp, g)
    body = fp.RigidWakeBody(g, sh)
    V = fp.direction(5.0)
    fp.solve(body, np.tile(V[:, None], (1, body.ncells)), np.tile(V[:, None], (1, body.nsheddings)),
             np.tile(V[:, None], (1, body.nsheddings)))
    ob = types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=g.cell_idx, ncells=g.ncells,
                               CPoffset=body.CPoffset, charlen_kind=0)
    mode = 1 if tri else 2
    out = fp.calcfield_Ugradmu_node(body, sharpTE=True)
    ref, nodal = orc.ugradmu_node(ob, body.strength, sh, sharp_te=mode)
    assert rel(out, ref) < 1e-12
    assert rel(body.get_field("gamma_node"), nodal) < 1e-13
    plain, _ = orc.ugradmu_node(ob, body.strength)
    assert rel(plain, ref) > 1e-6            # the trailing-edge jump is no longer averaged away
    outc = fp.calcfield_Ugradmu(body, sharpTE=True, force_cellTE=False)
    refc = orc.ugradmu_combined(ob, body.strength, sh, force_cell_te=False, anglecrit=30.0, sharp_te=mode)
    assert rel(outc, refc) < 1e-12


@pytest.mark.parametrize("tri", [False, True])
def test_loads_match_oracle(fp, geo, orc, tri):
    """calcfield_F (normals flipped by the sign of CPoffset, Kutta override), Ftot, Mtot
    (src/FLOWPanel_postprocess.jl:1199-1307, 1543-1784) and the sectional force (:1411-1542) on the
    device against the oracle, 1e-12."""
    g, sh = geo.loft_wing(airfoil="2412", n_chord=20, n_span=10, span=5.0, sweep_deg=20.0)
    if tri:
        g, sh = geo.triangulate(g, sh)
    body = fp.RigidWakeBody(g, sh, CPoffset=-1e-14 if tri else 1e-14)
    V = fp.direction(5.0)
    fp.solve(body, np.tile(V[:, None], (1, body.ncells)), np.tile(V[:, None], (1, body.nsheddings)),
             np.tile(V[:, None], (1, body.nsheddings)))
    fp.calcfield_U(body, body)
    Cp = fp.calcfield_Cp(body, 1.0)
    normals, cps, areas = orc.geometry(g.nodes, g.cell_ptr, g.cell_idx, body.CPoffset, 0)
    s = -1.0 if body.CPoffset < 0 else 1.0
    Xac = np.array([0.25, 0.1, -0.05])
    kutta = tri                        # the oracle's Kutta pattern is the triangulated one (pi, pi+1, pj, pj-1)
    F, Ftot, Mtot = fp.loads(body, Cp, 1.0, 1.2, Xac=Xac, correct_kuttacondition=kutta)
    Fo = orc.forces(areas, s * normals, Cp, 1.0, 1.2, sh, kutta=kutta)
    assert rel(F, Fo) < 1e-12
    assert rel(Ftot, Fo.sum(axis=1)) < 1e-12
    assert rel(Mtot, orc.mtot(cps, Fo, Xac)) < 1e-12
    # moments at caller points (the control points moved by an offset)
    X = cps + 0.05 * normals
    _, _, M2 = fp.loads(body, Cp, 1.0, 1.2, Xac=Xac, X=X, correct_kuttacondition=kutta)
    assert rel(M2, orc.mtot(X, Fo, Xac)) < 1e-12
    # the field-based mirrors agree with the device sums
    fp.calcfield_F(body, 1.0, 1.2, correct_kuttacondition=kutta)
    assert rel(fp.calcfield_Ftot(body), Ftot) < 1e-12
    assert rel(fp.calcfield_Mtot(body, Xac), Mtot) < 1e-10
    if not tri:
        n1, n2 = g.dims
        f, pos = fp.calcfield_sectionalforce(body)
        fo, po = orc.sectionalforce(F, cps, n1, n2, [0.0, 1.0, 0.0])
        assert rel(pos, po) < 1e-13 and rel(f, fo) < 1e-12
