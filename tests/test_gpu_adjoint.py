"""Lift/drag objective and its hand-written adjoint on the GPU.

* the device objective (CL, CDi) against the oracle's composition of the same
  reference steps (geometry, AIC, LU solve, offset-probe velocities, Cp with
  Kutta averaging, forces, L/D projections): 1e-10 relative;
* the adjoint gradient (all nodes and Uinf) against central finite differences
  (h = 1e-6) of the device forward and of the oracle's forward: 1e-5 relative,
  measured norm-wise over the sampled components, and component-wise against
  1e-5 of the largest sampled component;
* the mixed-precision adjoint (FP32 inner solves, FP64 refinement) against the
  FP64 one.
The reference's ForwardDiff gradient needs Julia and is not available here
(parity unpinned); finite differences of the FP64 forward stand in for it.
"""
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-6
GRAD_TOL = 1e-5


def _obody(b):
    return types.SimpleNamespace(nodes=b.grid.nodes, cell_ptr=b.grid.cell_ptr, cell_idx=b.grid.cell_idx,
                                 ncells=b.ncells, element="vortexring", shedding=b.shedding,
                                 nsheddings=b.nsheddings, kerneloffset=b.kerneloffset,
                                 kernelcutoff=b.kernelcutoff, CPoffset=b.CPoffset,
                                 charlen_kind={"unitary": 0, "bbox": 1, "maxdist": 2, "sqrtarea": 3}[
                                     b.characteristiclength])


def _kutta_np(Cp, shedding, mode):
    Cp = Cp.copy()
    for e in range(shedding.shape[1]):
        pi, pj = shedding[0, e], shedding[3, e]
        if mode == "reference":
            ids = [pi, pi + 1] + ([pj, pj - 1] if pj != -1 else [])
        elif mode == "pair" and pj != -1:
            ids = [pi, pj]
        else:
            continue
        Cp[ids] = np.mean(Cp[ids])
    return Cp


def oracle_lift_drag(orc, body, nodes, U, probe_offset, probe_length, span, S, kutta):
    """The reference's offset-probe workflow composed from the oracle's steps."""
    ob = _obody(body)
    ob.nodes = np.asfortranarray(nodes)
    U = np.asarray(U, float)
    un = np.linalg.norm(U)
    d = U / un
    D = np.tile(d[:, None], (1, body.nsheddings))
    G, normals, cps = orc.assemble([ob], [D], [D])
    _, _, areas = orc.geometry(ob.nodes, ob.cell_ptr, ob.cell_idx, ob.CPoffset, ob.charlen_kind)
    gam = orc.solve_lu(G, orc.rhs(np.tile(U[:, None], (1, body.ncells)), normals))
    _, probes, _ = orc.geometry(ob.nodes, ob.cell_ptr, ob.cell_idx, probe_offset * probe_length, 0)
    V = U[:, None] + orc.uind(ob, gam, probes, D, D)
    if kutta == "reference":
        Cp = orc.cp(V, un, body.shedding, kutta=True)
    else:
        Cp = _kutta_np(orc.cp(V, un, None, kutta=False), body.shedding, kutta)
    F = orc.forces(areas, normals, Cp, un, 1.0, None, kutta=False)
    q = 0.5 * un ** 2
    w = np.cross(d, span)
    Lh = w / np.linalg.norm(w)
    Ft = F.sum(axis=1)
    return Ft.dot(Lh) / (q * S), Ft.dot(d) / (q * S)


def _wing(fp, geo, n_chord=16, n_span=8, tri=False, **kw):
    g, sh = geo.loft_wing(airfoil="2412", n_chord=n_chord, n_span=n_span, span=4.0, sweep_deg=15.0, taper=0.7,
                          twist_root_deg=0.0, twist_tip_deg=-3.0)
    if tri:
        g, sh = geo.triangulate(g, sh)
    return fp.RigidWakeBody(g, sh, **kw)


def _fd_nodes(fp, body, U, idx, kw):
    """Central differences of (CL, CDi) w.r.t. flattened node coordinates idx."""
    X0 = np.array(body.grid.nodes)
    out = np.zeros((2, len(idx)))
    for c, k in enumerate(idx):
        v, d = divmod(k, 3)
        vals = []
        for s in (+1, -1):
            X = X0.copy()
            X[d, v] += s * H
            fp.set_nodes(body, X)
            cl, cd, _ = fp.lift_drag(body, U, **kw)
            vals.append((cl, cd))
        out[:, c] = [(vals[0][0] - vals[1][0]) / (2 * H), (vals[0][1] - vals[1][1]) / (2 * H)]
    fp.set_nodes(body, X0)
    return out


def _fd_uinf(fp, body, U, kw):
    out = np.zeros((2, 3))
    for d in range(3):
        vals = []
        for s in (+1, -1):
            Up = np.array(U, float)
            Up[d] += s * H
            cl, cd, _ = fp.lift_drag(body, Up, **kw)
            vals.append((cl, cd))
        out[:, d] = [(vals[0][0] - vals[1][0]) / (2 * H), (vals[0][1] - vals[1][1]) / (2 * H)]
    return out


def _check(adj, fd, what):
    nrm = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
    cmp = np.abs(adj - fd).max() / np.abs(fd).max()
    print("%s: norm-wise rel %.2e, max component / max|g| %.2e" % (what, nrm, cmp), flush=True)
    assert nrm < GRAD_TOL and cmp < GRAD_TOL, (what, nrm, cmp)


KW = dict(probe_offset=0.02, probe_length=0.5, span_dir=(0, 1, 0), ref_area=2.6)


@pytest.mark.parametrize("tri", [False, True])
def test_lift_drag_matches_oracle(fp, geo, orc, tri):
    body = _wing(fp, geo, tri=tri)
    U = fp.direction(4.2, 2.0)
    cl, cd, st = fp.lift_drag(body, U, **KW)
    ocl, ocd = oracle_lift_drag(orc, body, body.grid.nodes, U, KW["probe_offset"], KW["probe_length"],
                                np.array(KW["span_dir"], float), KW["ref_area"], body.kutta)
    print("CL %.12f (oracle %.12f)  CDi %.12f (oracle %.12f)" % (cl, ocl, cd, ocd), flush=True)
    assert abs(cl - ocl) < 1e-10 * abs(ocl)
    assert abs(cd - ocd) < 1e-10 * max(abs(ocd), 1e-3 * abs(ocl))
    assert 0.1 < cl < 1.0  # a lifting wing at 4.2 deg


@pytest.mark.parametrize("tri", [False, True])
def test_adjoint_matches_fd(fp, geo, tri):
    body = _wing(fp, geo, tri=tri)
    U = 1.3 * fp.direction(5.0, 3.0)
    r = fp.lift_drag_gradient(body, U, **KW)
    cl, cd, _ = fp.lift_drag(body, U, **KW)
    assert abs(r["CL"] - cl) < 1e-13 and abs(r["CDi"] - cd) < 1e-13
    rng = np.random.default_rng(41 + tri)
    idx = rng.choice(3 * body.grid.nnodes, 24, replace=False)
    fd = _fd_nodes(fp, body, U, idx, KW)
    adj = np.stack([r["dCL_dnodes"][0].T.ravel()[idx], r["dCDi_dnodes"][0].T.ravel()[idx]])
    _check(adj[0], fd[0], "dCL/dnodes")
    _check(adj[1], fd[1], "dCDi/dnodes")
    fdu = _fd_uinf(fp, body, U, KW)
    _check(r["dCL_dUinf"], fdu[0], "dCL/dUinf")
    _check(r["dCDi_dUinf"], fdu[1], "dCDi/dUinf")
    # invariances: a rigid translation of all nodes changes nothing
    for g in (r["dCL_dnodes"][0], r["dCDi_dnodes"][0]):
        assert np.abs(g.sum(axis=1)).max() < 1e-9 * np.abs(g).max() * g.shape[1] ** 0.5


def test_adjoint_matches_oracle_fd(fp, geo, orc):
    body = _wing(fp, geo, n_chord=12, n_span=6, tri=True)
    U = fp.direction(3.0)
    r = fp.lift_drag_gradient(body, U, **KW)
    rng = np.random.default_rng(5)
    idx = rng.choice(3 * body.grid.nnodes, 12, replace=False)
    X0 = np.array(body.grid.nodes)
    span = np.array(KW["span_dir"], float)
    fd = np.zeros((2, len(idx)))
    for c, k in enumerate(idx):
        v, d = divmod(k, 3)
        vals = []
        for s in (+1, -1):
            X = X0.copy()
            X[d, v] += s * H
            vals.append(oracle_lift_drag(orc, body, X, U, KW["probe_offset"], KW["probe_length"], span,
                                         KW["ref_area"], body.kutta))
        fd[:, c] = [(vals[0][0] - vals[1][0]) / (2 * H), (vals[0][1] - vals[1][1]) / (2 * H)]
    adj = np.stack([r["dCL_dnodes"][0].T.ravel()[idx], r["dCDi_dnodes"][0].T.ravel()[idx]])
    _check(adj[0], fd[0], "dCL/dnodes vs oracle FD")
    _check(adj[1], fd[1], "dCDi/dnodes vs oracle FD")


@pytest.mark.parametrize("cp_len,probe_len", [("sqrtarea", "maxdist"), ("maxdist", "sqrtarea")])
def test_adjoint_offsets_and_lengths(fp, geo, cp_len, probe_len):
    # control points and probes offset by area- and vertex-distance-based lengths,
    # no Kutta averaging: every geometric derivative path is exercised.  (The
    # bounding-box length is left out: on a structured loft vertices tie in a
    # coordinate, where min/max have no derivative and central differences
    # average the two one-sided slopes.)
    body = _wing(fp, geo, n_chord=12, n_span=6, CPoffset=0.01, characteristiclength=cp_len)
    kw = dict(KW, probe_charlen=probe_len, probe_length=0.05, kutta="none")
    U = fp.direction(6.0, -2.0)
    r = fp.lift_drag_gradient(body, U, **kw)
    rng = np.random.default_rng(9)
    idx = rng.choice(3 * body.grid.nnodes, 20, replace=False)
    fd = _fd_nodes(fp, body, U, idx, kw)
    adj = np.stack([r["dCL_dnodes"][0].T.ravel()[idx], r["dCDi_dnodes"][0].T.ravel()[idx]])
    _check(adj[0], fd[0], "dCL/dnodes (offsets)")
    _check(adj[1], fd[1], "dCDi/dnodes (offsets)")


def test_adjoint_multibody(fp, geo):
    w1 = _wing(fp, geo, n_chord=10, n_span=6)
    g, sh = geo.loft_wing(airfoil="0012", n_chord=8, n_span=4, span=1.5, sweep_deg=30.0, taper=0.6)
    g.nodes = np.asfortranarray(g.nodes + np.array([[3.0], [0.0], [0.4]]))
    w2 = fp.RigidWakeBody(g, sh)
    mb = fp.MultiBody([w1, w2])
    U = fp.direction(4.0)
    r = fp.lift_drag_gradient(mb, U, **KW)
    rng = np.random.default_rng(3)
    for bi, b in enumerate((w1, w2)):
        idx = rng.choice(3 * b.grid.nnodes, 10, replace=False)
        X0 = np.array(b.grid.nodes)
        fd = np.zeros((2, len(idx)))
        for c, k in enumerate(idx):
            v, d = divmod(k, 3)
            vals = []
            for s in (+1, -1):
                X = X0.copy()
                X[d, v] += s * H
                fp.set_nodes(b, X)
                cl, cd, _ = fp.lift_drag(mb, U, **KW)
                vals.append((cl, cd))
            fd[:, c] = [(vals[0][0] - vals[1][0]) / (2 * H), (vals[0][1] - vals[1][1]) / (2 * H)]
        fp.set_nodes(b, X0)
        adj = np.stack([r["dCL_dnodes"][bi].T.ravel()[idx], r["dCDi_dnodes"][bi].T.ravel()[idx]])
        _check(adj[0], fd[0], "multibody dCL/dnodes body %d" % bi)
        _check(adj[1], fd[1], "multibody dCDi/dnodes body %d" % bi)


@pytest.mark.parametrize("solver", ["solve_gmres_ir", "solve_lu32_ir", "solve_gmres"])
def test_mixed_adjoint_matches_fp64(fp, geo, solver):
    body = _wing(fp, geo, n_chord=24, n_span=20)
    U = fp.direction(4.2)
    ref = fp.lift_drag_gradient(body, U, **KW)
    opt = {"precond_block": 48} if solver in ("solve_gmres_ir", "solve_gmres") else {}
    if solver == "solve_gmres":
        opt["tol"] = 1e-12
    r = fp.lift_drag_gradient(body, U, solver=getattr(fp, solver), solver_optargs=opt, **KW)
    assert abs(r["CL"] - ref["CL"]) < 1e-10 * abs(ref["CL"])
    for key in ("dCL_dnodes", "dCDi_dnodes"):
        a, b = r[key][0], ref[key][0]
        assert np.linalg.norm(a - b) < 1e-8 * np.linalg.norm(b), key
    for key in ("dCL_dUinf", "dCDi_dUinf"):
        assert np.linalg.norm(r[key] - ref[key]) < 1e-8 * np.linalg.norm(ref[key]), key


def test_set_nodes_roundtrip(fp, geo):
    body = _wing(fp, geo, n_chord=8, n_span=4)
    n0 = fp.calc_normals(body).copy()
    X0 = np.array(body.grid.nodes)
    fp.set_nodes(body, X0 * 2.0)
    assert np.abs(fp.calc_normals(body) - n0).max() < 1e-13       # scaling keeps the normals
    assert np.abs(fp.calc_areas(body)).sum() > 0
    fp.set_nodes(body, X0)
    assert np.abs(fp.calc_normals(body) - n0).max() == 0


def test_lift_drag_errors(fp, geo):
    body = _wing(fp, geo, n_chord=8, n_span=4)
    with pytest.raises(fp.FPError):
        fp.lift_drag(body, [0.0, 0.0, 0.0], **KW)
    with pytest.raises(fp.FPError):
        fp.lift_drag(body, [0.0, 1.0, 0.0], **KW)  # span direction parallel to Uinf
    src = fp.NonLiftingBody(geo.sphere(n_meridional=6, n_azimuthal=8))
    with pytest.raises(fp.FPError):
        fp.lift_drag(src, [1.0, 0.0, 0.0], **KW)
