"""CPU FP64 oracle driver for the FLOWPanel panel-method hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product path
(the HIP library behind include/flowpanel_hip.h) never does.

It loads ``oracle/liboracle.so`` (built from ``oracle/fp_oracle.c`` by
``oracle/build_oracle.py``) and composes the reference's body-level algorithms
from it:

* ``assemble``      -- ``_G_Uvortexring!`` / ``_G_U!`` (src/FLOWPanel_liftingbody.jl:455-572,
                       src/FLOWPanel_nonliftingbody.jl:141-186), MultiBody column blocks
                       (src/FLOWPanel_multibody.jl:502-555)
* ``rhs``           -- ``calc_bc_noflowthrough`` (src/FLOWPanel_linearsolver.jl:297-328)
* ``solve_lu``      -- ``solve_ludiv!`` = LAPACK getrf + getrs (src/FLOWPanel_linearsolver.jl:57-78),
                       here scipy.linalg.lu_factor/lu_solve (the same LAPACK routines)
* ``uind``/``phi``  -- ``Uind!``/``phi!`` (src/FLOWPanel_liftingbody.jl:576-742,
                       src/FLOWPanel_nonliftingbody.jl:188-249)
* ``cp``, ``forces``, ``lds`` -- src/FLOWPanel_postprocess.jl:1059-1112, 1199-1307, 1615-1652

Bodies are passed as plain dicts/objects with attributes ``nodes`` (3 x nn,
float64), ``cell_ptr``/``cell_idx`` (0-based CSR), ``shedding`` (6 x ns,
0-based, -1 sentinel) and the kernel scalars.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int64)


def lib():
    """Load liboracle.so (build it first with oracle/build_oracle.py)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build_oracle  # noqa: F401  (relative import when used as a package)
        _LIB = ctypes.CDLL(path)
        L = _LIB
        i64, dbl, cint = ctypes.c_int64, ctypes.c_double, ctypes.c_int
        L.orc_panel_frame.argtypes = [_d, _i, _d, _d, _d]
        L.orc_geometry.argtypes = [_d, _i, _i, i64, dbl, cint, _d, _d, _d]
        L.orc_U_boundvortex.argtypes = [_d, _d, dbl, _d, i64, _d, _d, dbl, dbl]
        L.orc_U_vortexring.argtypes = [_d, _i, cint, dbl, _d, i64, _d, _d, dbl, dbl]
        L.orc_U_semiinfinite_vortex.argtypes = [_d, _d, dbl, _d, i64, _d, _d, dbl, dbl]
        L.orc_U_semiinfinite_vortex.restype = cint
        L.orc_U_semiinfinite_horseshoe.argtypes = [_d, _d, _d, _d, dbl, _d, i64, _d, _d, dbl, dbl]
        L.orc_U_semiinfinite_horseshoe.restype = cint
        L.orc_U_constant_source.argtypes = [_d, _i, cint, dbl, _d, i64, _d, _d, dbl]
        L.orc_phi_constant_source.argtypes = [_d, _i, cint, dbl, _d, i64, _d, dbl]
        L.orc_phi_constant_doublet.argtypes = [_d, _i, cint, dbl, _d, i64, _d]
        L.orc_phi_semiinfinite_doublet.argtypes = [_d, _d, _d, dbl, _d, i64, _d]
        L.orc_phi_semiinfinite_doublet.restype = cint
        L.orc_phi_semiinfinite_doublet2.argtypes = [_d, _d, _d, _d, dbl, _d, i64, _d, dbl]
        L.orc_phi_semiinfinite_doublet2.restype = cint
        L.orc_assemble_vortexring.argtypes = [_d, _i, _i, i64, _i, i64, _d, _d, _d, _d, i64,
                                              dbl, dbl, _d, i64]
        L.orc_assemble_vortexring.restype = cint
        L.orc_assemble_source.argtypes = [_d, _i, _i, i64, _d, _d, i64, dbl, _d, i64]
        L.orc_uind_vortexring.argtypes = [_d, _i, _i, i64, _i, i64, _d, _d, _d, _d, i64, _d, dbl, dbl]
        L.orc_uind_vortexring.restype = cint
        L.orc_uind_source.argtypes = [_d, _i, _i, i64, _d, _d, i64, _d, dbl]
        L.orc_phi_vortexring.argtypes = [_d, _i, _i, i64, _i, i64, _d, _d, _d, _d, i64, _d]
        L.orc_phi_vortexring.restype = cint
        L.orc_phi_source.argtypes = [_d, _i, _i, i64, _d, _d, i64, _d, dbl]
        L.orc_calcfield_cp.argtypes = [_d, i64, dbl, _i, i64, cint, _d]
        L.orc_calcfield_F.argtypes = [_d, _d, _d, i64, dbl, dbl, _i, i64, cint, _d]
        L.orc_ugradmu_cell.argtypes = [_d, _i, _i, i64, _i, cint, _d, _d, _d, _d, dbl, _i, i64, cint, _d]
    return _LIB


def _pd(a):
    return a.ctypes.data_as(_d) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(_i) if a is not None else None


def _f64(a, shape=None):
    a = np.asarray(a, dtype=np.float64, order="F")
    if shape is not None:
        a = a.reshape(shape, order="F")
    return np.asfortranarray(a)


def _i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _shed(sh):
    """6 x ns shedding matrix -> edge-major contiguous int64 (6 values per edge)."""
    sh = np.asarray(sh, dtype=np.int64).reshape(6, -1)
    return np.ascontiguousarray(sh.T).reshape(-1)


class OracleError(RuntimeError):
    pass


def _check(st):
    if st == 3:
        raise OracleError("Found non-unitary semi-infinite direction")
    if st == 4:
        raise OracleError("Invalid semi-infinite directions (opposite or orthogonal)")
    if st != 0:
        raise OracleError("oracle status %d" % st)


# ----------------------------------------------------------------------------- elements
def panel_frame(nodes, panel):
    nodes = _f64(nodes)
    panel = _i64(panel)
    t, o, n = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().orc_panel_frame(_pd(nodes), _pi(panel), _pd(t), _pd(o), _pd(n))
    return t, o, n


def U_boundvortex(pa, pb, strength, targets, out, dot_with=None, cutoff=1e-14, offset=1e-8):
    targets = _f64(targets)
    lib().orc_U_boundvortex(_pd(_f64(pa)), _pd(_f64(pb)), strength, _pd(targets), targets.shape[1],
                            _pd(out), _pd(None if dot_with is None else _f64(dot_with)), cutoff, offset)
    return out


def U_vortexring(nodes, panel, strength, targets, out, dot_with=None, cutoff=1e-14, offset=1e-8):
    nodes, targets, panel = _f64(nodes), _f64(targets), _i64(panel)
    dw = None if dot_with is None else _f64(dot_with)
    lib().orc_U_vortexring(_pd(nodes), _pi(panel), len(panel), strength, _pd(targets), targets.shape[1],
                           _pd(out), _pd(dw), cutoff, offset)
    return out


def U_semiinfinite_vortex(p, d, strength, targets, out, dot_with=None, cutoff=1e-14, offset=1e-8):
    targets = _f64(targets)
    dw = None if dot_with is None else _f64(dot_with)
    _check(lib().orc_U_semiinfinite_vortex(_pd(_f64(p)), _pd(_f64(d)), strength, _pd(targets),
                                           targets.shape[1], _pd(out), _pd(dw), cutoff, offset))
    return out


def U_semiinfinite_horseshoe(pa, pb, da, db, strength, targets, out, dot_with=None, cutoff=1e-14,
                             offset=1e-8):
    targets = _f64(targets)
    dw = None if dot_with is None else _f64(dot_with)
    _check(lib().orc_U_semiinfinite_horseshoe(_pd(_f64(pa)), _pd(_f64(pb)), _pd(_f64(da)), _pd(_f64(db)),
                                              strength, _pd(targets), targets.shape[1], _pd(out), _pd(dw),
                                              cutoff, offset))
    return out


def U_constant_source(nodes, panel, strength, targets, out, dot_with=None, offset=1e-8):
    nodes, targets, panel = _f64(nodes), _f64(targets), _i64(panel)
    dw = None if dot_with is None else _f64(dot_with)
    lib().orc_U_constant_source(_pd(nodes), _pi(panel), len(panel), strength, _pd(targets),
                                targets.shape[1], _pd(out), _pd(dw), offset)
    return out


def phi_constant_source(nodes, panel, strength, targets, out, offset=1e-8):
    nodes, targets, panel = _f64(nodes), _f64(targets), _i64(panel)
    lib().orc_phi_constant_source(_pd(nodes), _pi(panel), len(panel), strength, _pd(targets),
                                  targets.shape[1], _pd(out), offset)
    return out


def phi_constant_doublet(nodes, panel, strength, targets, out):
    nodes, targets, panel = _f64(nodes), _f64(targets), _i64(panel)
    lib().orc_phi_constant_doublet(_pd(nodes), _pi(panel), len(panel), strength, _pd(targets),
                                   targets.shape[1], _pd(out))
    return out


def phi_semiinfinite_doublet(pa, pb, d, strength, targets, out):
    targets = _f64(targets)
    _check(lib().orc_phi_semiinfinite_doublet(_pd(_f64(pa)), _pd(_f64(pb)), _pd(_f64(d)), strength,
                                              _pd(targets), targets.shape[1], _pd(out)))
    return out


def phi_semiinfinite_doublet2(pa, pb, da, db, strength, targets, out, infinitelength=1e12):
    targets = _f64(targets)
    _check(lib().orc_phi_semiinfinite_doublet2(_pd(_f64(pa)), _pd(_f64(pb)), _pd(_f64(da)), _pd(_f64(db)),
                                               strength, _pd(targets), targets.shape[1], _pd(out),
                                               infinitelength))
    return out


# ----------------------------------------------------------------------------- bodies
def geometry(nodes, cell_ptr, cell_idx, cpoffset=1e-14, charlen=0):
    """Normals, control points and areas (src/FLOWPanel_abstractbody.jl:635-927)."""
    nodes = _f64(nodes)
    cell_ptr, cell_idx = _i64(cell_ptr), _i64(cell_idx)
    n = len(cell_ptr) - 1
    normals = np.zeros((3, n), order="F")
    cps = np.zeros((3, n), order="F")
    areas = np.zeros(n)
    lib().orc_geometry(_pd(nodes), _pi(cell_ptr), _pi(cell_idx), n, cpoffset, charlen,
                       _pd(normals), _pd(cps), _pd(areas))
    return normals, cps, areas


def _body_geometry(b):
    return geometry(b.nodes, b.cell_ptr, b.cell_idx, b.CPoffset, b.charlen_kind)


def assemble(bodies, Das=None, Dbs=None, targets=None, normals=None):
    """Column-block AIC of one or several bodies (MultiBody composition).

    ``Das``/``Dbs`` are lists (one 3 x ns array per lifting body, None for
    non-lifting bodies).  Targets default to the concatenated control points.
    """
    if not isinstance(bodies, (list, tuple)):
        bodies, Das, Dbs = [bodies], [Das], [Dbs]
    geo = [_body_geometry(b) for b in bodies]
    if targets is None:
        targets = np.hstack([g[1] for g in geo])
        normals = np.hstack([g[0] for g in geo])
    targets, normals = _f64(targets), _f64(normals)
    M = targets.shape[1]
    N = sum(b.ncells for b in bodies)
    G = np.zeros((M, N), order="F")
    col = 0
    L = lib()
    for bi, b in enumerate(bodies):
        sub = G[:, col:col + b.ncells]
        assert sub.flags["F_CONTIGUOUS"]
        nodes = _f64(b.nodes)
        if b.element == "vortexring":
            sh = _shed(b.shedding)
            ns = b.nsheddings
            da = _f64(Das[bi], (3, ns)) if ns else np.zeros((3, 1), order="F")
            db = _f64(Dbs[bi], (3, ns)) if ns else np.zeros((3, 1), order="F")
            _check(L.orc_assemble_vortexring(_pd(nodes), _pi(_i64(b.cell_ptr)), _pi(_i64(b.cell_idx)),
                                             b.ncells, _pi(sh), ns, _pd(da), _pd(db), _pd(targets),
                                             _pd(normals), M, b.kerneloffset, b.kernelcutoff,
                                             _pd(sub), M))
        elif b.element == "source":
            L.orc_assemble_source(_pd(nodes), _pi(_i64(b.cell_ptr)), _pi(_i64(b.cell_idx)), b.ncells,
                                  _pd(targets), _pd(normals), M, b.kerneloffset, _pd(sub), M)
        else:
            raise OracleError("unsupported element " + str(b.element))
        col += b.ncells
    return G, normals, targets


def rhs(Uinfs, normals):
    """RHS_i = -Uinf_i . n_i (src/FLOWPanel_linearsolver.jl:297-328)."""
    return -np.einsum("ij,ij->j", np.asarray(Uinfs, dtype=np.float64), np.asarray(normals))


def solve_lu(G, b):
    """getrf + getrs (LAPACK), as ``solve_ludiv!``."""
    import scipy.linalg as sla
    lu = sla.lu_factor(G, check_finite=False)
    return sla.lu_solve(lu, b, check_finite=False)


def uind(body, strength, targets, Das=None, Dbs=None):
    """Induced velocity (3 x nt) of a solved body at ``targets``."""
    targets = _f64(targets)
    nt = targets.shape[1]
    out = np.zeros((3, nt), order="F")
    strength = np.ascontiguousarray(strength, dtype=np.float64)
    nodes = _f64(body.nodes)
    L = lib()
    if body.element == "vortexring":
        ns = body.nsheddings
        sh = _shed(body.shedding)
        da = _f64(Das, (3, ns)) if ns else np.zeros((3, 1), order="F")
        db = _f64(Dbs, (3, ns)) if ns else np.zeros((3, 1), order="F")
        _check(L.orc_uind_vortexring(_pd(nodes), _pi(_i64(body.cell_ptr)), _pi(_i64(body.cell_idx)),
                                     body.ncells, _pi(sh), ns, _pd(da), _pd(db), _pd(strength),
                                     _pd(targets), nt, _pd(out), body.kerneloffset, body.kernelcutoff))
    else:
        L.orc_uind_source(_pd(nodes), _pi(_i64(body.cell_ptr)), _pi(_i64(body.cell_idx)), body.ncells,
                          _pd(strength), _pd(targets), nt, _pd(out), body.kerneloffset)
    return out


def phi(body, strength, targets, Das=None, Dbs=None):
    targets = _f64(targets)
    nt = targets.shape[1]
    out = np.zeros(nt)
    strength = np.ascontiguousarray(strength, dtype=np.float64)
    nodes = _f64(body.nodes)
    L = lib()
    if body.element == "vortexring":
        ns = body.nsheddings
        sh = _shed(body.shedding)
        da = _f64(Das, (3, ns)) if ns else np.zeros((3, 1), order="F")
        db = _f64(Dbs, (3, ns)) if ns else np.zeros((3, 1), order="F")
        _check(L.orc_phi_vortexring(_pd(nodes), _pi(_i64(body.cell_ptr)), _pi(_i64(body.cell_idx)),
                                    body.ncells, _pi(sh), ns, _pd(da), _pd(db), _pd(strength),
                                    _pd(targets), nt, _pd(out)))
    else:
        L.orc_phi_source(_pd(nodes), _pi(_i64(body.cell_ptr)), _pi(_i64(body.cell_idx)), body.ncells,
                         _pd(strength), _pd(targets), nt, _pd(out), body.kerneloffset)
    return out


def cp(Us, uref, shedding=None, kutta=True):
    Us = _f64(Us)
    n = Us.shape[1]
    out = np.zeros(n)
    sh = None
    ns = 0
    if shedding is not None and np.asarray(shedding).size:
        sh = _shed(shedding)
        ns = len(sh) // 6
    lib().orc_calcfield_cp(_pd(Us), n, uref, _pi(sh), ns, int(bool(kutta and sh is not None)), _pd(out))
    return out


def forces(areas, normals, cps_, uinf, rho, shedding=None, kutta=True):
    normals = _f64(normals)
    n = normals.shape[1]
    out = np.zeros((3, n), order="F")
    sh = None
    ns = 0
    if shedding is not None and np.asarray(shedding).size:
        sh = _shed(shedding)
        ns = len(sh) // 6
    lib().orc_calcfield_F(_pd(np.ascontiguousarray(areas, dtype=np.float64)), _pd(normals),
                          _pd(np.ascontiguousarray(cps_, dtype=np.float64)), n, uinf, rho, _pi(sh), ns,
                          int(bool(kutta and sh is not None)), _pd(out))
    return out


def mtot(X, F, Xac):
    """Integrated moment about Xac of forces F applied at X (src/FLOWPanel_postprocess.jl:1715-1745)."""
    X, F = np.asarray(X, float), np.asarray(F, float)
    out = np.zeros(3)
    for i in range(F.shape[1]):
        r = X[:, i] - np.asarray(Xac, float)
        out += [r[1] * F[2, i] - r[2] * F[1, i], r[2] * F[0, i] - r[0] * F[2, i], r[0] * F[1, i] - r[1] * F[0, i]]
    return out


def sectionalforce(F, X, nchord, nspan, spandir):
    """Sectional force per unit span of a chord-fastest structured grid (src/FLOWPanel_postprocess.jl:1411-1475):
    station sums of F, station position = mean of spandir . X, divided by the central-difference spacing
    (one-sided at the two ends)."""
    F, X = np.asarray(F, float), np.asarray(X, float)
    outf = np.zeros((3, nspan))
    pos = np.zeros(nspan)
    for j in range(nspan):
        cells = [i + nchord * j for i in range(nchord)]
        for c in cells:
            outf[:, j] += F[:, c]
        pos[j] = np.mean([np.dot(spandir, X[:, c]) for c in cells])
    for j in range(nspan):
        d = pos[1] - pos[0] if j == 0 else (pos[j] - pos[j - 1] if j == nspan - 1 else (pos[j + 1] - pos[j - 1]) / 2)
        outf[:, j] /= abs(d)
    return outf, pos


def lds(F, Lhat, Dhat, Shat=None):
    """Lift/drag/sideslip projections (src/FLOWPanel_postprocess.jl:1615-1652)."""
    Lhat, Dhat = np.asarray(Lhat, float), np.asarray(Dhat, float)
    if Shat is None:
        Shat = np.cross(Lhat, Dhat)
    Ftot = np.asarray(F).sum(axis=1)
    return np.stack([Lhat * Ftot.dot(Lhat), Dhat * Ftot.dot(Dhat), Shat * Ftot.dot(Shat)], axis=1)


def cell_neighbors(cell_ptr, cell_idx):
    """neighbor[c, k] = the other cell holding edge (k, k+1) of cell c, -1 if none."""
    cell_ptr, cell_idx = np.asarray(cell_ptr), np.asarray(cell_idx)
    n = len(cell_ptr) - 1
    maxnv = int(np.diff(cell_ptr).max())
    owners = {}
    for c in range(n):
        p = cell_idx[cell_ptr[c]:cell_ptr[c + 1]]
        for k in range(len(p)):
            owners.setdefault(frozenset((int(p[k]), int(p[(k + 1) % len(p)]))), []).append(c)
    nb = -np.ones((n, maxnv), dtype=np.int64)
    for c in range(n):
        p = cell_idx[cell_ptr[c]:cell_ptr[c + 1]]
        for k in range(len(p)):
            other = [o for o in owners[frozenset((int(p[k]), int(p[(k + 1) % len(p)])))] if o != c]
            if other:
                nb[c, k] = other[0]
    return nb


def ugradmu_cell(body, strength, shedding=None, te=True):
    """Green-Gauss cell-centred -1/2 grad(mu) (src/FLOWPanel_postprocess.jl:139-497), 3 x n."""
    normals, cps, areas = _body_geometry(body)
    nb = cell_neighbors(body.cell_ptr, body.cell_idx)
    out = np.zeros((3, body.ncells), order="F")
    sh = _shed(shedding) if shedding is not None and np.asarray(shedding).size else None
    ns = 0 if sh is None else len(sh) // 6
    lib().orc_ugradmu_cell(_pd(_f64(body.nodes)), _pi(_i64(body.cell_ptr)), _pi(_i64(body.cell_idx)), body.ncells,
                           _pi(np.ascontiguousarray(nb)), nb.shape[1],
                           _pd(np.ascontiguousarray(strength, dtype=np.float64)), _pd(normals), _pd(cps), _pd(areas),
                           float(body.CPoffset), _pi(sh), ns, int(bool(te and sh is not None)), _pd(out))
    return out


def nodal_mean(body, strength, areas=None):
    """Nodal values of a cell field: the area-weighted mean over the cells holding
    each node (the ``gt.get_nodal_data(grid, Gammas; areas)`` call of
    src/FLOWPanel_postprocess.jl:651; GeometricTools is not vendored, so this
    restates that call as an area-weighted average)."""
    if areas is None:
        areas = _body_geometry(body)[2]
    nodes = np.asarray(body.nodes)
    cell_ptr, cell_idx = np.asarray(body.cell_ptr), np.asarray(body.cell_idx)
    s = np.zeros(nodes.shape[1])
    w = np.zeros(nodes.shape[1])
    G = np.asarray(strength, dtype=np.float64)
    for c in range(len(cell_ptr) - 1):
        for p in cell_idx[cell_ptr[c]:cell_ptr[c + 1]]:
            s[p] += areas[c] * G[c]
            w[p] += areas[c]
    return np.where(w > 0, s / np.where(w > 0, w, 1.0), 0.0)


def te_sides(body, shedding, pairs):
    """Trailing-edge nodes and the cells on either side of the trailing edge.

    Side A holds the shedding panels pi (with pi+1 when ``pairs``: the two
    triangles of a structured triangulated grid's last quad), side B the
    partners pj (with pj-1).  The TE nodes are the shedding edges' nodes (one
    set when the two sides share their trailing-edge nodes, as a grid closed
    around the section loop does).
    This is the sharp-trailing-edge split of src/FLOWPanel_postprocess.jl:692-748
    (upper/lower TE cells of the loop dimension) written on the shedding matrix.
    """
    cell_ptr, cell_idx = np.asarray(body.cell_ptr), np.asarray(body.cell_idx)
    n = len(cell_ptr) - 1
    A, B, te = [], [], []
    for pi, nia, nib, pj, nja, njb in np.asarray(shedding, dtype=np.int64).T:
        for c in ((pi, pi + 1) if pairs else (pi,)):
            if 0 <= c < n and c not in A:
                A.append(int(c))
        if pj >= 0:
            for c in ((pj, pj - 1) if pairs else (pj,)):
                if 0 <= c < n and c not in B:
                    B.append(int(c))
        ends = [cell_idx[cell_ptr[pi] + nia], cell_idx[cell_ptr[pi] + nib]]
        if pj >= 0:
            ends += [cell_idx[cell_ptr[pj] + nja], cell_idx[cell_ptr[pj] + njb]]
        for v in ends:
            if int(v) not in te:
                te.append(int(v))
    return A, B, te


def _node_grad(nodes, cell_ptr, cell_idx, c, vals, sgn):
    v = cell_idx[cell_ptr[c]:cell_ptr[c] + 3]
    X = nodes[:, v]
    e1 = X[:, 1] - X[:, 0]
    e1 /= np.linalg.norm(e1)
    w = X[:, 2] - X[:, 0]
    e2 = w - w.dot(e1) * e1
    e2 /= np.linalg.norm(e2)
    t = np.array([[0.0, 0.0], [(X[:, 1] - X[:, 0]).dot(e1), 0.0], [w.dot(e1), w.dot(e2)]])
    t0 = t.mean(axis=0)
    M = np.column_stack([np.ones(3), t - t0])
    a = np.linalg.solve(M, sgn * vals[v])
    return -0.5 * (a[1] * e1 + a[2] * e2)


def ugradmu_node(body, strength, shedding=None, sharp_te=0):
    """Node-based -1/2 grad(mu) (src/FLOWPanel_postprocess.jl:622-748), 3 x n, and the nodal values.

    Per cell: the plane through the first three vertices carrying the
    (CPoffset-signed) nodal values, solved as a 3 x 3 system in an in-plane frame
    (np.linalg.solve), its gradient mapped back to 3-D.  ``sharp_te`` (1: side
    cells (pi, pi+1) / (pj, pj-1), 2: pi / pj, see ``te_sides``): the cells on
    each side of the trailing edge take TE nodal values averaged over that side's
    cells only, as the reference's sharpTE pass (:692-743); the nodal values
    returned then hold side B's values at the TE nodes, as the reference leaves them.
    """
    normals, cps, areas = _body_geometry(body)
    nodal = nodal_mean(body, strength, areas)
    sgn = -1.0 if body.CPoffset < 0 else 1.0
    nodes = np.asarray(body.nodes, dtype=np.float64)
    cell_ptr, cell_idx = np.asarray(body.cell_ptr), np.asarray(body.cell_idx)
    n = len(cell_ptr) - 1
    out = np.zeros((3, n), order="F")
    for c in range(n):
        out[:, c] = _node_grad(nodes, cell_ptr, cell_idx, c, nodal, sgn)
    if sharp_te and shedding is not None and np.asarray(shedding).size:
        A, B, te = te_sides(body, shedding, pairs=sharp_te == 1)
        G = np.asarray(strength, dtype=np.float64)
        for cells in (A, B):
            side = nodal.copy()
            for v in te:
                s = w = 0.0
                for c in cells:
                    if v in cell_idx[cell_ptr[c]:cell_ptr[c + 1]]:
                        s += areas[c] * G[c]
                        w += areas[c]
                if w > 0:
                    side[v] = s / w
            for c in cells:
                out[:, c] = _node_grad(nodes, cell_ptr, cell_idx, c, side, sgn)
        nodal = side
    return out, nodal


def ugradmu_combined(body, strength, shedding=None, force_cell_te=True, anglecrit=30.0, sharp_te=0):
    """Combined -1/2 grad(mu) (src/FLOWPanel_postprocess.jl:756-921), 3 x n:
    node-based values on cells with an edge without neighbour or a neighbour normal
    more than ``anglecrit`` degrees away, cell-centred values (trailing-edge variant)
    elsewhere; with ``force_cell_te`` the shedding panels pi, pi+1, pj, pj-1 keep the
    cell-centred values."""
    cell = ugradmu_cell(body, strength, shedding, te=True)
    node, _ = ugradmu_node(body, strength, shedding, sharp_te)
    n = cell.shape[1]
    if force_cell_te and shedding is not None and np.asarray(shedding).size:
        for pi, nia, nib, pj, nja, njb in np.asarray(shedding, dtype=np.int64).T:
            for c in (pi, pi + 1) + ((pj, pj - 1) if pj >= 0 else ()):
                if 0 <= c < n:
                    node[:, c] = cell[:, c]
    normals = _body_geometry(body)[0]
    nb = cell_neighbors(body.cell_ptr, body.cell_idx)
    cosc = np.cos(np.deg2rad(anglecrit))
    cell_ptr = np.asarray(body.cell_ptr)
    out = np.zeros((3, n), order="F")
    for c in range(n):
        use_node = False
        for k in range(cell_ptr[c + 1] - cell_ptr[c]):
            m = nb[c, k]
            if m < 0 or normals[:, c].dot(normals[:, m]) < cosc:
                use_node = True
                break
        out[:, c] = node[:, c] if use_node else cell[:, c]
    return out
