"""Field evaluation and post-processing on the device.

* ``Uind``         -- ``Uind!(body, targets, out)`` (src/FLOWPanel_abstractbody.jl:101-110)
* ``phi``          -- ``phi!(body, targets, out)`` (src/FLOWPanel_abstractbody.jl:118-127)
* ``calcfield_U``  -- src/FLOWPanel_postprocess.jl:28-110 (at the control points)
* ``calcfield_Ugradmu_cell`` -- src/FLOWPanel_postprocess.jl:139-497 (Green-Gauss)
* ``calcfield_Cp`` -- src/FLOWPanel_postprocess.jl:1059-1176
* ``calcfield_F``, ``calcfield_LDS`` -- src/FLOWPanel_postprocess.jl:1199-1307, 1615-1689
"""
import numpy as np

from . import _native as nat
from .bodies import calc_areas, calc_normals


def _check_targets(targets):
    t = nat.f64(targets)
    if t.ndim != 2 or t.shape[0] != 3:
        raise ValueError("targets must be 3 x nt, got %s" % (t.shape,))
    return t


def Uind(body, targets, out):
    """out (3 x nt) += velocity induced by the solved body at targets."""
    t = _check_targets(targets)
    if out.shape != t.shape:
        raise ValueError("out must be 3 x %d" % t.shape[1])
    o = nat.f64(out)
    comps = body.bodies
    body_ctx = comps[0].ctx
    body_ctx.check(nat.lib().fp_uind(body_ctx.handle, nat.handles(comps), len(comps), nat.dptr(t), t.shape[1],
                                     nat.dptr(o)))
    if o is not out:
        out[...] = o
    return out


def phi(body, targets, out):
    """out (nt) += potential of the solved body at targets."""
    t = _check_targets(targets)
    o = np.ascontiguousarray(out, dtype=np.float64)
    if o.shape != (t.shape[1],):
        raise ValueError("out must have length %d" % t.shape[1])
    comps = body.bodies
    ctx = comps[0].ctx
    ctx.check(nat.lib().fp_phi(ctx.handle, nat.handles(comps), len(comps), nat.dptr(t), t.shape[1], nat.dptr(o)))
    if o is not out:
        out[...] = o
    return out


def _charlen(body, kind):
    """Characteristic length of every panel (src/FLOWPanel_abstractbody.jl:571-633)."""
    g = body.grid
    if kind == "unitary":
        return np.ones(g.ncells)
    if kind == "sqrtarea":
        return np.sqrt(calc_areas(body))
    out = np.zeros(g.ncells)
    for i in range(g.ncells):
        P = g.nodes[:, g.cell_idx[g.cell_ptr[i]:g.cell_ptr[i + 1]]]
        if kind == "bbox":
            out[i] = np.linalg.norm(P.max(axis=1) - P.min(axis=1))
        elif kind == "maxdist":
            out[i] = np.linalg.norm(P - P[:, :1], axis=0).max()
        else:
            raise ValueError("unknown characteristic length %r" % (kind,))
    return out


def _controlpoints(body, offset, characteristiclength):
    """calc_controlpoints(body, normals; off, characteristiclength) (src/FLOWPanel_abstractbody.jl:635-711):
    the vertex mean of every panel plus off * l * n."""
    cps = []
    for b in body.bodies:
        g = b.grid
        counts = np.diff(g.cell_ptr)
        sums = np.add.reduceat(g.nodes[:, g.cell_idx], g.cell_ptr[:-1], axis=1)
        off = b.CPoffset if offset is None else float(offset)
        kind = b.characteristiclength if characteristiclength is None else characteristiclength
        cps.append(sums / counts + off * _charlen(b, kind)[None, :] * calc_normals(b))
    return np.asfortranarray(np.hstack(cps))


def calcfield_U(sourcebody, targetbody=None, offset=None, characteristiclength=None, fieldname="U",
                addfield=True):
    """calcfield_U(sourcebody, targetbody; offset, characteristiclength) (src/FLOWPanel_postprocess.jl:28-110):
    U = Uinf + velocity induced by ``sourcebody`` at the control points of ``targetbody``, 3 x N.

    The target's freestream is its field "Uinf" (set by ``solve``).  ``offset`` and
    ``characteristiclength`` ("unitary", "bbox", "maxdist", "sqrtarea") move the control points
    along the normals as the reference does; by default they are the target body's own.
    """
    if targetbody is None:
        targetbody = sourcebody
    comps = sourcebody.bodies
    ctx = comps[0].ctx
    n = targetbody.ncells
    out = np.zeros((3, n), order="F")
    if targetbody is sourcebody and offset is None and characteristiclength is None:
        ctx.check(nat.lib().fp_calcfield_U(ctx.handle, nat.handles(comps), len(comps), nat.dptr(out)))
    else:
        try:
            U = np.hstack([np.asarray(b.get_field("Uinf")).reshape(3, -1) for b in targetbody.bodies])
        except KeyError:
            raise nat.FPError(nat.FP_ERR_ARG, "Target body doesn't have freestream field `Uinf`")
        out += U
        Uind(sourcebody, _controlpoints(targetbody, offset, characteristiclength), out)
    if addfield:
        _store(targetbody, fieldname, out)
    return out


def calcfield_Ugradmu_cell(body, fieldname="Ugradmu", addfield=True, te_correction=True):
    """-1/2 grad(mu) at the cells, Green-Gauss (src/FLOWPanel_postprocess.jl:139-497), 3 x N."""
    comps = body.bodies
    ctx = comps[0].ctx
    out = np.zeros((3, body.ncells), order="F")
    ctx.check(nat.lib().fp_calcfield_ugradmu(ctx.handle, nat.handles(comps), len(comps), int(bool(te_correction)),
                                             nat.dptr(out)))
    if addfield:
        body.add_field(fieldname, out)
        col = 0
        for b in comps:
            b.add_field(fieldname, out[:, col:col + b.ncells])
            col += b.ncells
    return out


def _store(body, fieldname, out):
    body.add_field(fieldname, out)
    col = 0
    for b in body.bodies:
        b.add_field(fieldname, out[:, col:col + b.ncells])
        col += b.ncells


def _sharp_mode(b, sharpTE):
    """FP_SHARPTE_* of one component: the reference's structured triangulated layout (pairs of
    trailing-edge triangles, Kutta mode "reference") or single trailing-edge cells."""
    if not sharpTE:
        return nat.FP_SHARPTE_OFF
    if sharpTE in ("pairs", "cells"):
        return nat.FP_SHARPTE_PAIRS if sharpTE == "pairs" else nat.FP_SHARPTE_CELLS
    return nat.FP_SHARPTE_PAIRS if getattr(b, "kutta", "pair") == "reference" else nat.FP_SHARPTE_CELLS


def calcfield_Ugradmu_node(body, fieldname="Ugradmu", addfield=True, sharpTE=False):
    """-1/2 grad(mu) from nodal values of Gamma (src/FLOWPanel_postprocess.jl:622-748), 3 x N.

    The nodal values (area-weighted means over the cells holding each node) are
    stored as the field "gamma_node".  ``sharpTE`` (True, or "pairs"/"cells")
    averages the trailing-edge nodal values over each side's cells separately
    (:692-743): side cells (pi, pi+1)/(pj, pj-1) on triangulated structured grids,
    pi/pj otherwise.
    """
    out = np.zeros((3, body.ncells), order="F")
    nodals = []
    col = 0
    for b in body.bodies:
        o = np.zeros((3, b.ncells), order="F")
        nodal = np.zeros(max(int(b.nnodes), 1))
        b.ctx.check(nat.lib().fp_calcfield_ugradmu_node(b.ctx.handle, nat.handles([b]), 1, _sharp_mode(b, sharpTE),
                                                        nat.dptr(o), nat.dptr(nodal)))
        out[:, col:col + b.ncells] = o
        nodals.append(nodal[:b.nnodes])
        col += b.ncells
    if addfield:
        _store(body, fieldname, out)
        body.add_field("gamma_node", np.concatenate(nodals))
    return out


def calcfield_Ugradmu(body, fieldname="Ugradmu", addfield=True, sharpTE=False, force_cellTE=True, anglecrit=30):
    """Combined -1/2 grad(mu) (src/FLOWPanel_postprocess.jl:756-1000), 3 x N: the
    cell-centred value, except on cells at an open edge or a sharp edge (normals
    more than ``anglecrit`` degrees apart), which take the node-based value (with
    ``sharpTE`` as in ``calcfield_Ugradmu_node``); with ``force_cellTE`` the
    trailing-edge panels keep the cell-centred value."""
    out = np.zeros((3, body.ncells), order="F")
    col = 0
    for b in body.bodies:
        o = np.zeros((3, b.ncells), order="F")
        b.ctx.check(nat.lib().fp_calcfield_ugradmu_combined(b.ctx.handle, nat.handles([b]), 1, int(bool(force_cellTE)),
                                                            float(anglecrit), _sharp_mode(b, sharpTE), nat.dptr(o)))
        out[:, col:col + b.ncells] = o
        col += b.ncells
    if addfield:
        _store(body, fieldname, out)
    return out


def addfields(body, name1, name2):
    """field name2 += field name1 (src/FLOWPanel_abstractbody.jl add_field helpers)."""
    f = np.asarray(body.get_field(name2)) + np.asarray(body.get_field(name1))
    body.add_field(name2, f)
    col = 0
    for b in body.bodies:
        if b is not body:
            b.add_field(name2, f[..., col:col + b.ncells])
        col += b.ncells
    return f


def calcfield_Cp(body, Uref, Us=None, correct_kuttacondition=True, clip=None, fieldname="Cp", addfield=True,
                 U_fieldname="U"):
    """Cp = 1 - (|U|/Uref)^2 with Kutta averaging over each body's shedding edges."""
    comps = body.bodies
    if Us is None:
        Us = body.get_field(U_fieldname)
    Us = nat.f64(Us)
    out = np.zeros(body.ncells)
    col = 0
    for b in comps:
        o = np.zeros(b.ncells)
        U = nat.f64(Us[:, col:col + b.ncells])
        mode = nat.FP_KUTTA_NONE
        if correct_kuttacondition and b.nsheddings:
            mode = {"reference": nat.FP_KUTTA_REFERENCE, "pair": nat.FP_KUTTA_PAIR,
                    "none": nat.FP_KUTTA_NONE}[b.kutta]
        sh = np.ascontiguousarray(b.shedding.T) if b.nsheddings else None
        b.ctx.check(nat.lib().fp_calcfield_cp(b.ctx.handle, nat.dptr(U), b.ncells, float(Uref), nat.iptr(sh),
                                              b.nsheddings, 0, mode, nat.dptr(o)))
        if clip is not None:
            o = np.array([clip(v) for v in o])
        out[col:col + b.ncells] = o
        if addfield:
            b.add_field(fieldname, o)
        col += b.ncells
    if addfield:
        body.add_field(fieldname, out)
    return out


def _kutta_mode(b, correct):
    if not correct or not b.nsheddings or b.element.code != nat.FP_ELEM_VORTEXRING:
        return nat.FP_KUTTA_NONE
    return {"reference": nat.FP_KUTTA_REFERENCE, "pair": nat.FP_KUTTA_PAIR, "none": nat.FP_KUTTA_NONE}[b.kutta]


def loads(body, Cps, Uinf, rho, Xac=(0.0, 0.0, 0.0), X=None, correct_kuttacondition=True):
    """Element forces and integrated loads on the device (fp_calcfield_F): returns (F 3 x N, Ftot 3, Mtot 3),
    Mtot about ``Xac`` with every force applied at ``X`` (3 x N, default the control points)."""
    Cps = np.ascontiguousarray(Cps, dtype=np.float64)
    if Cps.shape != (body.ncells,):
        raise ValueError("Cps must have length %d" % body.ncells)
    Xh = None if X is None else nat.f64(X)
    if Xh is not None and Xh.shape != (3, body.ncells):
        raise ValueError("X must be 3 x %d" % body.ncells)
    F = np.zeros((3, body.ncells), order="F")
    Ftot, Mtot = np.zeros(3), np.zeros(3)
    xac = np.ascontiguousarray(Xac, dtype=np.float64)
    col = 0
    for b in body.bodies:
        n = b.ncells
        Fb = np.zeros((3, n), order="F")
        b.ctx.check(nat.lib().fp_calcfield_F(b.ctx.handle, nat.handles([b]), 1, nat.dptr(Cps[col:col + n]),
                                             float(Uinf), float(rho), _kutta_mode(b, correct_kuttacondition),
                                             nat.dptr(xac), nat.dptr(None if Xh is None else nat.f64(Xh[:, col:col + n])),
                                             nat.dptr(Fb), nat.dptr(Ftot), nat.dptr(Mtot)))
        F[:, col:col + n] = Fb
        col += n
    return F, Ftot, Mtot


def calcfield_F(body, Uinf, rho, Cps=None, areas=None, normals=None, correct_kuttacondition=True, fieldname="F",
                addfield=True, Cp_fieldname="Cp"):
    """calcfield_F (src/FLOWPanel_postprocess.jl:1199-1392): F_i = -Cp_i q A_i n_i, 3 x N, with the Kutta
    override of the trailing-edge cells.  Without ``areas``/``normals`` it runs on the device with the
    body's areas and its normals flipped by the sign of CPoffset (calc_normals(body; flipbyCPoffset=true));
    with them, on the host as the reference's calcfield_F!(out, body, areas, normals, Cps, Uinf, rho)."""
    if Cps is None:
        Cps = body.get_field(Cp_fieldname)
    Cps = np.asarray(Cps, dtype=np.float64)
    if areas is None and normals is None:
        F, Ftot, _ = loads(body, Cps, Uinf, rho, correct_kuttacondition=correct_kuttacondition)
    else:
        if areas is None:
            areas = calc_areas(body)
        if normals is None:
            normals = calc_normals(body)
        q = 0.5 * rho * Uinf ** 2
        F = -(Cps * q * np.asarray(areas))[None, :] * np.asarray(normals)
        if correct_kuttacondition:
            col = 0
            for b in body.bodies:
                for e in range(b.nsheddings):
                    pi, pj = b.shedding[0, e] + col, b.shedding[3, e]
                    if b.kutta == "reference":
                        ids = [pi, pi + 1] + ([pj + col, pj + col - 1] if pj != -1 else [])
                    else:
                        ids = [pi] + ([pj + col] if pj != -1 else [])
                    if len(ids) < 2:
                        continue
                    ave = np.mean([Cps[i] for i in ids])
                    for i in ids:
                        F[:, i] = -ave * q * areas[i] * normals[:, i]
                col += b.ncells
        Ftot = F.sum(axis=1)
    if addfield:
        _store(body, fieldname, F)
        body.add_field("Ftot", Ftot)
    return F


def calcfield_Ftot(body, Fs=None, F_fieldname="F", fieldname="Ftot", addfield=True):
    """Integrated force (src/FLOWPanel_postprocess.jl:1543-1590), 3."""
    F = np.asarray(body.get_field(F_fieldname) if Fs is None else Fs)
    out = F.sum(axis=1)
    if addfield:
        body.add_field(fieldname, out)
    return out


def _moment(body, Xac, F_fieldname, offset, characteristiclength, Fs):
    F = np.asarray(body.get_field(F_fieldname) if Fs is None else Fs)
    X = _controlpoints(body, offset, characteristiclength)
    return np.cross(X - np.asarray(Xac, dtype=np.float64).reshape(3, 1), F, axis=0).sum(axis=1)


def calcfield_Mtot(body, Xac, F_fieldname="F", offset=None, characteristiclength=None, Fs=None, fieldname="Mtot",
                   addfield=True):
    """Integrated moment about Xac (src/FLOWPanel_postprocess.jl:1715-1784): sum of (X_i - Xac) x F_i over
    the control points X (moved by ``offset``/``characteristiclength`` as calc_controlpoints does)."""
    out = _moment(body, Xac, F_fieldname, offset, characteristiclength, Fs)
    if addfield:
        body.add_field(fieldname, out)
    return out


def calcfield_lmn(body, Xac, lhat, mhat, nhat=None, F_fieldname="F", offset=None, characteristiclength=None,
                  Fs=None, addfield=True):
    """Rolling, pitching and yawing moments (src/FLOWPanel_postprocess.jl:1804-1904): 3 x 3, columns the
    projections of Mtot on lhat, mhat, nhat (default lhat x mhat)."""
    lhat, mhat = np.asarray(lhat, float), np.asarray(mhat, float)
    nhat = np.cross(lhat, mhat) if nhat is None else np.asarray(nhat, float)
    for v, nm in ((lhat, "lhat"), (mhat, "mhat"), (nhat, "nhat")):
        if abs(np.linalg.norm(v) - 1) > 2 * np.finfo(float).eps:
            raise ValueError("%s=%s is not a unitary vector" % (nm, v))
    M = _moment(body, Xac, F_fieldname, offset, characteristiclength, Fs)
    out = np.stack([lhat * M.dot(lhat), mhat * M.dot(mhat), nhat * M.dot(nhat)], axis=1)
    if addfield:
        body.add_field("Mroll", out[:, 0])
        body.add_field("Mpitch", out[:, 1])
        body.add_field("Myaw", out[:, 2])
    return out


def calcfield_sectionalforce(body, F_fieldname="F", Fs=None, dimspan=2, dimchord=1, spandirection=(0.0, 1.0, 0.0),
                             offset=None, characteristiclength=None, fieldname="sectionalforce", addfield=True):
    """Sectional force per unit span (src/FLOWPanel_postprocess.jl:1411-1542) of a structured grid, on the
    device: returns (3 x nspan, nspan positions).  The grid's ``dims`` = (cells along dimension 1, along 2)."""
    if len(body.bodies) != 1:
        raise ValueError("sectional loads take a single structured body")
    b = body.bodies[0]
    dims = getattr(b.grid, "dims", None)
    if dims is None:
        raise ValueError("sectional loads need a structured grid (grid.dims)")
    n1, n2 = dims
    F = nat.f64(body.get_field(F_fieldname) if Fs is None else Fs)
    X = _controlpoints(body, offset, characteristiclength)
    if (dimchord, dimspan) == (2, 1):      # chordwise index slow: reorder to chord-fastest
        order = np.arange(n1 * n2).reshape(n2, n1).T.ravel()
        F, X = nat.f64(F[:, order]), nat.f64(X[:, order])
        n1, n2 = n2, n1
    elif (dimchord, dimspan) != (1, 2):
        raise ValueError("dimchord/dimspan must be (1, 2) or (2, 1)")
    outF = np.zeros((3, n2), order="F")
    pos = np.zeros(n2)
    sd = np.ascontiguousarray(spandirection, dtype=np.float64)
    b.ctx.check(nat.lib().fp_calcfield_sectionalforce(b.ctx.handle, b.handle, nat.dptr(F), nat.dptr(nat.f64(X)),
                                                      n1, n2, nat.dptr(sd), nat.dptr(outF), nat.dptr(pos)))
    if addfield:
        body.add_field(fieldname, outF)
        body.add_field(fieldname + "-pos", pos)
    return outF, pos


def calcfield_LDS(body, Lhat, Dhat, Shat=None, F=None):
    """3 x 3 [L D S] projections of the integrated force."""
    if F is None:
        F = body.get_field("F")
    Lhat, Dhat = np.asarray(Lhat, float), np.asarray(Dhat, float)
    for v, nm in ((Lhat, "Lhat"), (Dhat, "Dhat")):
        if abs(np.linalg.norm(v) - 1) > 2 * np.finfo(float).eps:
            raise ValueError("%s=%s is not a unitary vector" % (nm, v))
    if Shat is None:
        Shat = np.cross(Lhat, Dhat)
    Ftot = np.asarray(F).sum(axis=1)
    out = np.stack([Lhat * Ftot.dot(Lhat), Dhat * Ftot.dot(Dhat), Shat * Ftot.dot(Shat)], axis=1)
    body.add_field("L", out[:, 0])
    body.add_field("D", out[:, 1])
    body.add_field("S", out[:, 2])
    return out
