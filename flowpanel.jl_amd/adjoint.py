"""Lift/drag coefficients of a lifting body and their hand-written adjoint.

``lift_drag`` evaluates, on the device, the reference's offset-probe force
workflow (test/runtests_solvers.jl:136-166):

    solve(body, Uinf at every control point, Das = Dbs = Uinf/|Uinf|)
    calcfield_U(body, body; offset, characteristiclength)    # probes off the surface
    calcfield_Cp(body, |Uinf|)                                # Kutta-averaged
    calcfield_F(body, |Uinf|, rho); calcfield_LDS(body, Lhat, Dhat)
    CL = L.Lhat / (q S),  CDi = D.Dhat / (q S)

with Dhat = Uinf/|Uinf| and Lhat = normalize(Dhat x span_dir).

``lift_drag_gradient`` returns the same coefficients and their gradients with
respect to every mesh node and to Uinf from one reverse pass (transpose solve
plus the reverse of the geometry, influence and post-processing kernels).  It
is what a ChainRulesCore ``rrule`` of ``(nodes, Uinf) -> (CL, CDi)`` returns as
its pullback: ``pullback(dCL, dCDi) = dCL * grad_CL + dCDi * grad_CDi``.

``set_nodes`` moves the mesh (same connectivity) in place, like editing
``grid._nodes`` in the reference.
"""
import ctypes

import numpy as np

from . import _native as nat
from .solvers import _CODES, solve_ludiv

KUTTA = {"none": nat.FP_KUTTA_NONE, "reference": nat.FP_KUTTA_REFERENCE, "pair": nat.FP_KUTTA_PAIR}


def set_nodes(body, nodes):
    """Replace the node coordinates (3 x nnodes) of a single-component body."""
    comps = body.bodies
    if len(comps) != 1:
        raise ValueError("set_nodes takes one component body")
    b = comps[0]
    X = nat.f64(nodes)
    if X.shape != (3, b.grid.nnodes):
        raise ValueError("nodes must be 3 x %d, got %s" % (b.grid.nnodes, X.shape))
    b.ctx.check(nat.lib().fp_body_set_nodes(b.ctx.handle, b.handle, nat.dptr(X)))
    b.grid.nodes = np.array(X)
    b._solved = False


def _opts(body, probe_offset, probe_length, probe_charlen, span_dir, ref_area, kutta, solver, solver_optargs):
    if solver not in _CODES:
        raise ValueError("solver must be one of %s" % [f.__name__ for f in _CODES])
    o = nat.LiftDragOpts()
    nat.lib().fp_liftdrag_opts_default(ctypes.byref(o), _CODES[solver])
    for k, v in (solver_optargs or {}).items():
        if not hasattr(o.solve, k):
            raise TypeError("unknown solver option %r" % k)
        setattr(o.solve, k, v)
    o.probe_offset = float(probe_offset)
    o.probe_length = float(probe_length)
    if probe_charlen not in nat.CHARLEN:
        raise ValueError("probe_charlen must be one of %s" % list(nat.CHARLEN))
    o.probe_charlen = nat.CHARLEN[probe_charlen]
    for d in range(3):
        o.span_dir[d] = float(span_dir[d])
    o.ref_area = float(ref_area)
    if kutta is None:
        kutta = body.bodies[0].kutta
    o.kutta_mode = KUTTA[kutta]
    return o


def _uinf(Uinf):
    U = np.ascontiguousarray(Uinf, dtype=np.float64).reshape(3)
    return U


def _mark_solved(body):
    for b in body.bodies:
        g = np.zeros(b.ncells)
        b.ctx.check(nat.lib().fp_body_get_strength(b.ctx.handle, b.handle, nat.dptr(g)))
        b.strength = g
        b._solved = True
        b.add_field(b.element.field, g)


def lift_drag(body, Uinf, probe_offset=0.02, probe_length=1.0, probe_charlen="unitary", span_dir=(0, 1, 0),
              ref_area=1.0, kutta=None, solver=solve_ludiv, solver_optargs=None):
    """(CL, CDi, stats) of the body at freestream Uinf (3,)."""
    o = _opts(body, probe_offset, probe_length, probe_charlen, span_dir, ref_area, kutta, solver, solver_optargs)
    U = _uinf(Uinf)
    cl, cd = np.zeros(1), np.zeros(1)
    s = nat.LiftDragStats()
    comps = body.bodies
    ctx = comps[0].ctx
    ctx.check(nat.lib().fp_lift_drag(ctx.handle, nat.handles(comps), len(comps), nat.dptr(U), ctypes.byref(o),
                                     nat.dptr(cl), nat.dptr(cd), ctypes.byref(s)))
    _mark_solved(body)
    return float(cl[0]), float(cd[0]), s.as_dict()


def lift_drag_gradient(body, Uinf, probe_offset=0.02, probe_length=1.0, probe_charlen="unitary",
                       span_dir=(0, 1, 0), ref_area=1.0, kutta=None, solver=solve_ludiv, solver_optargs=None):
    """CL, CDi and their gradients.

    Returns a dict with ``CL``, ``CDi``, ``dCL_dnodes`` / ``dCDi_dnodes`` (one
    3 x nnodes array per component body), ``dCL_dUinf`` / ``dCDi_dUinf`` (3,)
    and ``stats`` (device milliseconds of the forward and reverse passes).
    """
    o = _opts(body, probe_offset, probe_length, probe_charlen, span_dir, ref_area, kutta, solver, solver_optargs)
    U = _uinf(Uinf)
    comps = body.bodies
    ctx = comps[0].ctx
    nn = [b.grid.nnodes for b in comps]
    gl = np.zeros(3 * sum(nn))
    gd = np.zeros(3 * sum(nn))
    ul, ud = np.zeros(3), np.zeros(3)
    s = nat.LiftDragStats()
    ctx.check(nat.lib().fp_lift_drag_adjoint(ctx.handle, nat.handles(comps), len(comps), nat.dptr(U),
                                             ctypes.byref(o), nat.dptr(gl), nat.dptr(gd), nat.dptr(ul),
                                             nat.dptr(ud), ctypes.byref(s)))
    _mark_solved(body)
    out = {"CL": s.CL, "CDi": s.CDi, "dCL_dUinf": ul, "dCDi_dUinf": ud, "stats": s.as_dict(),
           "dCL_dnodes": [], "dCDi_dnodes": []}
    off = 0
    for k in nn:
        out["dCL_dnodes"].append(gl[3 * off:3 * (off + k)].reshape(k, 3).T.copy())
        out["dCDi_dnodes"].append(gd[3 * off:3 * (off + k)].reshape(k, 3).T.copy())
        off += k
    return out
