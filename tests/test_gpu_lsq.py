"""Watertight least-squares formulation (RigidWakeBody{VortexRing, 2}) on the GPU.

The reference prescribes strengths (elprescribe, default the first element = 0),
moves them to the right-hand side, drops their columns and solves the normal
equations G_red^T G_red x = G_red^T b (src/FLOWPanel_liftingbody.jl:236-447).
Checked against the oracle's AIC with numpy's normal equations (same system:
1e-9 relative) and, loosely, against a QR least-squares solve.
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


def _reference_lsq(orc, body, U, pres, D=None):
    G, normals, _ = orc.assemble([_obody(body)], [D], [D])
    rhs = orc.rhs(U, normals)
    keep = np.ones(body.ncells, bool)
    for i, v in pres:
        rhs = rhs - v * G[:, i]
        keep[i] = False
    Gr = G[:, keep]
    x = np.linalg.solve(Gr.T @ Gr, Gr.T @ rhs)
    xq = np.linalg.lstsq(Gr, rhs, rcond=None)[0]
    out, outq = np.zeros(body.ncells), np.zeros(body.ncells)
    out[keep], outq[keep] = x, xq
    for i, v in pres:
        out[i] = outq[i] = v
    return out, outq


@pytest.mark.parametrize("pres", [None, [(0, 0.3), (57, -0.1)]])
@pytest.mark.parametrize("solver", ["solve_ludiv", "solve_lu32_ir"])
def test_watertight_sphere_lsq(fp, geo, orc, pres, solver):
    g = geo.sphere(n_meridional=10, n_azimuthal=16)
    body = fp.RigidWakeBody(g, None, formulation=2)
    U = np.tile([[1.0], [0.2], [0.0]], (1, body.ncells))
    st = fp.solve(body, U, solver=getattr(fp, solver), elprescribe=pres)
    ref, refq = _reference_lsq(orc, body, U, pres if pres is not None else [(0, 0.0)])
    print("lsq %s: rel err %.2e (normal eq.), %.2e (QR), residual %.1e" %
          (solver, rel(body.strength, ref), rel(body.strength, refq), st["residual"]), flush=True)
    assert rel(body.strength, ref) < 1e-9
    assert rel(body.strength, refq) < 1e-6
    for i, v in (pres or [(0, 0.0)]):
        assert body.strength[i] == v


def test_lsq_errors(fp, geo):
    g = geo.sphere(n_meridional=6, n_azimuthal=8)
    body = fp.RigidWakeBody(g, None, formulation=2)
    U = np.tile([[1.0], [0.0], [0.0]], (1, body.ncells))
    with pytest.raises(fp.FPError):
        fp.solve(body, U, elprescribe=[(body.ncells, 0.0)])
    with pytest.raises(fp.FPError):
        fp.solve(body, U, elprescribe=[(3, 0.0), (3, 1.0)])
