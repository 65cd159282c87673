"""The reference's own swept-wing meshes solved end to end on the GPU.

examples/data/sweptwing-{left,right}000.vtu (the two halves of the 45-deg swept
RAE 101 wing of the reference's tests, test/runtests_meshes.jl:195-257) are read
from the fixtures, their trailing edges found by calc_shedding, and the lift is
evaluated with the offset-probe workflow of test/runtests_solvers.jl:136-166
(probe offset 0.02 x b/ar, Kutta averaging, L/D projections).  A loose,
physics-level pin: the reference test accepts CL within 3 % of the experimental
0.238 on its own spline-lofted mesh (test/runtests_solvers.jl:32, 166-170; the
lofting needs un-vendored GeometricTools/Dierckx), and the reference's published
values for this wing range from 0.2335 to 0.272 with the surface-velocity scheme
(docs/resources/data/sweptwing000-*.md).  On the stored meshes the offset probes
give CL = 0.2247 (-5.6 %); the check is CL within 10 % of 0.238.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
B = 98 * 0.0254
AR = 5.0


def _te(side):
    y = -1.2446 if side == "left" else 1.2446
    X1 = np.array([1.74244, y, 1.38094e-07])
    X2 = np.array([0.49784, 0.0, 1.38094e-07])
    te = np.stack([X1 + v * (X2 - X1) for v in np.linspace(0, 1, 1000)], axis=1)
    key = te[1] if side == "left" else -te[1]
    return te[:, np.argsort(key, kind="stable")]


def _wing(fp):
    bodies = []
    for side in ("left", "right"):
        g = fp.read_vtu(os.path.join(GOLDEN, "sweptwing-%s000.vtu" % side))
        if side == "right":
            # this half is stored with inward normals (the reference flips it with a
            # negative CPoffset); reverse its cells so that every normal points out
            g = fp.Grid(g.nodes, [g.cell(c)[::-1] for c in range(g.ncells)])
        te = _te(side)
        sh = fp.calc_shedding(g, te, tolerance=0.001 * np.ptp(te[1]))
        # structured triangulation (96 triangles per spanwise station): the reference's
        # Kutta pattern (pi, pi+1, pj, pj-1) applies
        bodies.append(fp.RigidWakeBody(g, sh, kutta="reference"))
    return fp.MultiBody(bodies, ["L", "R"])


def test_sweptwing_reference_meshes_lift(fp):
    body = _wing(fp)
    assert body.ncells == 2880 and body.nsheddings == 30
    aoa = 4.2
    U = 30.0 * np.array([np.cos(np.deg2rad(aoa)), 0.0, np.sin(np.deg2rad(aoa))])
    kw = dict(probe_offset=0.02, probe_length=B / AR, span_dir=(0, 1, 0), ref_area=B ** 2 / AR)
    cl, cd, _ = fp.lift_drag(body, U, **kw)
    clp, _, _ = fp.lift_drag(body, U, kutta="pair", **kw)
    cln, _, _ = fp.lift_drag(body, U, kutta="none", **kw)
    print("Kutta variants: reference %.4f, pair %.4f, none %.4f" % (cl, clp, cln), flush=True)
    g64 = body.strength.copy()
    cl2, cd2, st = fp.lift_drag(body, U, solver=fp.solve_gmres_ir, solver_optargs={"precond_block": 96}, **kw)
    print("swept wing (reference meshes, 2,880 panels): CL %.4f (exp. 0.238), CDi %.5f (exp. CD 0.005); "
          "GMRES-IR CL %.10f" % (cl, cd, cl2), flush=True)
    assert abs(cl - 0.238) / 0.238 <= 0.10
    assert abs(cl2 - cl) < 1e-9 * abs(cl)
    assert np.abs(body.strength - g64).max() < 1e-9 * np.abs(g64).max()
    # symmetric geometry and flow: the two halves carry mirrored lift
    nL = body.bodies[0].ncells
    assert abs(body.strength[:nL].sum() - body.strength[nL:].sum()) < 1e-6 * np.abs(body.strength).sum()


def test_sweptwing_surface_velocity_workflow(fp):
    """The reference example's post-processing chain on the same wing
    (examples/sweptwing.jl:120-150: calcfield_U, the combined calcfield_Ugradmu,
    addfields, calcfield_Cp, calcfield_F, then L and D), whose published result is
    CL 0.272, CD 0.0066 (docs/resources/data/sweptwing000-CLCD.md).  On the stored
    meshes this gives CL 0.2760 (+1.5 %) and CD 0.0084; the nodal averaging and
    neighbour search of GeometricTools are restated, not run, so the pins are loose:
    CL within 3 % and CD within 35 % of the published values.  The cell-centred
    and node-based schemes alone are printed beside it."""
    body = _wing(fp)
    aoa, mag, rho = 4.2, 30.0, 1.225
    U = mag * np.array([np.cos(np.deg2rad(aoa)), 0.0, np.sin(np.deg2rad(aoa))])
    D = np.tile((U / mag)[:, None], (1, body.nsheddings))
    fp.solve(body, np.tile(U[:, None], (1, body.ncells)), D, D)
    Dhat = U / mag
    Lhat = np.cross(Dhat, np.array([0.0, 1.0, 0.0]))
    qS = 0.5 * rho * mag ** 2 * B ** 2 / AR
    res = {}
    for name, fn in (("combined", fp.calcfield_Ugradmu), ("cell", fp.calcfield_Ugradmu_cell),
                     ("node", fp.calcfield_Ugradmu_node)):
        fp.calcfield_U(body, body)
        fn(body)
        fp.addfields(body, "Ugradmu", "U")
        fp.calcfield_Cp(body, mag)
        fp.calcfield_F(body, mag, rho)
        LDS = fp.calcfield_LDS(body, Lhat, Dhat)
        res[name] = (LDS[:, 0].dot(Lhat) / qS, LDS[:, 1].dot(Dhat) / qS)
    print("surface-velocity workflow: " + ", ".join("%s CL %.4f CD %.5f" % (k, *v) for k, v in res.items()),
          flush=True)
    cl, cd = res["combined"]
    assert abs(cl - 0.272) / 0.272 <= 0.03
    assert abs(cd - 0.0066) / 0.0066 <= 0.35
