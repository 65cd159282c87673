"""The reference's blended-wing-body example end to end on the GPU.

examples/blendedwingbody.jl: the Gmsh half mesh of the ZEROe BWB
(examples/data/zeroebwb.msh, 11,125 triangles) translated, rotated by
RotZ(-90 deg) RotX(90 deg), scaled by 2e-3 and mirrored across y = 0 (22,250
panels); its trailing edge (zeroebwb-TE.msh) sorted along the span and found
within 0.001 of the span; RigidWakeBody{VortexRing, 2} (least squares) at
AoA 10 deg, 30 m/s; surface velocity U + U_gradmu, Cp clipped at
1 - 342/30, forces and CL/CD on Sref = 3.23^2 / 8.  The reference publishes no
CL for it (docs/src/examples/blendedwingbody-aero.md:207-208 gives the panel
count and ~10 s "with GPU"), so the checks are: the FP64 LU and the mixed
GMRES-IR least-squares solutions (the latter to a 3e-13 residual, near the FP64 floor of this system: the LU leaves 5e-13; its condition is about 5e6)
give the same Cp (5e-5) and CL, CD (1e-6 of CL), the panel order does not change
the answer (a Morton-ordered copy of the mesh, mapped back, 1e-9 on the strengths
and 1e-7 on Cp away from the cells that force_cellTE picks by index, pi+1 and pj-1,
which are other cells in the other order), the lift is positive and the drag small.  GMRES-IR runs with its block-Jacobi blocks and
coarse chunks along the library's spatial order (FP_ORDER_SPATIAL) on the mesh
as read, and along index ranges on the Morton-ordered copy; with index ranges
on the mesh as read, the preconditioner groups panels that are not neighbours
and GMRES stalls (0.09-0.16 relative residual after 300 iterations in a numpy
restatement), which is what the spatial order is for.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rot():
    c, s = np.cos(np.pi / 2), np.sin(np.pi / 2)
    Rx = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    Rz = np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])      # RotZ(-90 deg)
    return Rz @ Rx


def _zeroebwb(fp, order=False):
    g = fp.read_msh(os.path.join(GOLDEN, "zeroebwb.msh"))
    te = fp.read_msh_nodes(os.path.join(GOLDEN, "zeroebwb-TE.msh"))
    R = _rot()
    g = fp.geometry.transform(g, R, scaling=2e-3)
    te = 2e-3 * (R @ te)
    g = fp.geometry.mirror(g, 1)
    te = te[:, np.argsort(te[1], kind="stable")]
    span = te[1].max() - te[1].min()
    sh = fp.calc_shedding(g, te, tolerance=0.001 * span)
    perm = None
    if order:
        g, sh, perm = fp.geometry.spatial_order(g, sh)
    return g, sh, perm


def _aero(fp, body, Vinf, magV, rho):
    fp.calcfield_U(body, body)
    fp.calcfield_Ugradmu(body)
    fp.addfields(body, "Ugradmu", "U")
    fp.calcfield_Cp(body, magV, clip=lambda cp: max(1 - 342.0 / magV, cp))
    fp.calcfield_F(body, magV, rho)
    Dhat = Vinf / magV
    Lhat = np.cross(Dhat, [0.0, 1.0, 0.0])
    LDS = fp.calcfield_LDS(body, Lhat, Dhat)
    nondim = 0.5 * rho * magV ** 2 * 3.23 ** 2 / 8.0
    CL = np.sign(LDS[:, 0].dot(Lhat)) * np.linalg.norm(LDS[:, 0]) / nondim
    CD = np.sign(LDS[:, 1].dot(Dhat)) * np.linalg.norm(LDS[:, 1]) / nondim
    return CL, CD


def test_zeroebwb_example_least_squares(fp):
    aoa, magV, rho = 10.0, 30.0, 1.225
    Vinf = magV * np.array([np.cos(np.deg2rad(aoa)), 0.0, np.sin(np.deg2rad(aoa))])
    krylov = {"precond_block": 1024, "coarse_block": 50, "tol": 3e-13, "restart": 100}
    res = {}
    index_neighbours = []
    for order in (False, True):
        g, sh, perm = _zeroebwb(fp, order)
        index_neighbours.append((sh, perm))
        body = fp.RigidWakeBody(g, sh, formulation=2, CPoffset=1e-14)
        assert body.ncells == 22250 and body.nsheddings > 100
        U = np.tile(Vinf[:, None], (1, body.ncells))
        D = np.tile((Vinf / magV)[:, None], (1, body.nsheddings))
        # the same physical panel prescribed in both orders (the least-squares null mode is a uniform strength)
        pres = [(0, 0.0)] if perm is None else [(int(np.nonzero(perm == 0)[0][0]), 0.0)]
        # the mesh's own order: preconditioner blocks taken along the library's spatial order;
        # the Morton-ordered mesh: its index ranges already are compact patches
        runs = (("solve_ludiv", {}),
                ("solve_gmres_ir", dict(krylov, precond_order="index" if order else "spatial")))
        out = {}
        for solver, opts in runs:
            t = time.perf_counter()
            st = fp.solve(body, U, D, D, solver=getattr(fp, solver), solver_optargs=opts, elprescribe=pres)
            dt = time.perf_counter() - t
            x = body.strength.copy()
            if perm is not None:
                back = np.empty_like(x)
                back[perm] = x
                x = back
            CL, CD = _aero(fp, body, Vinf, magV, rho)
            cp = np.asarray(body.get_field("Cp")).copy()
            if perm is not None:
                back = np.empty_like(cp)
                back[perm] = cp
                cp = back
            out[solver] = (x, CL, CD, st, cp)
            print("zeroebwb %s mesh, %s %s: %.3f s, CL %.4f, CD %.5f, iterations %s, residual %.1e" %
                  ("Morton-ordered" if order else "as read", solver, opts.get("precond_order", ""), dt, CL, CD,
                   st.get("iterations"), st.get("residual")), flush=True)
        # the bordered least-squares system of this mesh is ill-conditioned (a residual of 1e-10 left
        # strength errors near 4e-4, condition about 5e6), so GMRES-IR is driven to a 3e-13 residual and
        # the solvers are compared on what the strengths produce: Cp at the control points, CL and CD
        (xl, CLl, CDl, _, cpl), (xg, CLg, CDg, _, cpg) = out["solve_ludiv"], out["solve_gmres_ir"]
        print("  LU vs GMRES-IR: max|dGamma|/max|Gamma| %.1e, max|dCp| %.1e, dCL %.1e, dCD %.1e" %
              (np.abs(xl - xg).max() / np.abs(xl).max(), np.abs(cpl - cpg).max(), CLg - CLl, CDg - CDl), flush=True)
        assert np.abs(cpl - cpg).max() < 5e-5
        assert abs(CLg - CLl) < 1e-6 * abs(CLl) and abs(CDg - CDl) < 1e-6 * abs(CLl)
        res[order] = out
    x0, x1 = res[False]["solve_ludiv"][0], res[True]["solve_ludiv"][0]
    # two FP64 LU solves of the same system in two row/column orders: they differ by rounding
    # amplified by the condition number (~5e6 above): measured 1.4e-9
    assert np.abs(x0 - x1).max() / np.abs(x0).max() < 1e-8
    # force_cellTE keeps the cell-centred gradient on the shedding panels and on their index
    # neighbours pi+1, pj-1 (src/FLOWPanel_postprocess.jl:875-882); on an unstructured mesh those
    # neighbours are other cells in the other order, so Cp is compared everywhere else
    keep = np.ones(len(x0), bool)
    for sh, perm in index_neighbours:
        cells = np.concatenate([sh[0] + 1, sh[3][sh[3] >= 0] - 1])
        cells = cells[(cells >= 0) & (cells < len(x0))]
        keep[cells if perm is None else perm[cells]] = False
    print("  Morton vs as read: max|dCp| %.1e off the %d index-neighbour cells" %
          (np.abs(res[False]["solve_ludiv"][4] - res[True]["solve_ludiv"][4])[keep].max(), (~keep).sum()), flush=True)
    assert (~keep).sum() <= 4 * len(sh[0])
    # the two orders are two LU factorisations of a system of condition about 5e6: strengths agree to
    # 1e-9 of their maximum and Cp (|Cp| up to about 10) to 1e-7
    assert np.abs(res[False]["solve_ludiv"][4] - res[True]["solve_ludiv"][4])[keep].max() < 1e-7
    CL, CD = res[False]["solve_ludiv"][1:3]
    assert 0.2 < CL < 2.0 and abs(CD) < 0.2 * CL


def test_spatial_order_needs_bodies(fp):
    # solve_mat has no control points: a spatial preconditioner order is an argument error
    A = np.eye(50) * 4.0 + np.random.default_rng(0).normal(size=(50, 50)) * 0.1
    y = np.zeros(50)
    with pytest.raises(fp.FPError) as e:
        fp.solve_gmres_ir(y, A, np.ones(50), precond_block=16, precond_order="spatial")
    assert e.value.code == 1 and "control points" in str(e.value)
