"""Block-Jacobi inverses started beside the assembly.

For GMRES-IR on one vortex-ring body with index-range blocks, the library
assembles the AIC's diagonal blocks on their own (ring and wake terms, FP64
rounded to FP32, the last block padded with the identity) and inverts them on a
stream of their own while the whole matrix is assembled; the coarse space
follows once the matrix is whole.  The solution must match the oracle's LU
(1e-10 of max|Gamma|) and the solve with the inverses taken from the assembled
matrix (the default; FLOWPANEL_BJ_EARLY=1 selects the early set-up), with the
iteration counts within one of each
other, on quads with a wake, triangles with a wake and a ragged last block,
blocks in registers (<= 256 rows) and blocked Gauss-Jordan (> 256 rows), with and
without the coarse space, for the matrix-free and the stored FP64 residual and
for several right-hand sides.
"""
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def _obody(b):
    return types.SimpleNamespace(nodes=b.grid.nodes, cell_ptr=b.grid.cell_ptr, cell_idx=b.grid.cell_idx,
                                 ncells=b.ncells, element="vortexring", shedding=b.shedding,
                                 nsheddings=b.nsheddings, kerneloffset=b.kerneloffset,
                                 kernelcutoff=b.kernelcutoff, CPoffset=b.CPoffset, charlen_kind=0)


def _body(fp, geo, case):
    if case == "quads":
        g, sh = geo.loft_wing(n_chord=48, n_span=30, span=6.0, sweep_deg=20.0)       # 1,440 panels
    else:
        g, sh = geo.triangulate(*geo.loft_wing(n_chord=34, n_span=22, span=6.0))    # 1,496 panels
    return fp.RigidWakeBody(g, sh)


@pytest.mark.parametrize("case,block,coarse,resid", [
    ("quads", 200, 0, "matrix_free"),
    ("quads", 512, 60, "stored"),
    ("triangles", 300, 0, "stored"),
    ("triangles", 128, 44, "matrix_free"),
])
def test_early_block_jacobi_matches_oracle_and_late(fp, geo, orc, case, block, coarse, resid, monkeypatch):
    body = _body(fp, geo, case)
    V = fp.direction(5.0)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings))
    opts = {"precond_block": block, "coarse_block": coarse, "tol": 1e-12,
            "residual_mode": fp._native.FP_RESID_STORED if resid == "stored" else fp._native.FP_RESID_MATRIX_FREE}
    print("case %s, %d panels, block %d, coarse %d, %s" % (case, body.ncells, block, coarse, resid), flush=True)
    monkeypatch.setenv("FLOWPANEL_BJ_EARLY", "1")
    st_e = fp.solve(body, U, D, D, solver=fp.solve_gmres_ir, solver_optargs=opts)
    x_e = body.strength.copy()
    monkeypatch.setenv("FLOWPANEL_BJ_EARLY", "0")
    st_l = fp.solve(body, U, D, D, solver=fp.solve_gmres_ir, solver_optargs=opts)
    x_l = body.strength.copy()
    Go, normals, _ = orc.assemble([_obody(body)], [D], [D])
    x_o = orc.solve_lu(Go, orc.rhs(U, normals))
    print("  early: %s\n  late:  %s\n  vs oracle %.1e / %.1e" % (st_e, st_l, rel(x_e, x_o), rel(x_l, x_o)), flush=True)
    assert rel(x_e, x_o) < 1e-10 and rel(x_l, x_o) < 1e-10
    assert st_e["residual"] < 1e-11
    assert abs(st_e["iterations"] - st_l["iterations"]) <= max(1, st_l["refinements"])


def test_early_block_jacobi_several_freestreams(fp, geo, orc, monkeypatch):
    monkeypatch.setenv("FLOWPANEL_BJ_EARLY", "1")
    body = _body(fp, geo, "quads")
    aoas = (0.0, 4.0, 8.0)
    Us = [np.tile(fp.direction(a)[:, None], (1, body.ncells)) for a in aoas]
    D = np.tile(fp.direction(4.0)[:, None], (1, body.nsheddings))
    opts = {"precond_block": 256, "coarse_block": 48, "tol": 1e-12}
    X, st = fp.solve_freestreams(body, np.stack(Us, axis=2), D, D, solver=fp.solve_gmres_ir, solver_optargs=opts)
    print("  stats %s" % (st,), flush=True)
    Go, normals, _ = orc.assemble([_obody(body)], [D], [D])
    for c, U in enumerate(Us):
        assert rel(X[:, c], orc.solve_lu(Go, orc.rhs(U, normals))) < 1e-10
