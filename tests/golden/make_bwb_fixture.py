"""Write the oracle's FP64 solution of the 30,000-panel blended-wing-body workload.

The workload: ``geometry.bwb(seed=5)``, 250 chordwise x 120 spanwise vortex-ring
quads with a rigid wake along the freestream, AoA 5 deg, U = 1.  The oracle
(oracle/fp_oracle.c: `_G_Uvortexring!` restatement, OpenMP over columns)
assembles the 30,000 x 30,000 FP64 AIC, LAPACK getrf/getrs solves it
(`solve_ludiv!`), and the oracle's `Uind!` gives the surface velocity and
Cp = 1 - |U|^2 (no Kutta averaging) at the control points.

Output ``bwb30k_oracle.npz``: strengths ``gamma`` (30,000), ``cp`` (30,000),
the 2-norm relative residual of the oracle's own solve, and a SHA-256 of the
generated node coordinates so that a test can tell whether the generator still
builds the same body.  Runs on the CPU in a few minutes with ~16 GB of memory:

    python tests/golden/make_bwb_fixture.py [n_chord n_span]
"""
import hashlib
import os
import sys
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def node_hash(nodes):
    return hashlib.sha256(np.ascontiguousarray(nodes, dtype=np.float64).tobytes()).hexdigest()


def main():
    nc, ns = (int(sys.argv[1]), int(sys.argv[2])) if len(sys.argv) > 2 else (250, 120)
    import __graft_entry__
    fp = __graft_entry__.load_package()
    from oracle import oracle as orc
    import scipy.linalg as sla
    grid, sh = fp.geometry.bwb(seed=5, n_chord=nc, n_span=ns)
    N = grid.ncells
    b = types.SimpleNamespace(nodes=grid.nodes, cell_ptr=grid.cell_ptr, cell_idx=grid.cell_idx, ncells=N,
                              element="vortexring", shedding=sh, nsheddings=sh.shape[1], kerneloffset=1e-8,
                              kernelcutoff=1e-14, CPoffset=1e-14, charlen_kind=0)
    V = fp.direction(5.0)
    U = np.tile(V[:, None], (1, N))
    D = np.tile(V[:, None], (1, sh.shape[1]))
    t = time.time()
    G, normals, cps = orc.assemble([b], [D], [D])
    print("assembly %.1f s" % (time.time() - t), flush=True)
    rhs = orc.rhs(U, normals)
    t = time.time()
    lu = sla.lu_factor(G, overwrite_a=False, check_finite=False)
    gam = sla.lu_solve(lu, rhs, check_finite=False)
    del lu
    print("LU %.1f s" % (time.time() - t), flush=True)
    res = float(np.linalg.norm(G @ gam - rhs) / np.linalg.norm(rhs))
    del G
    t = time.time()
    Us = U + orc.uind(b, gam, cps, D, D)
    cp = orc.cp(Us, 1.0)
    print("Uind %.1f s, residual %.3e" % (time.time() - t, res), flush=True)
    out = os.path.join(HERE, "bwb%dx%d_oracle.npz" % (nc, ns) if (nc, ns) != (250, 120) else "bwb30k_oracle.npz")
    np.savez_compressed(out, gamma=gam, cp=cp, residual=np.array(res), n_chord=np.array(nc),
                        n_span=np.array(ns), node_sha256=np.array(node_hash(grid.nodes)))
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
