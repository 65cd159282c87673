"""The shared-term vortex-ring assemblies on the device.

Triangle and quad bodies assemble their AIC with the kernel that shares each
mesh node's and mesh edge's terms between the block's panels through LDS (the
default) or with the edge-pair kernel (each lane's path of two mesh edges
evaluated in registers, only the edge terms through LDS; FLOWPANEL_ASSEMBLY=pairs)
(src/FLOWPanel_liftingbody.jl:455-572 restated edge by edge).  Both are checked against the oracle (1e-12 of max|G| in FP64,
2e-5 in FP32 storage) and against the panel-wise kernel (FLOWPANEL_ASSEMBLY=panel,
1e-13) on quads with a wake,
triangles, a quad/triangle mix, a sphere with collapsed pole edges, disjoint
panels (node-limited blocks), a reduced blended-wing-body loft and a
two-body column block (a wing beside a second vortex-ring body).
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


def _mixed(fp, geo):
    g, sh = geo.loft_wing(n_chord=12, n_span=10)
    cells = []
    for k, c in enumerate(g.cell_idx.reshape(-1, 4)):
        cells += [c[[0, 1, 2]], c[[0, 2, 3]]] if k % 3 == 1 else [c]
    return fp.Grid(g.nodes, cells)


def _disjoint(fp, n=300, seed=1):
    rng = np.random.default_rng(seed)
    nodes, cells = [], []
    for k in range(n):
        c = rng.uniform(-3, 3, 3)
        u, v = rng.normal(size=3), rng.normal(size=3)
        nodes += [c, c + 0.1 * u, c + 0.1 * (u + v), c + 0.1 * v]
        cells.append(np.arange(4 * k, 4 * k + 4))
    return fp.Grid(np.array(nodes).T, cells)


def _body(fp, geo, case):
    if case == "quads_wake":
        g, sh = geo.loft_wing(n_chord=16, n_span=12)
        return fp.RigidWakeBody(g, sh)
    if case == "triangles_wake":
        g, sh = geo.triangulate(*geo.loft_wing(n_chord=16, n_span=10))
        return fp.RigidWakeBody(g, sh)
    if case == "bwb":
        g, sh = geo.bwb(n_chord=50, n_span=24)
        return fp.RigidWakeBody(g, sh)
    if case == "mixed":
        return fp.RigidWakeBody(_mixed(fp, geo), None)
    if case == "sphere":
        return fp.RigidWakeBody(geo.sphere(n_meridional=8, n_azimuthal=12), None)
    return fp.RigidWakeBody(_disjoint(fp), None)


@pytest.mark.parametrize("case", ["quads_wake", "triangles_wake", "bwb", "mixed", "sphere", "disjoint"])
def test_shared_edge_assembly_matches_oracle_and_panelwise(fp, geo, orc, case, monkeypatch):
    body = _body(fp, geo, case)
    V = fp.direction(4.0)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings)) if body.nsheddings else None
    fp.solvers.set_freestream(body, U, D, D)
    print("case %s: %d panels" % (case, body.ncells), flush=True)
    G = fp.assemble(body, "f64")
    G32 = fp.assemble(body, "f32")
    Go, _, _ = orc.assemble([_obody(body)], [D], [D])
    monkeypatch.setenv("FLOWPANEL_ASSEMBLY", "pairs")
    Gn = fp.assemble(body, "f64")
    monkeypatch.setenv("FLOWPANEL_ASSEMBLY", "panel")
    Gp = fp.assemble(body, "f64")
    print("  vs oracle %.1e, vs panel-wise %.1e, vs edge pairs %.1e, f32 %.1e" %
          (rel(G, Go), rel(G, Gp), rel(G, Gn), rel(G32, Go)), flush=True)
    assert rel(G, Go) < 1e-12 and rel(Gn, Go) < 1e-12
    assert rel(G, Gp) < 1e-13 and rel(Gn, Gp) < 1e-13
    assert rel(G32, Go) < 2e-5


def test_two_vortex_ring_bodies_column_blocks(fp, geo, orc):
    g1, s1 = geo.loft_wing(n_chord=12, n_span=8, span=3.0)
    g2, s2 = geo.triangulate(*geo.loft_wing(n_chord=10, n_span=6, span=2.0))
    g2.nodes[2] -= 1.0
    w1, w2 = fp.RigidWakeBody(g1, s1), fp.RigidWakeBody(g2, s2)
    mb = fp.MultiBody([w1, w2], ["upper", "lower"])
    V = fp.direction(3.0)
    U = np.tile(V[:, None], (1, mb.ncells))
    D = np.tile(V[:, None], (1, mb.nsheddings))
    fp.solve(mb, U, D, D)
    D1, D2 = D[:, :w1.nsheddings], D[:, w1.nsheddings:]
    Go, normals, _ = orc.assemble([_obody(w1), _obody(w2)], [D1, D2], [D1, D2])
    G = fp.assemble(mb, "f64")
    assert rel(G, Go) < 1e-12
    x = orc.solve_lu(Go, orc.rhs(U, normals))
    assert rel(mb.strength, x) < 1e-10
