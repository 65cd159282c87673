"""CPU-only checks of the host side: the C ABI library loads and exports every
symbol the header declares, host-only entry points behave, and the synthetic
geometry generators produce the workloads' shapes and conventions.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "flowpanel_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fp_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(fp):
    from flowpanel_amd import _native
    lib = ctypes.CDLL(_native.LIBPATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(_native.EXPORTS) == syms


def test_solve_opts_defaults(fp):
    from flowpanel_amd import _native as nat
    o = nat.opts(nat.FP_SOLVER_GMRES)
    assert abs(o.tol - np.sqrt(np.finfo(float).eps)) < 1e-20 and o.restart == 50
    o = nat.opts(nat.FP_SOLVER_GMRES_IR, tol=1e-11)
    assert o.tol == 1e-11 and o.precond_block == 512
    with pytest.raises(TypeError):
        nat.opts(nat.FP_SOLVER_LU, bogus=1)


def test_sphere_generator(geo):
    g = geo.sphere()
    assert g.ncells == 1152
    nv = g.nvertices()
    assert (nv == 3).sum() == 96 and (nv == 4).sum() == 1056
    # outward normals: (p2-p1) x (p3-p1) points away from the centre
    for c in range(g.ncells):
        p = g.nodes[:, g.cell(c)]
        n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        assert np.dot(n, p.mean(axis=1)) > 0


def test_wing_generator_and_shedding(geo):
    g, sh = geo.loft_wing(n_chord=60, n_span=80)
    assert g.ncells == 4800 and sh.shape == (6, 80)
    for e in range(sh.shape[1]):
        pu, pl = sh[0, e], sh[3, e]
        cu, cl = g.cell(pu), g.cell(pl)
        te_u = g.nodes[:, [cu[sh[1, e]], cu[sh[2, e]]]]
        te_l = g.nodes[:, [cl[sh[4, e]], cl[sh[5, e]]]]
        # the lower edge runs the other way along the same trailing edge
        assert np.abs(te_u[:, 0] - te_l[:, 1]).max() < 1e-12
        assert np.abs(te_u[:, 1] - te_l[:, 0]).max() < 1e-12
    # upper-surface normals point up, lower down (mid-span, mid-chord)
    j = 40
    up = g.nodes[:, g.cell(45 + 60 * j)]
    lo = g.nodes[:, g.cell(15 + 60 * j)]
    assert np.cross(up[:, 1] - up[:, 0], up[:, 2] - up[:, 0])[2] > 0
    assert np.cross(lo[:, 1] - lo[:, 0], lo[:, 2] - lo[:, 0])[2] < 0


def test_triangulate_keeps_te_pattern(geo):
    g, sh = geo.loft_wing(n_chord=8, n_span=3, span=2.0)
    gt, st = geo.triangulate(g, sh)
    assert gt.ncells == 2 * g.ncells
    for e in range(st.shape[1]):
        cu, cl = gt.cell(st[0, e]), gt.cell(st[3, e])
        a = gt.nodes[:, [cu[st[1, e]], cu[st[2, e]]]]
        b = gt.nodes[:, [cl[st[4, e]], cl[st[5, e]]]]
        assert np.abs(a[:, 0] - b[:, 1]).max() < 1e-12 and np.abs(a[:, 1] - b[:, 0]).max() < 1e-12
        # the Kutta neighbours pi+1 and pj-1 are the other halves of the same quads
        assert st[0, e] % 2 == 0 and st[3, e] % 2 == 1


def test_bwb_generator(geo):
    g, sh = geo.bwb()
    assert g.ncells == 30000 and sh.shape[1] == 120
    g2, _ = geo.bwb()
    assert np.array_equal(g.nodes, g2.nodes)  # deterministic


def test_naca_section(geo):
    c = geo.naca4("0012", 60)
    assert c.shape == (2, 61)
    assert np.allclose(c[:, 0], c[:, -1])          # sharp trailing edge
    assert abs(c[1].max() - 0.06) < 2e-3             # 12 % thick
    assert np.allclose(geo.direction(0.0), [1, 0, 0])


GOLDEN = os.path.join(ROOT, "tests", "golden")


def _sweptwing_te(side):
    """The trailing-edge lines of the reference's mesh test (test/runtests_meshes.jl:215-245)."""
    y = -1.2446 if side == "left" else 1.2446
    X1 = np.array([1.74244, y, 1.38094e-07])
    X2 = np.array([0.49784, -1.38178e-16 if side == "left" else 1.38178e-16, 1.38094e-07])
    te = np.stack([X1 + v * (X2 - X1) for v in np.linspace(0, 1, 1000)], axis=1)
    key = te[1] if side == "left" else -te[1]
    return te[:, np.argsort(key, kind="stable")]


def test_read_vtu_reference_meshes(fp):
    for side in ("left", "right"):
        g = fp.read_vtu(os.path.join(GOLDEN, "sweptwing-%s000.vtu" % side))
        assert g.nnodes == 768 and g.ncells == 1440
        assert (g.nvertices() == 3).all()
        # the tip trailing-edge node and the span of each half
        assert np.isclose(np.abs(g.nodes[1]).max(), 1.2446)
        assert np.allclose(g.nodes[0].max(), 1.74244)
        # a closed surface: every edge is shared by exactly two triangles (open only at the root and tip)
        edges = {}
        for c in range(g.ncells):
            cell = g.cell(c)
            for k in range(3):
                e = tuple(sorted((int(cell[k]), int(cell[(k + 1) % 3]))))
                edges[e] = edges.get(e, 0) + 1
        assert max(edges.values()) == 2


def test_calc_shedding_reference_meshes(fp):
    for side in ("left", "right"):
        g = fp.read_vtu(os.path.join(GOLDEN, "sweptwing-%s000.vtu" % side))
        te = _sweptwing_te(side)
        span = np.ptp(te[1])
        sh = fp.calc_shedding(g, te, tolerance=0.001 * span)
        assert sh.shape == (6, 15)
        for e in range(sh.shape[1]):
            pi, nia, nib, pj, nja, njb = sh[:, e]
            ci, cj = g.cell(pi), g.cell(pj)
            # the two panels share the edge with opposite orientation
            assert ci[nia] == cj[njb] and ci[nib] == cj[nja]
            assert (list(ci).index(ci[nib]) + 1) % 3 == nia
            # both end points lie on the trailing edge line
            for v in (ci[nia], ci[nib]):
                p = g.nodes[:, v]
                d = np.cross(te[:, -1] - te[:, 0], p - te[:, 0])
                assert np.linalg.norm(d) / np.linalg.norm(te[:, -1] - te[:, 0]) < 1e-6
        # consecutive shedding edges are chained along the edge
        for e in range(sh.shape[1] - 1):
            assert g.cell(sh[0, e])[sh[1, e]] == g.cell(sh[0, e + 1])[sh[2, e + 1]]


def test_read_msh_synthetic(fp, tmp_path):
    txt = """$MeshFormat
4.1 0 8
$EndMeshFormat
$Nodes
1 4 1 4
2 1 0 4
10
11
12
13
0 0 0
1 0 0
1 1 0
0 1 0
$EndNodes
$Elements
2 3 1 3
1 1 1 1
7 10 11
2 1 2 2
5 10 11 12
6 10 12 13
$EndElements
"""
    p = tmp_path / "two.msh"
    p.write_text(txt)
    g = fp.read_msh(str(p))
    assert g.nnodes == 4 and g.ncells == 2
    assert list(g.cell(0)) == [0, 1, 2] and list(g.cell(1)) == [0, 2, 3]


def test_multisection_spacing_and_airfoil_rediscretisation(fp):
    g = fp.geometry
    nd = [(0.25, 10, 10.0, False), (0.50, 10, 1.0, True), (0.25, 10, 0.1, False)]
    u = g._sections_spacing(nd)
    assert len(u) == 31 and u[0] == 0.0 and abs(u[-1] - 1.0) < 1e-15 and np.all(np.diff(u) > 0)
    d = np.diff(u)
    # section 1 grows by 10 from its first to its last cell, section 3 shrinks by 10, the middle is uniform
    assert abs(d[9] / d[0] - 10.0) < 1e-9 and abs(d[29] / d[20] - 0.1) < 1e-9
    assert np.allclose(d[10:20], d[10]) and abs(u[10] - 0.25) < 1e-12 and abs(u[20] - 0.75) < 1e-12
    # a circle through the rediscretisation stays on the circle and keeps its end points
    t = np.linspace(0, 2 * np.pi, 81)
    x, y = 0.5 + 0.5 * np.cos(t), 0.5 * np.sin(t)
    xs, ys = g.rediscretize_airfoil(x, y, nd, nd)
    assert len(xs) == 61 and abs(xs[0] - 1.0) < 1e-12 and abs(xs[-1] - 1.0) < 1e-12
    assert np.abs(np.hypot(xs - 0.5, ys) - 0.5).max() < 1e-4
    assert abs(xs.min()) < 1e-12                               # the leading edge is kept


def test_revolution_liftbody_is_closed_and_outward(fp, orc):
    g = fp.geometry
    t = np.linspace(0, 2 * np.pi, 41)
    pts = np.c_[0.5 + 0.5 * np.cos(t), 1.0 + 0.1 * np.sin(t)]       # a torus-like duct section
    grid, sh = g.revolution_liftbody(pts, 24)
    tg, tsh = g.triangulate(grid, sh)
    for gg in (grid, tg):
        n, c, a = orc.geometry(gg.nodes, gg.cell_ptr, gg.cell_idx)
        # closed and outward: the vector areas cancel, sum(n . x A) = 3 V > 0 (Pappus: V = 2 pi R (pi a b) for the
        # elliptic section of semi-axes a = 0.5, b = 0.1 at radius R = 1; the polygonal surface falls a little short)
        assert np.abs((n * a).sum(axis=1)).max() < 1e-12
        vol = (np.einsum("ij,ij->j", n, c) * a).sum() / 3
        exact = 2 * np.pi ** 2 * 1.0 * 0.5 * 0.1
        assert 0.95 * exact < vol < exact
    # every azimuthal strip sheds: coincident trailing-edge nodes of the first and last contour cells
    assert sh.shape[1] == 24 and tsh.shape[1] == 24
    X = tg.nodes
    for pi, nia, nib, pj, nja, njb in tsh.T:
        ci, cj = tg.cell(pi), tg.cell(pj)
        assert np.allclose(X[:, ci[nia]], X[:, cj[njb]]) and np.allclose(X[:, ci[nib]], X[:, cj[nja]])
