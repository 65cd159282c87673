"""Body-level pins of the oracle (CPU only): closed forms and invariants.

* sphere of constant-source panels: Cp -> 1 - 9/4 sin^2(theta) (examples/sphere.jl:86-87)
  with the error shrinking under refinement, and zero net source strength;
* a watertight vortex-ring body: a uniform ring strength induces no velocity
  (the singular mode noted at src/FLOWPanel_abstractliftingbody.jl:113-121);
* assembly: column j equals n . U_vortexring(panel j) at every control point;
  the wake only touches the shedding columns;
* LU solve agrees with numpy; Kutta/force/LDS small cases by hand.
"""
import types

import numpy as np
import pytest


def obody(grid, element, shedding=None):
    sh = np.zeros((6, 0), np.int64) if shedding is None else shedding
    return types.SimpleNamespace(nodes=grid.nodes, cell_ptr=grid.cell_ptr, cell_idx=grid.cell_idx,
                                 ncells=grid.ncells, element=element, shedding=sh, nsheddings=sh.shape[1],
                                 kerneloffset=1e-8, kernelcutoff=1e-14, CPoffset=1e-14, charlen_kind=0)


def sphere_cp_error(orc, geo, nm, na):
    b = obody(geo.sphere(n_meridional=nm, n_azimuthal=na), "source")
    G, normals, cps = orc.assemble(b)
    U = np.tile([[1.0], [0.0], [0.0]], (1, b.ncells))
    sig = orc.solve_lu(G, orc.rhs(U, normals))
    Us = U + orc.uind(b, sig, cps)
    cp = orc.cp(Us, 1.0)
    x = cps[0] / np.linalg.norm(cps, axis=0)
    ana = 1 - 9 / 4 * (1 - x ** 2)
    _, _, areas = orc.geometry(b.nodes, b.cell_ptr, b.cell_idx)
    return np.abs(cp - ana).max(), abs(np.dot(sig, areas)) / np.abs(sig * areas).sum()


def test_sphere_closed_form_and_refinement(orc, geo):
    e1, f1 = sphere_cp_error(orc, geo, 12, 24)
    e2, f2 = sphere_cp_error(orc, geo, 24, 48)
    assert e2 < 0.02
    assert e2 < 0.6 * e1          # converges under refinement
    assert f1 < 1e-6 and f2 < 1e-6  # closed body: no net source


def test_watertight_vortex_ring_null_mode(orc, geo):
    b = obody(geo.sphere(n_meridional=10, n_azimuthal=16), "vortexring")
    G, normals, cps = orc.assemble([b], [None], [None])
    # a uniform ring strength on a closed surface cancels edge by edge
    assert np.abs(G @ np.ones(b.ncells)).max() < 1e-10 * np.abs(G).max()
    T = np.asfortranarray(np.array([[3.0, 0.2, 0.1], [0.1, 0.2, 0.3], [-2.0, 1.0, 1.5]]).T)
    U = orc.uind(b, np.ones(b.ncells), T)
    assert np.abs(U).max() < 1e-12


def test_assembly_columns_and_wake(orc, geo):
    g, sh = geo.loft_wing(n_chord=10, n_span=4, span=3.0)
    b = obody(g, "vortexring", sh)
    V = np.array([1.0, 0.0, 0.0])
    D = np.tile(V[:, None], (1, b.nsheddings))
    G, normals, cps = orc.assemble([b], [D], [D])
    b0 = obody(g, "vortexring")
    G0, _, _ = orc.assemble([b0], [None], [None])
    for j in range(b.ncells):
        col = np.zeros(b.ncells)
        orc.U_vortexring(g.nodes, g.cell(j), 1.0, cps, col, dot_with=normals)
        assert np.abs(G0[:, j] - col).max() == 0.0
    te = set(sh[0]) | set(sh[3])
    diff = np.abs(G - G0).max(axis=0)
    assert all((diff[j] > 0) == (j in te) for j in range(b.ncells))


def test_solve_matches_numpy(orc, geo):
    g, sh = geo.loft_wing(n_chord=12, n_span=6, span=4.0)
    b = obody(g, "vortexring", sh)
    V = geo.direction(5.0)
    D = np.tile(V[:, None], (1, b.nsheddings))
    G, normals, cps = orc.assemble([b], [D], [D])
    U = np.tile(V[:, None], (1, b.ncells))
    rhs = orc.rhs(U, normals)
    assert np.abs(orc.solve_lu(G, rhs) - np.linalg.solve(G, rhs)).max() < 1e-10
    assert np.abs(rhs + np.einsum("ij,ij->j", U, normals)).max() == 0


def test_kutta_cp_reference_indexing(orc):
    Us = np.zeros((3, 6))
    Us[0] = [0.0, 1.0, 2.0, 0.5, 0.1, 0.3]
    sh = np.array([[4], [2], [1], [1], [0], [2]])  # pi=4, pj=1
    cp = orc.cp(Us, 2.0, sh, kutta=True)
    raw = 1 - (Us[0] / 2.0) ** 2
    ave = (raw[4] + raw[5] + raw[1] + raw[0]) / 4
    assert np.allclose(cp[[0, 1, 4, 5]], ave)
    assert np.allclose(cp[[2, 3]], raw[[2, 3]])
    one = orc.cp(Us, 2.0, np.array([[4], [2], [1], [-1], [0], [0]]), kutta=True)
    assert np.allclose(one[[4, 5]], (raw[4] + raw[5]) / 2)


def test_forces_and_lds(orc):
    areas = np.array([2.0, 3.0])
    normals = np.asfortranarray(np.array([[0, 0, 1.0], [0, 1.0, 0]]).T)
    cps = np.array([0.5, -1.0])
    F = orc.forces(areas, normals, cps, 2.0, 1.2)
    q = 0.5 * 1.2 * 4.0
    assert np.allclose(F[:, 0], [0, 0, -0.5 * q * 2])
    assert np.allclose(F[:, 1], [0, 1.0 * q * 3, 0])
    L = orc.lds(F, [0, 0, 1.0], [1.0, 0, 0])
    assert np.allclose(L[:, 0], [0, 0, F[2].sum()])
    assert np.allclose(L[:, 2], np.cross([0, 0, 1.0], [1.0, 0, 0]) * F[1].sum())


def _cells(*cells):
    ptr = np.cumsum([0] + [len(c) for c in cells]).astype(np.int64)
    return ptr, np.concatenate([np.asarray(c, np.int64) for c in cells])


def test_panel_area_definition(orc):
    """Panel area = |sum of the fan cross products| / 2, the magnitude of the vector area.

    Reference panels are triangles (every body holds a gt.GridTriangleSurface,
    src/FLOWPanel_abstractbody.jl:905-919), where all definitions agree with
    |(p2-p1) x (p3-p1)|/2.  A polygon of more vertices takes the magnitude of its
    vector area, which depends only on its boundary; a planted mistake (the sum of
    the triangle magnitudes, or a dropped fan triangle) fails one of the cases.
    """
    h = 0.6
    nodes = np.asfortranarray(np.array([
        [0, 0, 0], [1, 0, 0], [1, 1, h], [0, 1, 0],                          # non-planar quad
        [0, 0, 5], [2, 0, 5], [2.5, 1.5, 5], [1, 2.5, 5], [-0.5, 1, 5],       # planar pentagon (z = 5)
        [3, 0, 0], [4, 0.5, 1], [3.2, 2, -1],                                 # triangle
    ], float).T)
    ptr, idx = _cells([0, 1, 2, 3], [4, 5, 6, 7, 8], [9, 10, 11])
    _, _, areas = orc.geometry(nodes, ptr, idx)
    # non-planar quad: fan vectors (0,-h,1)/2 + (-h,0,1)/2 -> |.| = sqrt(1 + h^2/2)
    assert abs(areas[0] - np.sqrt(1 + h * h / 2)) < 1e-15
    assert abs(areas[0] - np.sqrt(1 + h * h)) > 0.05      # the sum of triangle magnitudes differs
    # planar pentagon: shoelace formula in its plane
    x, y = nodes[0, 4:9], nodes[1, 4:9]
    shoelace = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
    assert abs(areas[1] - shoelace) < 1e-14
    # triangle: |cross| / 2
    p = nodes[:, 9:12]
    assert abs(areas[2] - 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]))) < 1e-15
    # a quad made of the reference's two triangles (p1,p2,p3), (p1,p3,p4) carries, at
    # one pressure, the force of its vector area: A n summed over the pair
    t1 = 0.5 * np.cross(nodes[:, 1] - nodes[:, 0], nodes[:, 2] - nodes[:, 0])
    t2 = 0.5 * np.cross(nodes[:, 2] - nodes[:, 0], nodes[:, 3] - nodes[:, 0])
    assert abs(np.linalg.norm(t1 + t2) - areas[0]) < 1e-15


def test_moment_and_sectional_force_hand_cases(orc):
    # one force at a known arm: M = r x F
    X = np.array([[1.0], [2.0], [0.5]])
    F = np.array([[0.0], [0.0], [3.0]])
    assert np.allclose(orc.mtot(X, F, [1.0, 0.0, 0.0]), np.cross([0.0, 2.0, 0.5], [0.0, 0.0, 3.0]))
    # a pure couple: equal and opposite forces give the same moment about any point
    X = np.array([[0.0, 1.0], [0.0, 0.0], [0.0, 0.0]])
    F = np.array([[0.0, 0.0], [1.0, -1.0], [0.0, 0.0]])
    assert np.allclose(orc.mtot(X, F, [0, 0, 0]), orc.mtot(X, F, [3.0, -2.0, 7.0]))
    assert np.allclose(orc.mtot(X, F, [0, 0, 0]), [0, 0, -1.0])
    # uniform load w per unit span on stations 0.5 apart: every station reads w per unit span
    nchord, nspan, w, dy = 4, 6, 2.5, 0.5
    Xc = np.zeros((3, nchord * nspan))
    Fc = np.zeros((3, nchord * nspan))
    for j in range(nspan):
        for i in range(nchord):
            c = i + nchord * j
            Xc[:, c] = [0.1 * i, dy * j, 0.0]
            Fc[2, c] = w * dy / nchord
    f, pos = orc.sectionalforce(Fc, Xc, nchord, nspan, [0.0, 1.0, 0.0])
    assert np.allclose(pos, dy * np.arange(nspan))
    assert np.allclose(f[2], w) and np.allclose(f[:2], 0.0)
    # a mistaken spacing (the full distance between neighbours) halves the interior stations
    assert not np.allclose(f[2] / 2, w)
