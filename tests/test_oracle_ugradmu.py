"""Pins of the oracle's cell-centred Green-Gauss -1/2 grad(mu)
(src/FLOWPanel_postprocess.jl:139-497), CPU only:

* a linear strength on a uniform planar quad grid: the Green-Gauss gradient of
  interior cells is exact, so the velocity is exactly -1/2 grad(Gamma);
* the sign follows the sign of CPoffset (normals pointing inward);
* a uniform strength on a closed body gives zero (the edge normals of every
  planar panel sum to zero);
* trailing-edge panels lose all their neighbour terms (RigidWakeBody variant).
"""
import types

import numpy as np


def _planar(nx=6, ny=5, h=0.3, cpoffset=1e-14):
    xs, ys = np.meshgrid(np.arange(nx + 1) * h, np.arange(ny + 1) * h * 1.5, indexing="ij")
    nodes = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)])
    vid = lambda i, j: i * (ny + 1) + j
    cells = []
    for i in range(nx):
        for j in range(ny):
            cells.append([vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)])   # CCW, n = +z
    ptr = np.arange(0, 4 * len(cells) + 1, 4)
    return types.SimpleNamespace(nodes=np.asfortranarray(nodes), cell_ptr=ptr, cell_idx=np.concatenate(cells),
                                 ncells=len(cells), CPoffset=cpoffset, charlen_kind=0), nx, ny


def test_linear_strength_exact_on_planar_grid(orc):
    a, b = 0.7, -1.3
    for cpoff, sgn in ((1e-14, 1.0), (-1e-14, -1.0)):
        body, nx, ny = _planar(cpoffset=cpoff)
        n, cp, _ = orc.geometry(body.nodes, body.cell_ptr, body.cell_idx, 0.0, 0)
        G = a * cp[0] + b * cp[1]
        out = orc.ugradmu_cell(body, G)
        for i in range(1, nx - 1):
            for j in range(1, ny - 1):
                c = i * ny + j
                assert np.abs(out[:, c] - sgn * (-0.5) * np.array([a, b, 0.0])).max() < 1e-12


def test_uniform_strength_closed_body_is_zero(orc, geo):
    g = geo.sphere(n_meridional=8, n_azimuthal=12)
    body = types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=g.cell_idx, ncells=g.ncells,
                                 CPoffset=1e-14, charlen_kind=0)
    nb = orc.cell_neighbors(g.cell_ptr, g.cell_idx)
    assert (nb >= 0).sum() == np.diff(g.cell_ptr).sum()   # closed: every edge has a neighbour
    out = orc.ugradmu_cell(body, np.full(g.ncells, 2.5))
    assert np.abs(out).max() < 1e-12


def test_trailing_edge_panels_drop_their_terms(orc, geo):
    g, sh = geo.loft_wing(n_chord=12, n_span=4, span=2.0)
    body = types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=g.cell_idx, ncells=g.ncells,
                                 CPoffset=1e-14, charlen_kind=0)
    G = np.random.default_rng(1).normal(size=g.ncells)
    full = orc.ugradmu_cell(body, G, sh, te=False)
    te = orc.ugradmu_cell(body, G, sh, te=True)
    tecells = set(sh[0]) | set(sh[3][sh[3] >= 0])
    for c in range(g.ncells):
        if c in tecells:
            assert np.abs(te[:, c]).max() == 0.0
        else:
            assert np.array_equal(te[:, c], full[:, c])


# ---- node-based and combined schemes (src/FLOWPanel_postprocess.jl:622-1000) ----
#
# GeometricTools (nodal averaging, planar projection, triangle gradient) is not
# vendored, so these schemes are pinned by what the mathematics fixes: a linear
# strength makes every interior nodal mean (point-symmetric ring of equal-area
# cells) equal the node's own linear value, so the plane through any interior
# cell's vertices is the linear field itself; uniform strengths give zero; the
# sign follows CPoffset; the combined scheme picks node values exactly on open
# or sharp edges.

def _split(body):
    """each quad (a, b, c, d) -> triangles (a, b, c), (a, c, d)"""
    tris = []
    for c in range(body.ncells):
        a, b, cc, d = body.cell_idx[body.cell_ptr[c]:body.cell_ptr[c + 1]]
        tris += [[a, b, cc], [a, cc, d]]
    return types.SimpleNamespace(nodes=body.nodes, cell_ptr=np.arange(0, 3 * len(tris) + 1, 3),
                                 cell_idx=np.concatenate(tris), ncells=len(tris), CPoffset=body.CPoffset,
                                 charlen_kind=0)


def _interior_cells(body, nx, ny, h=0.3):
    """cells whose first three vertices are interior nodes of the planar grid"""
    x, y = body.nodes[0], body.nodes[1]
    inner = (x > 1e-9) & (x < nx * h - 1e-9) & (y > 1e-9) & (y < ny * h * 1.5 - 1e-9)
    return [c for c in range(body.ncells) if inner[body.cell_idx[body.cell_ptr[c]:body.cell_ptr[c] + 3]].all()]


def test_node_scheme_linear_strength_exact(orc):
    a, b = 0.7, -1.3
    for tri in (False, True):
        for cpoff, sgn in ((1e-14, 1.0), (-1e-14, -1.0)):
            body, nx, ny = _planar(nx=7, ny=6, cpoffset=cpoff)
            if tri:
                body = _split(body)
            n, cp, _ = orc.geometry(body.nodes, body.cell_ptr, body.cell_idx, 0.0, 0)
            G = a * cp[0] + b * cp[1]
            out, nodal = orc.ugradmu_node(body, G)
            cells = _interior_cells(body, nx, ny)
            assert len(cells) > 10
            for c in cells:
                assert np.abs(out[:, c] - sgn * (-0.5) * np.array([a, b, 0.0])).max() < 1e-12, (tri, c)
            # interior nodal values are the linear field at the node
            x, y = body.nodes[0], body.nodes[1]
            for v in set(np.concatenate([body.cell_idx[body.cell_ptr[c]:body.cell_ptr[c] + 3] for c in cells])):
                assert abs(nodal[v] - (a * x[v] + b * y[v])) < 1e-12


def test_node_scheme_uniform_strength_is_zero(orc, geo):
    g, sh = geo.loft_wing(n_chord=10, n_span=4, span=2.0)
    body = types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=g.cell_idx, ncells=g.ncells,
                                 CPoffset=1e-14, charlen_kind=0)
    out, nodal = orc.ugradmu_node(body, np.full(g.ncells, -1.75))
    assert np.abs(out).max() < 1e-12
    used = np.unique(g.cell_idx)
    assert np.abs(nodal[used] + 1.75).max() < 1e-14


def test_combined_scheme_selects_node_values_on_open_and_sharp_edges(orc):
    body, nx, ny = _planar(nx=6, ny=5)
    # fold the grid along x = 3h by 90 degrees: z = x - 3h beyond the fold, x clamped
    h = 0.3
    x = body.nodes[0].copy()
    fold = x > 3 * h + 1e-12
    body.nodes[2, fold] = x[fold] - 3 * h
    body.nodes[0, fold] = 3 * h
    G = np.random.default_rng(7).normal(size=body.ncells)
    cell = orc.ugradmu_cell(body, G, None, te=True)
    node, _ = orc.ugradmu_node(body, G)
    comb = orc.ugradmu_combined(body, G, None, anglecrit=30.0)
    nb = orc.cell_neighbors(body.cell_ptr, body.cell_idx)
    for c in range(body.ncells):
        i = c // ny
        open_edge = (nb[c] < 0).any()
        at_fold = i in (2, 3)
        want = node[:, c] if (open_edge or at_fold) else cell[:, c]
        assert np.array_equal(comb[:, c], want), c
    # a critical angle above the fold angle keeps cell values at the fold
    comb95 = orc.ugradmu_combined(body, G, None, anglecrit=95.0)
    for c in range(body.ncells):
        if not (nb[c] < 0).any():
            assert np.array_equal(comb95[:, c], cell[:, c])


# ---- sharp trailing edge (src/FLOWPanel_postprocess.jl:692-748) ----
#
# A strength that is one constant on the cells of the lower half of the section
# loop and another on the upper half jumps across the trailing edge, as a lifting
# solution does.  The plain nodal mean mixes both sides at the TE nodes and puts
# a spurious gradient into the TE cells; the sharp-TE pass averages each side's
# TE nodal values over that side's cells only, so every TE cell sees a uniform
# field and gets exactly zero.  Cells away from the TE are untouched.

def _split_strength(g, nc, cl=-0.8, cu=1.7):
    i = np.arange(g.ncells) % nc            # chordwise position of each cell in its strip
    return np.where(i < nc // 2, cl, cu)


def _wing_body(g):
    """The wing with coincident nodes merged: the section loop closes at the trailing
    edge, so upper and lower TE cells share their TE nodes (a grid closed in its loop
    dimension, the case the reference's sharpTE pass is written for)."""
    X = np.asarray(g.nodes)
    _, first, inv = np.unique(np.round(X.T, 12), axis=0, return_index=True, return_inverse=True)
    rep = first[inv.ravel()]                    # every node -> its first coincident copy
    return types.SimpleNamespace(nodes=g.nodes, cell_ptr=g.cell_ptr, cell_idx=rep[np.asarray(g.cell_idx)],
                                 ncells=g.ncells, CPoffset=1e-14, charlen_kind=0)


def test_sharp_te_quads_zero_gradient_at_the_trailing_edge(orc, geo):
    nc = 12
    g, sh = geo.loft_wing(n_chord=nc, n_span=5, span=2.5, sweep_deg=10.0)
    body = _wing_body(g)
    G = _split_strength(g, nc)
    plain, nod0 = orc.ugradmu_node(body, G)
    sharp, nod1 = orc.ugradmu_node(body, G, sh, sharp_te=2)
    A, B, te = orc.te_sides(body, sh, pairs=False)
    assert sorted(A) == sorted(sh[0].tolist()) and sorted(B) == sorted(sh[3].tolist())
    for c in A + B:
        assert np.abs(sharp[:, c]).max() < 1e-12
        assert np.abs(plain[:, c]).max() > 1e-2          # the mixing the pass removes
    rest = [c for c in range(g.ncells) if c not in A + B]
    assert np.array_equal(sharp[:, rest], plain[:, rest])
    # the nodal values left behind carry side B's values at the TE nodes
    assert np.allclose(nod1[te], G[sh[3][0]]) and np.array_equal(np.delete(nod1, te), np.delete(nod0, te))


def test_sharp_te_triangle_pairs(orc, geo):
    nc = 12
    g, sh = geo.loft_wing(n_chord=nc, n_span=4, span=2.0)
    gt, sht = geo.triangulate(g, sh)
    body = _wing_body(gt)
    G = _split_strength(gt, 2 * nc)
    sharp, _ = orc.ugradmu_node(body, G, sht, sharp_te=1)
    plain, _ = orc.ugradmu_node(body, G)
    A, B, _ = orc.te_sides(body, sht, pairs=True)
    assert len(A) == 2 * sh.shape[1] and len(B) == 2 * sh.shape[1]
    for c in A + B:
        assert np.abs(sharp[:, c]).max() < 1e-12
    assert max(np.abs(plain[:, c]).max() for c in A + B) > 1e-2


def test_sharp_te_combined_uses_it(orc, geo):
    # the combined scheme hands sharpTE to its node-based part: with force_cellTE off, the TE cells
    # (open-edge or sharp-angle cells take node values) see the sharp-TE values
    nc = 12
    g, sh = geo.loft_wing(n_chord=nc, n_span=5, span=2.5)
    body = _wing_body(g)
    G = _split_strength(g, nc)
    node, _ = orc.ugradmu_node(body, G, sh, sharp_te=2)
    comb = orc.ugradmu_combined(body, G, sh, force_cell_te=False, anglecrit=30.0, sharp_te=2)
    A, B, _ = orc.te_sides(body, sh, pairs=False)
    for c in A + B:     # a TE cell's neighbour across the sharp edge is more than 30 degrees away
        assert np.array_equal(comb[:, c], node[:, c])
