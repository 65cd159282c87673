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
