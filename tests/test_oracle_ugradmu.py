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
