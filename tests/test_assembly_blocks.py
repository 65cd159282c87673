"""Shared node/edge assembly blocks of the vortex-ring AIC, checked on the CPU.

The device kernel evaluates, per target and column block, each mesh node's
r = X - p and 1/(|r| + offset) and each mesh edge's regularised segment term
once; a panel adds its edges' terms with the sign of its traversal.  The host
builds the blocks (fp_assembly_blocks, no device involved).  Here a numpy
restatement of the kernel's arithmetic over those tables must reproduce the
oracle's AIC (src/FLOWPanel_liftingbody.jl:455-572, panel by panel through
U_vortexring, src/FLOWPanel_elements.jl:442-467) to 1e-12 of max|G| on quads,
triangles, a mix, a sphere with collapsed pole edges and disjoint panels (the
node limit); every column sits in exactly one block and the packing limits
(256 panels, 384 nodes, 640 edges) hold.
A planted mistake in the tables (a flipped traversal sign, a swapped edge end,
a dropped edge) fails the comparison.
"""
import ctypes
import types

import numpy as np
import pytest

P, NMAX, EMAX, BLOCK = 256, 512, 768, 2048
OFF_NODE, OFF_EDGE, OFF_COL, OFF_REF = 0, NMAX, NMAX + EMAX, NMAX + EMAX + P


def blocks(fp, grid):
    from flowpanel_amd import _native
    L = _native.lib()
    nodes = np.asfortranarray(grid.nodes, dtype=np.float64)
    ptr = np.ascontiguousarray(grid.cell_ptr, dtype=np.int64)
    idx = np.ascontiguousarray(grid.cell_idx, dtype=np.int64)
    n = len(ptr) - 1
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    terms = np.zeros(2, np.int64)
    nb = L.fp_assembly_blocks(dp(nodes), nodes.shape[1], ip(ptr), ip(idx), n, 0, None, 0, ip(terms))
    assert nb > 0
    tab = np.zeros(nb * BLOCK, np.int32)
    nb2 = L.fp_assembly_blocks(dp(nodes), nodes.shape[1], ip(ptr), ip(idx), n, 0,
                               tab.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nb, ip(terms))
    assert nb2 == nb
    return tab.reshape(nb, BLOCK), terms


def emulate(grid, tab, X, nrm, off=1e-8, cut=1e-14):
    """The kernel's arithmetic over the block tables, in numpy (FP64)."""
    M = X.shape[1]
    n = len(grid.cell_ptr) - 1
    G = np.full((M, n), np.nan)
    for T in tab:
        nref = T[OFF_NODE:OFF_NODE + NMAX]
        nn = int((nref >= 0).sum())
        assert np.all(nref[:nn] >= 0) and np.all(nref[nn:] < 0)
        pan, slot = nref[:nn] >> 3, nref[:nn] & 7
        gid = grid.cell_idx[grid.cell_ptr[pan] + slot]
        Pn = grid.nodes[:, gid]                                     # 3 x nn
        r = X[:, :, None] - Pn[:, None, :]                          # 3 x M x nn
        iv = 1.0 / (np.sqrt((r * r).sum(0)) + off)                  # M x nn
        ed = T[OFF_EDGE:OFF_EDGE + EMAX]
        ne = int((ed >= 0).sum())
        assert np.all(ed[:ne] >= 0) and np.all(ed[ne:] < 0)
        a, b = ed[:ne] & 0xFFFF, (ed[:ne] >> 16) & 0xFFFF
        assert a.max(initial=0) < nn and b.max(initial=0) < nn
        e = Pn[:, b] - Pn[:, a]
        A, B = r[:, :, a], r[:, :, b]
        c = np.cross(A, B, axis=0)
        c2 = (c * c).sum(0)
        proj = (e[:, None, :] * (A * iv[None, :, a] - B * iv[None, :, b])).sum(0)
        cn = (c * nrm[:, :, None]).sum(0)
        val = np.where(c2 > cut * cut, proj / (4 * np.pi * (c2 + off * off)) * cn, 0.0)
        cols = T[OFF_COL:OFF_COL + P]
        refs = T[OFF_REF:OFF_REF + 2 * P].view(np.uint32)
        for lane in np.nonzero(cols >= 0)[0]:
            r4 = [refs[2 * lane] & 0xFFFF, refs[2 * lane] >> 16, refs[2 * lane + 1] & 0xFFFF, refs[2 * lane + 1] >> 16]
            g = np.zeros(M)
            for q in r4:
                if q != 0xFFFF:
                    assert (q >> 1) < ne
                    g += (-1.0 if q & 1 else 1.0) * val[:, q >> 1]
            assert np.isnan(G[0, cols[lane]]), "column in two blocks"
            G[:, cols[lane]] = g
    assert not np.isnan(G).any(), "column in no block"
    return G


def oracle_G(orc, grid, X=None, nrm=None):
    b = types.SimpleNamespace(nodes=grid.nodes, cell_ptr=grid.cell_ptr, cell_idx=grid.cell_idx,
                              ncells=len(grid.cell_ptr) - 1, element="vortexring", shedding=np.zeros((6, 0), np.int64),
                              nsheddings=0, kerneloffset=1e-8, kernelcutoff=1e-14, CPoffset=1e-14, charlen_kind=0)
    return orc.assemble([b], [None], [None], targets=X, normals=nrm)


def _mixed(fp, geo):
    g, _ = geo.loft_wing(n_chord=12, n_span=10)
    q = g.cell_idx.reshape(-1, 4)
    cells = []
    for k, c in enumerate(q):
        if k % 3 == 1:
            cells += [c[[0, 1, 2]], c[[0, 2, 3]]]
        else:
            cells.append(c)
    return fp.Grid(g.nodes, cells)


def _disjoint(fp, n=300, seed=1):
    rng = np.random.default_rng(seed)
    nodes, cells = [], []
    for k in range(n):
        c = rng.uniform(-3, 3, 3)
        u, v = rng.normal(size=3), rng.normal(size=3)
        pts = [c, c + 0.1 * u, c + 0.1 * (u + v), c + 0.1 * v]
        cells.append(np.arange(4 * k, 4 * k + 4))
        nodes += pts
    return fp.Grid(np.array(nodes).T, cells)


def _cases(fp, geo):
    g1, _ = geo.loft_wing(n_chord=12, n_span=10)
    g2, _ = geo.triangulate(*geo.loft_wing(n_chord=12, n_span=8))
    return {"quads": g1, "triangles": g2, "mixed": _mixed(fp, geo), "sphere": geo.sphere(n_meridional=8, n_azimuthal=12),
            "disjoint": _disjoint(fp)}


@pytest.mark.parametrize("case", ["quads", "triangles", "mixed", "sphere", "disjoint"])
def test_blocks_reproduce_the_oracle_aic(fp, geo, orc, case):
    grid = _cases(fp, geo)[case]
    tab, terms = blocks(fp, grid)
    G, nrm, X = oracle_G(orc, grid)
    Ge = emulate(grid, tab, X, nrm)
    err = np.abs(Ge - G).max() / np.abs(G).max()
    print("%s: %d panels, %d blocks, %.2f nodes and %.2f edges per panel, err %.1e" %
          (case, G.shape[1], len(tab), terms[0] / G.shape[1], terms[1] / G.shape[1], err))
    assert err < 1e-12
    if case == "disjoint":
        # four nodes per panel: the 384-node packing limit gives 96 panels per block
        assert len(tab) == int(np.ceil(G.shape[1] / 96))


def test_block_limits_and_sharing_on_a_large_loft(fp, geo, orc):
    g, _ = geo.loft_wing(n_chord=60, n_span=80)       # 4,800 quads
    tab, terms = blocks(fp, g)
    n = len(g.cell_ptr) - 1
    cols = tab[:, OFF_COL:OFF_COL + P]
    assert np.array_equal(np.sort(cols[cols >= 0]), np.arange(n))
    assert (tab[:, OFF_NODE:OFF_NODE + NMAX] >= 0).sum(1).max() <= 384
    assert (tab[:, OFF_EDGE:OFF_EDGE + EMAX] >= 0).sum(1).max() <= 640
    # the point of the blocks: far fewer node and edge terms than 4 + 4 per panel
    assert terms[0] / n < 1.5 and terms[1] / n < 2.6, terms / n
    # a sample of targets against the oracle
    G, nrm, X = oracle_G(orc, g)
    sel = np.arange(0, n, 97)
    Ge = emulate(g, tab, X[:, sel], nrm[:, sel])
    assert np.abs(Ge - G[sel]).max() / np.abs(G[sel]).max() < 1e-12


@pytest.mark.parametrize("plant", ["sign", "swap", "drop"])
def test_planted_table_mistakes_fail(fp, geo, orc, plant):
    grid = _cases(fp, geo)["quads"]
    tab, _ = blocks(fp, grid)
    G, nrm, X = oracle_G(orc, grid)
    bad = tab.copy()
    if plant == "sign":
        r = bad[0, OFF_REF:OFF_REF + 2].view(np.uint32)
        r[0] ^= 1
    elif plant == "swap":
        e = int(bad[0, OFF_EDGE + 5])
        bad[0, OFF_EDGE + 5] = ((e & 0xFFFF) << 16) | ((e >> 16) & 0xFFFF)
    else:
        r = bad[0, OFF_REF + 2:OFF_REF + 4].view(np.uint32)
        r[0] |= 0xFFFF
    Ge = emulate(grid, bad, X, nrm)
    assert np.abs(Ge - G).max() / np.abs(G).max() > 1e-6
