"""Edge-pair assembly tables of the vortex-ring AIC, checked on the CPU.

The device's edge-pair vortex-ring assembly (FLOWPANEL_ASSEMBLY=pairs) takes the
shared-term column blocks with each block's distinct edges paired at shared nodes: lane l evaluates the
node terms of its path n0 - n1 - n2 in registers and writes the segment terms of
n0 -> n1 (edge slot l) and n1 -> n2 (slot 256 + l); each column sums its edges'
slots with the sign of its traversal.  The host builds the tables
(fp_assembly_pairs, no device involved).  Here: every lane's edges are distinct
mesh edges of its block's columns, every such edge sits in exactly one slot, at
most 256 lanes per block, far fewer lanes than edges (pairing); and a numpy
restatement of the kernel's arithmetic over the tables reproduces the oracle's
AIC (src/FLOWPanel_liftingbody.jl:455-572 through U_vortexring,
src/FLOWPanel_elements.jl:442-467) to 1e-12 of max|G| on quads, triangles, a
mix, a sphere with collapsed pole edges, disjoint quads and disjoint triangles
(the pairing's odd edges).  Planted mistakes (a flipped sign, a swapped path end,
a dropped reference) fail the comparison.
"""
import ctypes

import numpy as np
import pytest

from test_assembly_blocks import _cases, oracle_G

P, BLOCK = 256, 6 * 256
OFF_COL, OFF_REF, OFF_NODE = 0, P, 3 * P


def pairs(grid):
    from flowpanel_amd import _native
    L = _native.lib()
    nodes = np.asfortranarray(grid.nodes, dtype=np.float64)
    ptr = np.ascontiguousarray(grid.cell_ptr, dtype=np.int64)
    idx = np.ascontiguousarray(grid.cell_idx, dtype=np.int64)
    n = len(ptr) - 1
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    terms = np.zeros(2, np.int64)
    nb = L.fp_assembly_pairs(dp(nodes), nodes.shape[1], ip(ptr), ip(idx), n, 0, None, 0, ip(terms))
    assert nb > 0
    tab = np.zeros(nb * BLOCK, np.int32)
    nb2 = L.fp_assembly_pairs(dp(nodes), nodes.shape[1], ip(ptr), ip(idx), n, 0,
                              tab.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nb, ip(terms))
    assert nb2 == nb
    return tab.reshape(nb, BLOCK), terms


def node_of(grid, ref):
    return grid.cell_idx[grid.cell_ptr[ref >> 3] + (ref & 7)]


def check_structure(grid, tab):
    """Lanes hold distinct edges of their block, every edge of the block's columns once."""
    ncol = len(grid.cell_ptr) - 1
    seen_cols = np.zeros(ncol, int)
    lanes_total = 0
    for T in tab:
        n0, n1, n2 = T[OFF_NODE:OFF_NODE + P], T[OFF_NODE + P:OFF_NODE + 2 * P], T[OFF_NODE + 2 * P:OFF_NODE + 3 * P]
        used = n0 >= 0
        k = int(used.sum())
        assert np.all(used[:k]) and not np.any(used[k:]), "lanes packed from 0"
        assert np.all(n1[:k] >= 0) and np.all(n1[k:] < 0) and np.all(n2[k:] < 0)
        lanes_total += k
        slots = {}
        for l in range(k):
            a, b = node_of(grid, n0[l]), node_of(grid, n1[l])
            slots[l] = (a, b)
            if n2[l] >= 0:
                slots[P + l] = (b, node_of(grid, n2[l]))
        und = [tuple(sorted(e)) for e in slots.values()]
        assert len(und) == len(set(und)), "an edge in two slots"
        assert all(a != b for a, b in slots.values())
        cols = T[OFF_COL:OFF_COL + P]
        want = set()
        for c in cols[cols >= 0]:
            seen_cols[c] += 1
            cc = grid.cell(c)
            for q in range(len(cc)):
                u, w = cc[q], cc[(q + 1) % len(cc)]
                if u != w:
                    want.add(tuple(sorted((u, w))))
        assert set(und) == want, "slots are not the block's edges"
    assert np.all(seen_cols == 1), "every column in exactly one block"
    return lanes_total


def emulate(grid, tab, X, nrm, off=1e-8, cut=1e-14):
    """The kernel's arithmetic over the pair tables, in numpy (FP64)."""
    M = X.shape[1]
    n = len(grid.cell_ptr) - 1
    G = np.full((M, n), np.nan)
    for T in tab:
        q0, q1, q2 = T[OFF_NODE:OFF_NODE + P], T[OFF_NODE + P:OFF_NODE + 2 * P], T[OFF_NODE + 2 * P:OFF_NODE + 3 * P]
        f0 = np.where(q0 >= 0, q0, T[OFF_NODE] if T[OFF_NODE] >= 0 else 0)
        f1 = np.where(q1 >= 0, q1, f0)
        f2 = np.where(q2 >= 0, q2, f1)
        p0, p1, p2 = (grid.nodes[:, node_of(grid, f)] for f in (f0, f1, f2))      # 3 x P each
        r0, r1, r2 = (X[:, :, None] - p[:, None, :] for p in (p0, p1, p2))      # 3 x M x P
        iv0, iv1, iv2 = (1.0 / (np.sqrt((r * r).sum(0)) + off) for r in (r0, r1, r2))

        def seg(A, ia, B, ib, e):
            c = np.cross(A, B, axis=0)
            c2 = (c * c).sum(0)
            proj = (e[:, None, :] * (A * ia[None] - B * ib[None])).sum(0)
            cn = (c * nrm[:, :, None]).sum(0)
            return np.where(c2 > cut * cut, proj / (4 * np.pi * (c2 + off * off)) * cn, 0.0)

        se = np.concatenate([seg(r0, iv0, r1, iv1, p1 - p0), seg(r1, iv1, r2, iv2, p2 - p1)], axis=1)  # M x 2P
        cols = T[OFF_COL:OFF_COL + P]
        refs = T[OFF_REF:OFF_REF + 2 * P].view(np.uint32)
        for lane in np.nonzero(cols >= 0)[0]:
            r4 = [refs[2 * lane] & 0xFFFF, refs[2 * lane] >> 16, refs[2 * lane + 1] & 0xFFFF, refs[2 * lane + 1] >> 16]
            g = np.zeros(M)
            for q in r4:
                if q != 0xFFFF:
                    assert (q >> 1) < 2 * P
                    g += (-1.0 if q & 1 else 1.0) * se[:, q >> 1]
            assert np.isnan(G[0, cols[lane]]), "column in two blocks"
            G[:, cols[lane]] = g
    assert not np.isnan(G).any(), "column in no block"
    return G


def _disjoint_triangles(fp, n=260, seed=2):
    rng = np.random.default_rng(seed)
    nodes, cells = [], []
    for k in range(n):
        c = rng.uniform(-3, 3, 3)
        u, v = rng.normal(size=3), rng.normal(size=3)
        nodes += [c, c + 0.1 * u, c + 0.1 * v]
        cells.append(np.arange(3 * k, 3 * k + 3))
    return fp.Grid(np.array(nodes).T, cells)


@pytest.mark.parametrize("case", ["quads", "triangles", "mixed", "sphere", "disjoint", "disjoint_triangles"])
def test_pairs_reproduce_the_oracle_aic(fp, geo, orc, case):
    grid = _disjoint_triangles(fp) if case == "disjoint_triangles" else _cases(fp, geo)[case]
    tab, terms = pairs(grid)
    lanes = check_structure(grid, tab)
    assert lanes == terms[0]
    G, nrm, X = oracle_G(orc, grid)
    Ge = emulate(grid, tab, X, nrm)
    err = np.abs(Ge - G).max() / np.abs(G).max()
    n = G.shape[1]
    print("%s: %d panels, %d blocks, %.2f lanes and %.2f edges per panel, err %.1e" %
          (case, n, len(tab), terms[0] / n, terms[1] / n, err))
    assert err < 1e-12
    # pairing: a connected patch leaves at most one odd edge, so lanes ~ edges / 2
    if case in ("quads", "triangles", "mixed"):
        assert terms[0] <= terms[1] / 2 + 2 * len(tab)


def test_pair_blocks_on_a_large_loft(fp, geo, orc):
    g, _ = geo.loft_wing(n_chord=60, n_span=80)       # 4,800 quads
    tab, terms = pairs(g)
    n = len(g.cell_ptr) - 1
    check_structure(g, tab)
    assert (tab[:, OFF_NODE:OFF_NODE + P] >= 0).sum(1).max() <= P
    # two edges per lane on almost every lane: about 1.1 lanes per panel
    assert terms[0] / n < 1.2, terms / n
    G, nrm, X = oracle_G(orc, g)
    sel = np.arange(0, n, 97)
    Ge = emulate(g, tab, X[:, sel], nrm[:, sel])
    assert np.abs(Ge - G[sel]).max() / np.abs(G[sel]).max() < 1e-12


@pytest.mark.parametrize("plant", ["sign", "swap", "drop"])
def test_planted_pair_mistakes_fail(fp, geo, orc, plant):
    grid = _cases(fp, geo)["quads"]
    tab, _ = pairs(grid)
    G, nrm, X = oracle_G(orc, grid)
    bad = tab.copy()
    if plant == "sign":
        r = bad[0, OFF_REF:OFF_REF + 2].view(np.uint32)
        r[0] ^= 1
    elif plant == "swap":       # reverse lane 3's path: both edges now run the other way
        a, c = int(bad[0, OFF_NODE + 3]), int(bad[0, OFF_NODE + 2 * P + 3])
        assert c >= 0
        bad[0, OFF_NODE + 3], bad[0, OFF_NODE + 2 * P + 3] = c, a
    else:
        r = bad[0, OFF_REF + 2:OFF_REF + 4].view(np.uint32)
        r[0] |= 0xFFFF
    Ge = emulate(grid, bad, X, nrm)
    assert np.abs(Ge - G).max() / np.abs(G).max() > 1e-6
