"""Edge trails of the vortex-ring velocity sum, checked on the CPU.

The device evaluates U(X) = sum_j Gamma_j U_vortexring_j(X) (the reference's
Uind!, src/FLOWPanel_liftingbody.jl:576-661) as a sum over the mesh's edges,
gamma_e = the signed strengths of the panels holding edge e, walked as trails
so that each step reuses the previous node's terms.  The host builds the trails
(fp_field_trails, no device involved).  Here a numpy restatement of the
kernel's arithmetic over the trails must reproduce the oracle's induced
velocity (1e-12 of max|U|) on quads, triangles, a quad/triangle mix, a sphere
with collapsed pole edges and hexagons; every mesh edge is walked exactly once
and each step starts where the previous one ended.  A planted mistake (one
contributor's sign flipped, one contributor dropped) fails the comparison.
"""
import ctypes
import types

import numpy as np
import pytest


def trails(grid):
    from flowpanel_amd import _native
    L = _native.lib()
    ptr = np.ascontiguousarray(grid.cell_ptr, dtype=np.int64)
    idx = np.ascontiguousarray(grid.cell_idx, dtype=np.int64)
    n = len(ptr) - 1
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    i32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    nc = np.zeros(1, np.int64)
    S = L.fp_field_trails(ip(ptr), ip(idx), n, grid.nodes.shape[1], 0, None, None, None, 0, 0, ip(nc))
    assert S > 0
    ref, cptr, cpan = np.zeros(S, np.int32), np.zeros(S + 1, np.int32), np.zeros(max(1, nc[0]), np.int32)
    S2 = L.fp_field_trails(ip(ptr), ip(idx), n, grid.nodes.shape[1], 0, i32(ref), i32(cptr), i32(cpan), S, nc[0],
                           ip(nc))
    assert S2 == S
    return ref, cptr, cpan[:nc[0]]


def node_of(grid, ref):
    return grid.cell_idx[grid.cell_ptr[ref >> 3] + (ref & 7)]


def emulate(grid, tr, strength, X, off=1e-8, cut=1e-14):
    ref, cptr, cpan = tr
    P = grid.nodes[:, node_of(grid, ref)]
    r = X[:, :, None] - P[:, None, :]                       # 3 x M x S
    iv = 1.0 / (np.sqrt((r * r).sum(0)) + off)
    U = np.zeros(X.shape)
    for m in range(1, len(ref)):
        c0, c1 = cptr[m], cptr[m + 1]
        if c1 == c0:
            continue
        g = sum(-strength[q >> 1] if q & 1 else strength[q >> 1] for q in cpan[c0:c1])
        e = P[:, m] - P[:, m - 1]
        a, b = r[:, :, m - 1], r[:, :, m]
        c = np.cross(a, b, axis=0)
        c2 = (c * c).sum(0)
        proj = (e[:, None] * (a * iv[None, :, m - 1] - b * iv[None, :, m])).sum(0)
        aux = np.where(c2 > cut * cut, g * proj / (4 * np.pi * (c2 + off * off)), 0.0)
        U += aux * c
    return U


def obody(grid):
    return types.SimpleNamespace(nodes=grid.nodes, cell_ptr=grid.cell_ptr, cell_idx=grid.cell_idx,
                                 ncells=len(grid.cell_ptr) - 1, element="vortexring",
                                 shedding=np.zeros((6, 0), np.int64), nsheddings=0, kerneloffset=1e-8,
                                 kernelcutoff=1e-14, CPoffset=1e-14, charlen_kind=0)


def _grids(fp, geo):
    g1, _ = geo.loft_wing(n_chord=12, n_span=10)
    g2, _ = geo.triangulate(*geo.loft_wing(n_chord=12, n_span=8))
    q = g1.cell_idx.reshape(-1, 4)
    mixed = []
    for k, c in enumerate(q):
        mixed += [c[[0, 1, 2]], c[[0, 2, 3]]] if k % 3 == 1 else [c]
    # hexagons: pairs of chordwise-adjacent quads (i, i+1) of one strip merged along their shared edge
    n1, n2 = g1.dims
    hexes = []
    for j in range(n2):
        for i in range(0, n1 - 1, 2):
            a, b = q[i + n1 * j], q[i + 1 + n1 * j]
            # quad a = (v0, v1, v2, v3) and b share a's edge (v1, v2) as b's (v0, v3) reversed
            if a[1] == b[0] and a[2] == b[3]:
                hexes.append([a[0], b[0], b[1], b[2], b[3], a[3]])
            else:
                hexes += [a, b]
    return {"quads": g1, "triangles": g2, "mixed": fp.Grid(g1.nodes, mixed),
            "sphere": geo.sphere(n_meridional=8, n_azimuthal=12), "hexagons": fp.Grid(g1.nodes, hexes)}


@pytest.mark.parametrize("case", ["quads", "triangles", "mixed", "sphere", "hexagons"])
def test_trails_reproduce_the_oracle_velocity(fp, geo, orc, case):
    grid = _grids(fp, geo)[case]
    tr = trails(grid)
    ref, cptr, cpan = tr
    n = len(grid.cell_ptr) - 1
    # every unique non-degenerate edge exactly once, steps chained node to node
    edges = set()
    for c in range(n):
        cc = grid.cell(c)
        for k in range(len(cc)):
            u, w = cc[k], cc[(k + 1) % len(cc)]
            if u != w:
                edges.add((min(u, w), max(u, w)))
    nodes = node_of(grid, ref)
    walked = [(min(nodes[m - 1], nodes[m]), max(nodes[m - 1], nodes[m])) for m in range(1, len(ref))
              if cptr[m + 1] > cptr[m]]
    assert cptr[1] == cptr[0] == 0                         # the stream starts a trail
    assert len(walked) == len(set(walked)) == len(edges) and set(walked) == edges
    starts = int(sum(cptr[m + 1] == cptr[m] for m in range(len(ref))))
    print("%s: %d panels, %d edges, %d trails, %.2f entries per panel" % (case, n, len(edges), starts, len(ref) / n))
    rng = np.random.default_rng(3)
    strength = rng.normal(size=n)
    X = np.vstack([rng.uniform(-1, 5, 40), rng.uniform(-4, 4, 40), rng.uniform(-1, 1, 40)])
    Uo = orc.uind(obody(grid), strength, X)
    Ue = emulate(grid, tr, strength, X)
    assert np.abs(Ue - Uo).max() / np.abs(Uo).max() < 1e-12


@pytest.mark.parametrize("plant", ["sign", "drop"])
def test_planted_trail_mistakes_fail(fp, geo, orc, plant):
    grid = _grids(fp, geo)["quads"]
    ref, cptr, cpan = trails(grid)
    n = len(grid.cell_ptr) - 1
    strength = np.random.default_rng(4).normal(size=n)
    X = np.array([[0.3, 0.5], [0.2, -1.0], [0.05, 0.1]])
    Uo = orc.uind(obody(grid), strength, X)
    m = next(k for k in range(1, len(ref)) if cptr[k + 1] - cptr[k] == 2)
    bad = cpan.copy()
    if plant == "sign":
        bad[cptr[m]] ^= 1
        Ue = emulate(grid, (ref, cptr, bad), strength, X)
    else:
        keep = np.ones(len(cpan), bool)
        keep[cptr[m]] = False
        c2 = cptr.copy()
        c2[m + 1:] -= 1
        Ue = emulate(grid, (ref, c2, cpan[keep]), strength, X)
    assert np.abs(Ue - Uo).max() / np.abs(Uo).max() > 1e-8
