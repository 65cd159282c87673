"""VTK output (src/FLOWPanel_liftingbody.jl:146-163, 1092-1196; src/FLOWPanel_abstractbody.jl:136-190) on
the CPU: the files are read back and every point, cell and field compared with what was written."""
import os
import types

import numpy as np


def _read_legacy(path):
    toks = open(path).read().split("\n")
    out = {"cell_data": {}, "point_data": {}}
    k = 0
    section = None
    while k < len(toks):
        ln = toks[k].split()
        if not ln:
            k += 1
            continue
        if ln[0] == "POINTS":
            n = int(ln[1])
            out["points"] = np.array([[float(v) for v in toks[k + 1 + i].split()] for i in range(n)]).T
            k += n + 1
        elif ln[0] == "CELLS":
            n = int(ln[1])
            out["cells"] = [[int(v) for v in toks[k + 1 + i].split()[1:]] for i in range(n)]
            k += n + 1
        elif ln[0] == "CELL_TYPES":
            n = int(ln[1])
            out["types"] = [int(toks[k + 1 + i]) for i in range(n)]
            k += n + 1
        elif ln[0] in ("CELL_DATA", "POINT_DATA"):
            section = "cell_data" if ln[0] == "CELL_DATA" else "point_data"
            count = int(ln[1])
            k += 1
        elif ln[0] == "SCALARS":
            out[section][ln[1]] = np.array([float(toks[k + 2 + i]) for i in range(count)])
            k += count + 2
        elif ln[0] == "VECTORS":
            out[section][ln[1]] = np.array([[float(v) for v in toks[k + 1 + i].split()] for i in range(count)]).T
            k += count + 1
        else:
            k += 1
    return out


def test_body_and_wake_round_trip(fp, geo, tmp_path):
    g, sh = geo.loft_wing(n_chord=8, n_span=4, span=2.0)
    n = g.ncells
    rng = np.random.default_rng(0)
    G = rng.standard_normal(n)
    U = rng.standard_normal((3, n))
    Da = np.tile([[1.0], [0.0], [0.1]], (1, sh.shape[1]))
    Db = np.tile([[1.0], [0.1], [0.0]], (1, sh.shape[1]))
    body = types.SimpleNamespace(grid=g, ncells=n, shedding=sh, nsheddings=sh.shape[1], strength=G,
                                 element=fp.VortexRing, fields={"Gamma": G, "U": U, "Da": Da, "Db": Db, "name": "x"})
    files = fp.vtk.save_base(body, "wing", str(tmp_path)).split(";")
    r = _read_legacy(files[0])
    assert np.array_equal(r["points"], g.nodes)
    assert r["cells"] == [list(g.cell(c)) for c in range(n)] and set(r["types"]) == {9}
    assert np.array_equal(r["cell_data"]["Gamma"], G) and np.array_equal(r["cell_data"]["U"], U)
    # horseshoes: 4-point lines (far a, a, b, far b), Gamma_pi - Gamma_pj
    w = _read_legacy(fp.vtk.save_wake(body, "wing", str(tmp_path), wake_len=2.0))
    ne = sh.shape[1]
    assert w["types"] == [4] * ne
    exp = G[sh[0]] - np.where(sh[3] != -1, G[np.maximum(sh[3], 0)], 0.0)
    assert np.allclose(w["cell_data"]["Gamma"], exp, rtol=0, atol=0)
    for e in range(ne):
        far_a, a, b, far_b = (w["points"][:, i] for i in w["cells"][e])
        cell = g.cell(sh[0, e])
        assert np.array_equal(a, g.nodes[:, cell[sh[1, e]]]) and np.array_equal(b, g.nodes[:, cell[sh[2, e]]])
        assert np.allclose(far_a - a, 2.0 * Da[:, e]) and np.allclose(far_b - b, 2.0 * Db[:, e])
    # wake as two sheets of panels: the upper one with mu = Gamma_pi, the lower reversed with Gamma_pj
    up, lo = fp.vtk.save_wake(body, "wing", str(tmp_path), wake_panel=True).split(";")
    ru, rl = _read_legacy(up), _read_legacy(lo)
    assert np.array_equal(ru["cell_data"]["mu"], G[sh[0]])
    assert all(c1 == c2[::-1] for c1, c2 in zip(ru["cells"], rl["cells"]))
    assert os.path.basename(up) == "wing_wake_up.vtk"
