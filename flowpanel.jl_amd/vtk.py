"""VTK output of bodies, control points and wakes (legacy ASCII .vtk, one file each).

Mirrors the reference's output functions: ``save(body, filename; out_wake,
wake_len, wake_panel, out_controlpoints)`` (src/FLOWPanel_liftingbody.jl:146-163,
src/FLOWPanel_nonliftingbody.jl:78-84, src/FLOWPanel_multibody.jl:264-273),
``save_base`` (src/FLOWPanel_abstractbody.jl:136-190), ``save_controlpoints``
(:192-262) and ``_savewake`` (src/FLOWPanel_liftingbody.jl:1092-1196).  The
body file carries every cell field the body holds (a length-N array as a
scalar, a 3 x N array as a vector); the wake is written as horseshoe lines
(field "Gamma" = Gamma_pi - Gamma_pj) or as two sheets of doublet panels of
length ``wake_len`` along Da/Db (field "mu", one file per side).  Paths of the
files written are returned joined by ';', like the reference.
"""
import os

import numpy as np

from .bodies import calc_controlpoints, calc_normals

_VTK_LINE, _VTK_TRIANGLE, _VTK_POLYGON, _VTK_QUAD, _VTK_POLY_LINE = 3, 5, 7, 9, 4


def _write(path, points, cells=(), cell_types=(), cell_data=None, point_data=None):
    points = np.asarray(points, dtype=np.float64)
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 4.0\nFLOWPanel\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write("POINTS %d double\n" % points.shape[1])
        for p in points.T:
            f.write("%.17g %.17g %.17g\n" % tuple(p))
        ncells = len(cells)
        size = sum(len(c) + 1 for c in cells)
        f.write("\nCELLS %d %d\n" % (ncells, size))
        for c in cells:
            f.write("%d %s\n" % (len(c), " ".join(str(int(v)) for v in c)))
        f.write("\nCELL_TYPES %d\n" % ncells)
        for t in cell_types:
            f.write("%d\n" % t)
        for kind, data, count in (("CELL_DATA", cell_data, ncells), ("POINT_DATA", point_data, points.shape[1])):
            if not data:
                continue
            f.write("\n%s %d\n" % (kind, count))
            for name, arr in data.items():
                arr = np.asarray(arr, dtype=np.float64)
                if arr.ndim == 1:
                    f.write("SCALARS %s double\nLOOKUP_TABLE default\n" % name)
                    f.write("\n".join("%.17g" % v for v in arr) + "\n")
                else:
                    f.write("VECTORS %s double\n" % name)
                    for v in arr.T:
                        f.write("%.17g %.17g %.17g\n" % tuple(v))
    return path


def _cell_type(nv):
    return {3: _VTK_TRIANGLE, 4: _VTK_QUAD}.get(nv, _VTK_POLYGON)


def _cell_fields(body):
    out = {}
    n = body.ncells
    for name, val in body.fields.items():
        a = np.asarray(val)
        if a.dtype.kind not in "fiu":
            continue
        if a.shape == (n,):
            out[name] = a
        elif a.shape == (3, n):
            out[name] = a
    out.setdefault(body.element.field, body.strength)
    return out


def save_base(body, filename, path=".", suffix="", out_cellindex=False, out_controlpoints=False):
    """The body's grid with its cell fields (and, with ``out_controlpoints``, the control points)."""
    g = body.grid
    cells = [g.cell(c) for c in range(g.ncells)]
    data = _cell_fields(body)
    if out_cellindex:
        data["cellindex"] = np.arange(1, g.ncells + 1, dtype=np.float64)
    files = [_write(os.path.join(path, filename + suffix + ".vtk"), g.nodes, cells, [_cell_type(len(c)) for c in cells],
                    cell_data=data)]
    if out_controlpoints:
        files.append(save_controlpoints(body, filename, path))
    return ";".join(files)


def save_controlpoints(body, filename, path=".", suffix="_cp"):
    """Control points as VTK vertices with the normals as point data."""
    cps = calc_controlpoints(body)
    n = cps.shape[1]
    return _write(os.path.join(path, filename + suffix + ".vtk"), cps, [[i] for i in range(n)], [1] * n,
                  point_data={"normal": calc_normals(body)})


def _wake_strengths(body):
    G = np.asarray(body.strength)
    sh = body.shedding
    s1 = G[sh[0]]
    s2 = np.where(sh[3] != -1, G[np.maximum(sh[3], 0)], 0.0)
    return s1, s2


def save_wake(body, filename, path=".", wake_len=1.0, wake_panel=False, suffix="_wake"):
    """The rigid wake: horseshoes as 4-point lines (Gamma_pi - Gamma_pj), or two sheets of panels (mu)."""
    if "Da" not in body.fields or "Db" not in body.fields:
        raise ValueError("Requested to save wake, but Da field was not found. Run solve(...) first")
    sh = body.shedding
    ne = sh.shape[1]
    g = body.grid
    Da, Db = np.asarray(body.fields["Da"]), np.asarray(body.fields["Db"])
    P = np.zeros((3, 4 * ne))
    for e in range(ne):
        cell = g.cell(sh[0, e])
        pa, pb = g.nodes[:, cell[sh[1, e]]], g.nodes[:, cell[sh[2, e]]]
        P[:, e], P[:, ne + e] = pa, pb
        P[:, 2 * ne + e], P[:, 3 * ne + e] = pa + wake_len * Da[:, e], pb + wake_len * Db[:, e]
    s1, s2 = _wake_strengths(body)
    quads = [[i + 2 * ne, i, i + ne, i + 3 * ne] for i in range(ne)]
    if not wake_panel:
        return _write(os.path.join(path, filename + suffix + ".vtk"), P, quads, [_VTK_POLY_LINE] * ne,
                      cell_data={"Gamma": s1 - s2})
    files = []
    for side, (nm, mu) in enumerate((("_up", s1), ("_lo", s2))):
        cells = quads if side == 0 else [c[::-1] for c in quads]
        normals = []
        for c in cells:
            a, b, cc = P[:, c[0]], P[:, c[1]], P[:, c[2]]
            nv = np.cross(b - a, cc - a)
            normals.append(nv / np.linalg.norm(nv))
        files.append(_write(os.path.join(path, filename + suffix + nm + ".vtk"), P, cells, [_VTK_QUAD] * ne,
                            cell_data={"mu": mu, "normal": np.array(normals).T}))
    return ";".join(files)


def save(body, filename, path=".", out_wake=True, wake_len=1.0, wake_panel=False, wake_suffix="_wake",
         out_controlpoints=False, out_cellindex=False):
    """save(body, filename; out_wake, wake_len, wake_panel): the body (every component of a MultiBody as
    filename_<name>) and, for lifting bodies, its wake."""
    os.makedirs(path, exist_ok=True)
    if hasattr(body, "names"):
        return ";".join(save(b, filename + "_" + nm, path, out_wake, wake_len, wake_panel, wake_suffix,
                             out_controlpoints, out_cellindex) for nm, b in zip(body.names, body.bodies))
    files = [save_base(body, filename, path, out_cellindex=out_cellindex, out_controlpoints=out_controlpoints)]
    if out_wake and body.nsheddings:
        files.append(save_wake(body, filename, path, wake_len, wake_panel, wake_suffix))
    return ";".join(files)
