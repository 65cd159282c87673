"""Mesh input: VTK unstructured grids (.vtu) and Gmsh 4.1 ASCII meshes (.msh).

The reference reads meshes through GeoIO/Meshes into GeometricTools'
``GridTriangleSurface`` (test/runtests_meshes.jl:209-213, examples/cessna.jl);
here they become a ``Grid`` (nodes 3 x nn, CSR cells, 0-based) directly.

* ``read_vtu``  -- XML UnstructuredGrid, raw appended data (optionally
  zlib-compressed with a UInt32/UInt64 block header) or ascii arrays; cells of
  VTK types triangle (5), quad (9) and polygon (7).
* ``read_msh``  -- Gmsh 4.1 ASCII: $Nodes and $Elements blocks, 3-node
  triangles (type 2) and 4-node quads (type 3) of the 2-D entities.
* ``read_msh_nodes`` -- every node of a Gmsh 4.1 ASCII file, 3 x n (a
  trailing-edge line mesh read as points, gt.vertices2nodes of the reference's
  examples/blendedwingbody.jl:84).
"""
import re
import zlib

import numpy as np

from .geometry import Grid

_VTK_DT = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "Int32": np.int32,
           "UInt64": np.uint64, "UInt32": np.uint32, "UInt8": np.uint8, "Int8": np.int8}
_SURFACE_TYPES = (5, 7, 9)  # VTK_TRIANGLE, VTK_POLYGON, VTK_QUAD


def _appended_array(data, off, dtype, compressed, htype):
    hsize = np.dtype(htype).itemsize
    if not compressed:
        nbytes = int(np.frombuffer(data[off:off + hsize], htype)[0])
        return np.frombuffer(data[off + hsize:off + hsize + nbytes], dtype)
    nb = int(np.frombuffer(data[off:off + hsize], htype)[0])
    hdr = np.frombuffer(data[off:off + hsize * (3 + nb)], htype)
    pos = off + hsize * (3 + nb)
    chunks = []
    for s in hdr[3:3 + nb]:
        chunks.append(zlib.decompress(data[pos:pos + int(s)]))
        pos += int(s)
    return np.frombuffer(b"".join(chunks), dtype)


def read_vtu(path):
    """Surface cells of a .vtu file as a ``Grid`` (non-surface cells are skipped)."""
    raw = open(path, "rb").read()
    head, sep, tail = raw.partition(b"<AppendedData")
    txt = head.decode("utf-8", errors="replace")
    compressed = "vtkZLibDataCompressor" in txt
    htype = np.uint64 if 'header_type="UInt64"' in txt else np.uint32
    data = b""
    if sep:
        data = tail[tail.index(b"_") + 1:]
    arrays = {}
    for m in re.finditer(r"<DataArray([^>]*?)(/>|>(.*?)</DataArray>)", txt, flags=re.S):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        name, typ = attrs.get("Name"), attrs.get("type")
        if name is None or typ not in _VTK_DT:
            continue
        if attrs.get("format") == "appended":
            arrays[name] = _appended_array(data, int(attrs["offset"]), _VTK_DT[typ], compressed, htype)
        elif attrs.get("format") == "ascii":
            arrays[name] = np.array(m.group(3).split(), dtype=_VTK_DT[typ])
        else:
            raise ValueError("unsupported DataArray format %r" % attrs.get("format"))
    pts = np.asarray(arrays["Points"], dtype=np.float64).reshape(-1, 3).T
    conn = np.asarray(arrays["connectivity"], dtype=np.int64)
    offs = np.asarray(arrays["offsets"], dtype=np.int64)
    types = np.asarray(arrays.get("types", np.full(len(offs), 7)), dtype=np.int64)
    start = np.concatenate([[0], offs[:-1]])
    cells = [conn[a:b] for a, b, t in zip(start, offs, types) if t in _SURFACE_TYPES]
    return Grid(pts, cells)


def _msh_nodes(lines, k):
    nblocks = int(lines[k + 1].split()[0])
    k += 2
    tags, xyz = [], []
    for _ in range(nblocks):
        _, _, _, nn = (int(v) for v in lines[k].split())
        k += 1
        tags += [int(lines[k + i]) for i in range(nn)]
        xyz += [[float(v) for v in lines[k + nn + i].split()[:3]] for i in range(nn)]
        k += 2 * nn
    return tags, xyz, k


def read_msh_nodes(path):
    """All nodes of a Gmsh 4.1 ASCII file, 3 x n, in tag order."""
    lines = open(path).read().split("\n")
    for k, ln in enumerate(lines):
        if ln.strip() == "$Nodes":
            tags, xyz, _ = _msh_nodes(lines, k)
            order = np.argsort(tags)
            return np.asfortranarray(np.asarray(xyz, dtype=np.float64)[order].T)
    raise ValueError("no $Nodes block")


def read_msh(path):
    """2-D elements (triangles, quads) of a Gmsh 4.1 ASCII file as a ``Grid``."""
    lines = open(path).read().split("\n")
    k = 0
    tags, xyz, cells = [], [], []
    while k < len(lines):
        ln = lines[k].strip()
        if ln == "$MeshFormat":
            ver = lines[k + 1].split()
            if not ver[0].startswith("4") or ver[1] != "0":
                raise ValueError("only Gmsh 4.x ASCII meshes are supported")
            k += 2
        elif ln == "$Nodes":
            t, c, k = _msh_nodes(lines, k)
            tags += t
            xyz += c
        elif ln == "$Elements":
            nblocks = int(lines[k + 1].split()[0])
            k += 2
            for _ in range(nblocks):
                dim, _, etype, ne = (int(v) for v in lines[k].split())
                k += 1
                for i in range(ne):
                    vals = [int(v) for v in lines[k + i].split()]
                    if dim == 2 and etype in (2, 3):
                        cells.append(vals[1:])
                k += ne
        else:
            k += 1
    if not cells:
        raise ValueError("no triangles or quads found")
    order = np.argsort(tags)
    tag_sorted = np.asarray(tags)[order]
    nodes = np.asarray(xyz, dtype=np.float64)[order].T
    index = {int(t): i for i, t in enumerate(tag_sorted)}
    cells = [np.array([index[t] for t in c], dtype=np.int64) for c in cells]
    return Grid(nodes, cells)
