"""Panel grids and deterministic synthetic geometry generators.

The reference builds its grids with the un-vendored GeometricTools package
(``gt.generate_loft``, ``gt.surface_revolution``, ``gt.GridTriangleSurface``)
and wraps them in bodies (src/FLOWPanel_abstractliftingbody.jl:73-176,
src/FLOWPanel_nonliftingbody.jl:827-858).  This module provides the grid
container the bodies need and deterministic generators for the workloads the
hot path is measured on:

* ``sphere``            -- body of revolution about x, pole rows as triangles
                           (a repeated vertex would make an edge of zero length,
                           which the Hess-Smith kernel divides by,
                           src/FLOWPanel_elements.jl:101-104)
* ``naca4``             -- NACA 4-digit section, cosine spacing, sharp trailing edge
* ``loft_wing``         -- structured quad loft with sweep, taper, twist and the
                           trailing-edge shedding list of ``generate_loft_liftbody``
                           (src/FLOWPanel_abstractliftingbody.jl:88-104), for quads
* ``bwb``               -- blended-wing-body loft of random NACA sections
* ``triangulate``       -- split every quad in two, as ``gt.GridTriangleSurface``
                           does, keeping the structured index pattern the
                           reference's Kutta averaging assumes
                           (src/FLOWPanel_postprocess.jl:1075-1097)

Layout follows the reference: nodes are a column-major 3 x nn float64 array;
cells are a CSR list of 0-based node indices (``cell_ptr``, ``cell_idx``);
the shedding matrix is 6 x ns ``(pi, nia, nib, pj, nja, njb)`` with 0-based
panel and local-vertex indices and ``pj = -1`` for a one-sided edge.
"""
import numpy as np


class Grid:
    """Unstructured panel grid (nodes + CSR cells), optionally structured.

    ``dims = (n1, n2)`` records the cell counts of a structured grid whose
    linear cell index is ``i + n1*j`` (first dimension fastest, as Julia's
    column-major ``LinearIndices``).
    """

    def __init__(self, nodes, cells, dims=None):
        nodes = np.asarray(nodes, dtype=np.float64)
        if nodes.ndim != 2 or nodes.shape[0] != 3:
            raise ValueError("nodes must be a 3 x nn array, got %s" % (nodes.shape,))
        self.nodes = np.asfortranarray(nodes)
        if isinstance(cells, tuple) and len(cells) == 2:
            ptr, idx = cells
            self.cell_ptr = np.ascontiguousarray(ptr, dtype=np.int64)
            self.cell_idx = np.ascontiguousarray(idx, dtype=np.int64)
        else:
            cells = [np.asarray(c, dtype=np.int64) for c in cells]
            self.cell_ptr = np.zeros(len(cells) + 1, dtype=np.int64)
            self.cell_ptr[1:] = np.cumsum([len(c) for c in cells])
            self.cell_idx = np.concatenate(cells) if cells else np.zeros(0, np.int64)
        nv = np.diff(self.cell_ptr)
        if len(nv) and nv.min() < 3:
            raise ValueError("every cell needs at least 3 vertices")
        if self.cell_idx.size and (self.cell_idx.min() < 0 or self.cell_idx.max() >= self.nnodes):
            raise ValueError("cell vertex index out of range")
        self.dims = dims

    @property
    def nnodes(self):
        return self.nodes.shape[1]

    @property
    def ncells(self):
        return len(self.cell_ptr) - 1

    def cell(self, i):
        return self.cell_idx[self.cell_ptr[i]:self.cell_ptr[i + 1]]

    def nvertices(self):
        return np.diff(self.cell_ptr)


# ----------------------------------------------------------------------------- sphere
def sphere(radius=1.0, n_meridional=24, n_azimuthal=48, center=(0.0, 0.0, 0.0)):
    """Sphere as a body of revolution about the x axis.

    ``n_meridional`` rows from x=-R to x=+R (the two pole rows are triangles),
    ``n_azimuthal`` cells around; normals point outward.  The unit sphere with
    24 x 48 gives 1152 panels.
    """
    nm, na = int(n_meridional), int(n_azimuthal)
    if nm < 2 or na < 3:
        raise ValueError("sphere needs n_meridional >= 2 and n_azimuthal >= 3")
    th = np.pi * np.arange(nm + 1) / nm
    ph = 2 * np.pi * np.arange(na) / na
    pts = [np.array([-radius, 0.0, 0.0])]
    for k in range(1, nm):
        for m in range(na):
            pts.append(np.array([-radius * np.cos(th[k]), radius * np.sin(th[k]) * np.cos(ph[m]),
                                 radius * np.sin(th[k]) * np.sin(ph[m])]))
    pts.append(np.array([radius, 0.0, 0.0]))
    nodes = np.array(pts).T + np.asarray(center, float).reshape(3, 1)

    def ring(k, m):  # k = 1..nm-1
        return 1 + (k - 1) * na + (m % na)

    last = len(pts) - 1
    cells = []
    for k in range(nm):
        for m in range(na):
            if k == 0:
                cells.append([0, ring(1, m + 1), ring(1, m)])
            elif k == nm - 1:
                cells.append([ring(k, m), ring(k, m + 1), last])
            else:
                cells.append([ring(k, m), ring(k, m + 1), ring(k + 1, m + 1), ring(k + 1, m)])
    # cell order: azimuth fastest within a meridional row
    return Grid(nodes, cells, dims=(na, nm))


# ----------------------------------------------------------------------------- airfoils
def naca4(digits="0012", n_panels=60, closed_te=True):
    """NACA 4-digit section with cosine spacing.

    Returns the 2 x (n_panels+1) contour (x, z) ordered lower trailing edge ->
    leading edge -> upper trailing edge (the two trailing-edge points coincide
    when ``closed_te``).  ``digits`` may be a string ``"mptt"`` or a tuple of
    fractions ``(m, p, t)``.
    """
    if isinstance(digits, str):
        m, p, t = int(digits[0]) / 100.0, int(digits[1]) / 10.0, int(digits[2:]) / 100.0
    else:
        m, p, t = digits
    if n_panels % 2:
        raise ValueError("n_panels must be even (half on each side)")
    nh = n_panels // 2
    x = 0.5 * (1 - np.cos(np.linspace(0.0, np.pi, nh + 1)))
    a4 = -0.1036 if closed_te else -0.1015
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 + a4 * x ** 4)
    if m > 0 and 0 < p < 1:
        yc = np.where(x < p, m / p ** 2 * (2 * p * x - x ** 2),
                      m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * x - x ** 2))
        dyc = np.where(x < p, 2 * m / p ** 2 * (p - x), 2 * m / (1 - p) ** 2 * (p - x))
    else:
        yc = np.zeros_like(x)
        dyc = np.zeros_like(x)
    th = np.arctan(dyc)
    xu, zu = x - yt * np.sin(th), yc + yt * np.cos(th)
    xl, zl = x + yt * np.sin(th), yc - yt * np.cos(th)
    if closed_te:
        zu[-1] = zl[-1] = yc[-1]
        xu[-1] = xl[-1] = x[-1]
    cx = np.concatenate([xl[::-1], xu[1:]])
    cz = np.concatenate([zl[::-1], zu[1:]])
    return np.vstack([cx, cz])


# ----------------------------------------------------------------------------- lofts
def _loft(sections, y, x_le, z_le, chord, twist_deg, xref=0.25):
    """Structured loft: ``sections[j]`` is the 2 x (nc+1) contour at station j."""
    ns1 = len(y)
    nc1 = sections[0].shape[1]
    nodes = np.zeros((3, nc1 * ns1))
    for j in range(ns1):
        cx, cz = sections[j]
        th = np.deg2rad(twist_deg[j])
        xs = (cx - xref) * np.cos(th) + cz * np.sin(th) + xref
        zs = -(cx - xref) * np.sin(th) + cz * np.cos(th)
        sl = slice(j * nc1, (j + 1) * nc1)
        nodes[0, sl] = x_le[j] + chord[j] * xs
        nodes[1, sl] = y[j]
        nodes[2, sl] = z_le[j] + chord[j] * zs
    nc, ns = nc1 - 1, ns1 - 1
    cells = np.zeros((nc * ns, 4), dtype=np.int64)
    shedding = np.zeros((6, ns), dtype=np.int64)
    for j in range(ns):
        for i in range(nc):
            a = i + nc1 * j
            cells[i + nc * j] = [a, a + 1, a + 1 + nc1, a + nc1]
        # upper TE cell (nc-1, j): its ring runs v1 -> v2 along the trailing
        # edge; lower TE cell (0, j): v3 -> v0.  The horseshoe's bound segment
        # runs against the ring (TE = (v2, v1) upper, (v0, v3) lower).
        shedding[:, j] = [(nc - 1) + nc * j, 2, 1, nc * j, 0, 3]
    ptr = np.arange(0, 4 * nc * ns + 1, 4, dtype=np.int64)
    return Grid(nodes, (ptr, cells.reshape(-1)), dims=(nc, ns)), shedding


def _span_stations(b, n_span, spacing):
    if spacing == "uniform":
        eta = np.linspace(-1.0, 1.0, n_span + 1)
    elif spacing == "cosine":
        eta = -np.cos(np.linspace(0.0, np.pi, n_span + 1))
    else:
        raise ValueError("unknown spacing " + str(spacing))
    return 0.5 * b * eta


def loft_wing(airfoil="0012", n_chord=60, n_span=80, span=7.5, root_chord=1.0, taper=0.5,
              sweep_deg=30.0, twist_root_deg=0.0, twist_tip_deg=0.0, dihedral_deg=0.0,
              span_spacing="uniform"):
    """Lofted wing from left tip (y=-b/2) to right tip (y=+b/2), open tips.

    ``n_chord`` panels around the section, ``n_span`` panels across the full
    span.  Returns ``(grid, shedding)``.  Leading-edge sweep and dihedral use
    |y|; chord varies linearly from root to tip; twist varies linearly in |y|
    about the quarter chord.
    """
    y = _span_stations(span, n_span, span_spacing)
    eta = np.abs(y) / (0.5 * span)
    chord = root_chord * (1 - (1 - taper) * eta)
    x_le = np.abs(y) * np.tan(np.deg2rad(sweep_deg))
    z_le = np.abs(y) * np.tan(np.deg2rad(dihedral_deg))
    twist = twist_root_deg + (twist_tip_deg - twist_root_deg) * eta
    sec = naca4(airfoil, n_chord)
    return _loft([sec] * len(y), y, x_le, z_le, chord, twist)


def bwb(seed=5, n_sections=12, n_chord=250, n_span=120, span=10.0):
    """Synthetic blended-wing-body loft.

    ``n_sections`` defining stations across the full span carry random NACA
    4-digit sections (thickness U[8,18] %, camber U[0,4] % at 40 % chord) drawn
    from ``numpy.random.Generator(PCG64(seed))``; the planform has a long
    centre body blending into a swept outer wing.  Between defining stations
    the contour, chord, leading edge and twist are interpolated linearly.
    ``n_chord`` x ``n_span`` quads (250 x 120 = 30,000 panels).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    eta_def = np.linspace(-1.0, 1.0, n_sections)
    thick = rng.uniform(0.08, 0.18, n_sections)
    camber = rng.uniform(0.0, 0.04, n_sections)
    a = np.abs(eta_def)
    chord_def = 0.25 + 0.75 * np.exp(-(a / 0.22) ** 2)           # centre body -> outer wing
    xle_def = 0.55 * span * 0.5 * a * np.tan(np.deg2rad(35.0)) / 0.55 * (0.4 + 0.6 * a)
    twist_def = -3.0 * a
    contours = [naca4((camber[k], 0.4, thick[k]), n_chord) for k in range(n_sections)]
    y = _span_stations(span, n_span, "uniform")
    eta = y / (0.5 * span)
    sections, chord, xle, twist = [], [], [], []
    for e in eta:
        k = min(int(np.searchsorted(eta_def, e, side="right")) - 1, n_sections - 2)
        k = max(k, 0)
        w = (e - eta_def[k]) / (eta_def[k + 1] - eta_def[k])
        sections.append((1 - w) * contours[k] + w * contours[k + 1])
        chord.append((1 - w) * chord_def[k] + w * chord_def[k + 1])
        xle.append((1 - w) * xle_def[k] + w * xle_def[k + 1])
        twist.append((1 - w) * twist_def[k] + w * twist_def[k + 1])
    chord = np.array(chord) * 3.0
    return _loft(sections, y, np.array(xle), np.zeros_like(y), chord, np.array(twist))


def triangulate(grid, shedding=None):
    """Split every quad of a structured loft into two triangles.

    Quad (i, j) of a (n1, n2) grid becomes triangles (2i, j) = (v0, v1, v2) and
    (2i+1, j) = (v0, v2, v3) of a (2 n1, n2) grid, so the upper trailing-edge
    quad maps to cells (pi, pi+1) and the lower one to (pj-1, pj) as the
    reference's Kutta averaging assumes (src/FLOWPanel_postprocess.jl:1083-1087).
    Returns ``(grid, shedding)``.
    """
    n1, n2 = grid.dims
    nv = grid.nvertices()
    if not np.all(nv == 4):
        raise ValueError("triangulate needs an all-quad grid")
    q = grid.cell_idx.reshape(-1, 4)
    tri = np.zeros((2 * len(q), 3), dtype=np.int64)
    for j in range(n2):
        for i in range(n1):
            v = q[i + n1 * j]
            tri[2 * i + 2 * n1 * j] = [v[0], v[1], v[2]]
            tri[2 * i + 1 + 2 * n1 * j] = [v[0], v[2], v[3]]
    ptr = np.arange(0, 3 * len(tri) + 1, 3, dtype=np.int64)
    g = Grid(grid.nodes, (ptr, tri.reshape(-1)), dims=(2 * n1, n2))
    if shedding is None:
        return g, None
    sh = np.zeros_like(shedding)
    for e in range(shedding.shape[1]):
        pu, pl = shedding[0, e], shedding[3, e]
        iu, ju = pu % n1, pu // n1
        il, jl = pl % n1, pl // n1
        # upper quad edge v1 -> v2 lies in triangle (v0, v1, v2) as local 1 -> 2
        # lower quad edge v3 -> v0 lies in triangle (v0, v2, v3) as local 2 -> 0
        sh[:, e] = [2 * iu + 2 * n1 * ju, 2, 1, 2 * il + 1 + 2 * n1 * jl, 0, 2]
    return g, sh


def calc_shedding(grid, trailingedge, tolerance=None, periodic=False):
    """Shedding matrix (6 x ne, 0-based, pj = -1 when single-sided) of the mesh
    edges lying on a trailing-edge polyline (src/FLOWPanel_liftingbody.jl:942-1043).

    ``trailingedge`` is 3 x k points ordered along the edge.  Nodes within
    ``tolerance`` of the polyline are ordered by their position along it;
    consecutive pairs that are mesh edges become shedding edges.  The panel in
    which the pair appears in the order (first, second) is ``pi`` (nia = local
    index of the second node, nib of the first); the panel on the other side is
    ``pj`` (nja = first, njb = second).
    """
    X = np.asarray(grid.nodes, dtype=np.float64)
    T = np.asarray(trailingedge, dtype=np.float64)
    if T.ndim != 2 or T.shape[0] != 3 or T.shape[1] < 2:
        raise ValueError("trailingedge must be 3 x k (k >= 2)")
    if tolerance is None:
        tolerance = 1e2 * np.finfo(float).eps
    # distance of every node to the polyline and its arc-length position
    seg0, seg1 = T[:, :-1], T[:, 1:]
    d = seg1 - seg0
    L = np.linalg.norm(d, axis=0)
    s0 = np.concatenate([[0.0], np.cumsum(L)[:-1]])
    best = np.full(X.shape[1], np.inf)
    pos = np.zeros(X.shape[1])
    for k in range(seg0.shape[1]):
        if L[k] == 0:
            continue
        t = np.clip(((X - seg0[:, k:k + 1]) * d[:, k:k + 1]).sum(axis=0) / L[k] ** 2, 0.0, 1.0)
        dist = np.linalg.norm(X - (seg0[:, k:k + 1] + t * d[:, k:k + 1]), axis=0)
        better = dist < best
        best[better] = dist[better]
        pos[better] = s0[k] + t[better] * L[k]
    te = np.nonzero(best <= tolerance)[0]
    te = te[np.argsort(pos[te], kind="stable")]
    if len(te) == 0:
        return np.zeros((6, 0), dtype=np.int64)
    if periodic:
        te = np.append(te, te[0])
    # directed edges of the cells
    owner = {}
    for c in range(grid.ncells):
        cell = grid.cell(c)
        for k in range(len(cell)):
            owner.setdefault((int(cell[k]), int(cell[(k + 1) % len(cell)])), []).append(c)
    cols = []
    for a, b in zip(te[:-1], te[1:]):
        a, b = int(a), int(b)
        fwd, bwd = owner.get((a, b), []), owner.get((b, a), [])
        if not fwd and not bwd:
            continue
        if fwd and bwd:
            pi, pj = fwd[0], bwd[0]
        else:
            pi, pj = (fwd or bwd)[0], -1
        ci = list(grid.cell(pi))
        nia, nib = ci.index(b), ci.index(a)
        if pj != -1:
            cj = list(grid.cell(pj))
            nja, njb = cj.index(a), cj.index(b)
        else:
            nja = njb = -1
        cols.append((pi, nia, nib, pj, nja, njb))
    return np.array(cols, dtype=np.int64).T.reshape(6, -1)


def transform(grid, rotation=None, translation=(0.0, 0.0, 0.0), scaling=1.0):
    """A copy of the grid with its nodes translated, then rotated (3 x 3), then scaled
    (the Meshes.Translate |> Rotate |> Scale chain of examples/blendedwingbody.jl:78)."""
    X = np.asarray(grid.nodes, dtype=np.float64) + np.asarray(translation, dtype=np.float64).reshape(3, 1)
    if rotation is not None:
        X = np.asarray(rotation, dtype=np.float64) @ X
    return Grid(X * scaling, (grid.cell_ptr, grid.cell_idx), dims=grid.dims)


def mirror(grid, coordinate, tol=1e-12):
    """The grid joined with its mirror image across the plane ``x[coordinate] = 0`` (0-based coordinate),
    the mirrored cells' vertex order reversed so that their normals keep pointing out, and the nodes that
    lie on the plane shared by both halves (the role of gt.mirror in examples/blendedwingbody.jl:82)."""
    X = np.asarray(grid.nodes, dtype=np.float64)
    nn = X.shape[1]
    Xm = X.copy()
    Xm[coordinate] *= -1
    on = np.abs(X[coordinate]) <= tol * max(1.0, np.abs(X).max())
    # mirrored node i -> itself when on the plane, else a new node
    new_index = np.where(on, np.arange(nn), nn + np.cumsum(~on) - 1)
    nodes = np.hstack([X, Xm[:, ~on]])
    cells = [grid.cell(c) for c in range(grid.ncells)]
    cells += [new_index[np.asarray(c)][::-1] for c in cells]
    return Grid(nodes, cells)


def spatial_order(grid, shedding=None, bits=10):
    """Cells reordered along a Morton (Z-order) curve of their centroids, so that every run of consecutive
    indices is a compact patch of the surface: the block-Jacobi blocks and coarse chunks of the Krylov
    solvers, which take consecutive index ranges, then group neighbouring panels on unstructured meshes.
    Returns ``(grid, shedding, perm)``: new cell k is old cell ``perm[k]`` (strengths map back with
    ``old[perm] = new``)."""
    n = grid.ncells
    counts = np.diff(grid.cell_ptr)
    cent = np.add.reduceat(grid.nodes[:, grid.cell_idx], grid.cell_ptr[:-1], axis=1) / counts
    lo, hi = cent.min(axis=1, keepdims=True), cent.max(axis=1, keepdims=True)
    q = np.floor((cent - lo) / np.maximum(hi - lo, 1e-300) * ((1 << bits) - 1)).astype(np.uint64)
    key = np.zeros(n, dtype=np.uint64)
    for b in range(bits):
        for d in range(3):
            key |= ((q[d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
    perm = np.argsort(key, kind="stable")
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    cells = [grid.cell(c) for c in perm]
    g = Grid(grid.nodes, cells)
    if shedding is None:
        return g, None, perm
    sh = np.array(shedding, dtype=np.int64)
    sh[0] = inv[sh[0]]
    sh[3] = np.where(sh[3] >= 0, inv[np.maximum(sh[3], 0)], -1)
    return g, sh, perm


def _sections_spacing(sections):
    """Normalised node positions in [0, 1] of a multi-section discretisation: each
    section ``(fraction, ncells, ratio, central)`` covers ``fraction`` of the length
    with ``ncells`` cells whose lengths grow geometrically by ``ratio`` (last over
    first) from its start to its end, or with ``central`` from both ends to its
    middle (the (sec, ndivs, expratio, central) convention the reference's
    examples pass to GeometricTools, restated; that package is not vendored)."""
    pos = [0.0]
    for frac, n, ratio, central in sections:
        n = int(n)
        if central:
            h1 = n // 2
            halves = [(h1, ratio), (n - h1, 1.0 / ratio)]
            lens = []
            for nh, r in halves:
                if nh == 0:
                    continue
                q = r ** (1.0 / (nh - 1)) if nh > 1 else 1.0
                lens += [q ** i for i in range(nh)]
            lens = np.array(lens)
        else:
            q = ratio ** (1.0 / (n - 1)) if n > 1 else 1.0
            lens = q ** np.arange(n)
        lens = frac * lens / lens.sum()
        pos += list(pos[-1] + np.cumsum(lens))
    pos = np.array(pos)
    return pos / pos[-1]


def rediscretize_airfoil(x, y, ndivs_up, ndivs_lo):
    """The closed airfoil contour (x, y) (a loop from the trailing edge round the
    leading edge, the point of least x, back to the trailing edge) re-sampled on
    each side by arc length with the multi-section spacing ``ndivs_*`` (see
    ``_sections_spacing``) through a cubic spline, as gt.rediscretize_airfoil in
    examples/duct.jl:78.  Returns (xs, ys), the first side first, leading-edge
    point shared."""
    from scipy.interpolate import CubicSpline
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    le = int(np.argmin(x))
    sides = []
    for k, (sx, sy, nd) in enumerate(((x[:le + 1], y[:le + 1], ndivs_up), (x[le:], y[le:], ndivs_lo))):
        # each side runs from the leading edge to the trailing edge
        if k == 0:
            sx, sy = sx[::-1], sy[::-1]
        t = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(sx), np.diff(sy)))])
        keep = np.concatenate([[True], np.diff(t) > 0])
        t, sx, sy = t[keep], sx[keep], sy[keep]
        u = _sections_spacing(nd) * t[-1]
        sides.append((CubicSpline(t, sx)(u), CubicSpline(t, sy)(u)))
    (ux, uy), (lx, ly) = sides
    xs = np.concatenate([ux[::-1], lx[1:]])
    ys = np.concatenate([uy[::-1], ly[1:]])
    return xs, ys


def revolution_liftbody(points, n_theta, axis_angle=270.0):
    """Lifting body of revolution about the x axis (generate_revolution_liftbody,
    src/FLOWPanel_abstractliftingbody.jl:124-167): the closed contour ``points``
    (k x 2: axial position, radius; first point = last = the trailing edge)
    revolved in ``n_theta`` azimuthal steps from ``axis_angle`` degrees.  A
    (contour cells) x (n_theta) quad grid like the lofts (contour dimension first,
    the trailing-edge nodes of the two sides separate), periodic in azimuth, normals
    pointing out of the closed contour's solid, the trailing edge of every azimuthal
    strip shedding between the last and the first contour cells.  Returns
    ``(grid, shedding)``; ``triangulate`` gives the reference's triangle grid."""
    P = np.asarray(points, dtype=np.float64)
    if P.ndim != 2 or P.shape[1] != 2:
        raise ValueError("points must be k x 2 (x, r)")
    nc1 = P.shape[0]
    nc = nc1 - 1
    nt = int(n_theta)
    th = np.deg2rad(axis_angle) + 2 * np.pi * np.arange(nt) / nt
    # orientation: the contour's signed area in the (x, r) plane decides the sense of revolution
    area = 0.5 * np.sum(P[:-1, 0] * P[1:, 1] - P[1:, 0] * P[:-1, 1])
    if area < 0:
        th = th[::-1].copy()
    nodes = np.zeros((3, nc1 * nt))
    for j in range(nt):
        sl = slice(j * nc1, (j + 1) * nc1)
        nodes[0, sl] = P[:, 0]
        nodes[1, sl] = P[:, 1] * np.cos(th[j])
        nodes[2, sl] = P[:, 1] * np.sin(th[j])
    cells = np.zeros((nc * nt, 4), dtype=np.int64)
    shedding = np.zeros((6, nt), dtype=np.int64)
    for j in range(nt):
        jn = (j + 1) % nt
        for i in range(nc):
            cells[i + nc * j] = [i + nc1 * j, i + 1 + nc1 * j, i + 1 + nc1 * jn, i + nc1 * jn]
        shedding[:, j] = [(nc - 1) + nc * j, 2, 1, nc * j, 0, 3]
    ptr = np.arange(0, 4 * nc * nt + 1, 4, dtype=np.int64)
    return Grid(nodes, (ptr, cells.reshape(-1)), dims=(nc, nt)), shedding


def direction(alpha_deg, beta_deg=0.0):
    """Freestream unit vector (cos a cos b, -sin b, sin a cos b)
    (src/FLOWPanel_utils.jl:34-37)."""
    a, b = np.deg2rad(alpha_deg), np.deg2rad(beta_deg)
    return np.array([np.cos(a) * np.cos(b), -np.sin(b), np.sin(a) * np.cos(b)])
