"""Body types: the host-side mirror of the reference's body API.

* ``NonLiftingBody``  -- ``NonLiftingBody{ConstantSource}(grid)``
  (src/FLOWPanel_nonliftingbody.jl:32-72)
* ``RigidWakeBody``   -- ``RigidWakeBody{VortexRing}(grid, shedding; CPoffset,
  kerneloffset, kernelcutoff, characteristiclength)`` (src/FLOWPanel_liftingbody.jl:43-134)
* ``MultiBody``       -- ``MultiBody(bodies, names)`` (src/FLOWPanel_multibody.jl:19-59),
  column-block composition of its components

A body owns a device handle (fp_body) holding its geometry, freestream, wake
directions and strengths.  Geometry queries (``calc_normals`` and friends) are
answered from the device.
"""
import ctypes

import numpy as np

from . import _native as nat


class ConstantSource:
    """Constant-strength source panel (Hess & Smith)."""
    code = nat.FP_ELEM_SOURCE
    field = "sigma"


class VortexRing:
    """Vortex ring panel (equivalent to a constant-strength doublet)."""
    code = nat.FP_ELEM_VORTEXRING
    field = "Gamma"


class ConstantDoublet(VortexRing):
    """Constant-strength doublet panel: for the velocity a vortex ring of the same
    strength (U_constant_doublet, src/FLOWPanel_elements.jl:740), with the doublet
    potential; its strength field is "mu"."""
    field = "mu"


def shedding_zero_based(shedding, index_base=0):
    """The 6 x ns shedding matrix (pi, nia, nib, pj, nja, njb) with 0-based entries.

    ``index_base=1`` takes a Julia (1-based) matrix as the reference builds it
    (src/FLOWPanel_liftingbody.jl:24-31): every entry moves down by one except the
    partner side of a one-sided edge, whose ``pj = -1`` sentinel and unused vertex
    slots keep their values.
    """
    sh = np.zeros((6, 0), dtype=np.int64) if shedding is None else np.array(shedding, dtype=np.int64)
    if sh.ndim != 2 or sh.shape[0] != 6:
        raise ValueError("shedding must be a 6 x ns matrix, got %s" % (sh.shape,))
    if index_base not in (0, 1):
        raise ValueError("index_base must be 0 or 1")
    if index_base:
        one_sided = sh[3] == -1
        sh[:3] -= index_base
        sh[3:, ~one_sided] -= index_base
    return sh


class AbstractBody:
    def __init__(self, grid, element, shedding=None, CPoffset=1e-14, kerneloffset=1e-8, kernelcutoff=1e-14,
                 characteristiclength="unitary", index_base=0, kutta=None, ctx=None):
        self.ctx = ctx or nat.context()
        self.grid = grid
        self.element = element
        self.CPoffset = float(CPoffset)
        self.kerneloffset = float(kerneloffset)
        self.kernelcutoff = float(kernelcutoff)
        if characteristiclength not in nat.CHARLEN:
            raise ValueError("characteristiclength must be one of %s" % list(nat.CHARLEN))
        self.characteristiclength = characteristiclength
        sh = shedding_zero_based(shedding, index_base)
        self.shedding = sh      # 0-based from here on (device tables, Kutta, forces)
        self.index_base = index_base
        self.fields = {}
        self.strength = np.zeros(grid.ncells)
        self._solved = False
        nodes = nat.f64(grid.nodes)
        shc = np.ascontiguousarray(sh.T) if sh.shape[1] else np.zeros((1, 6), dtype=np.int64)
        h = ctypes.c_void_p()
        self.ctx.check(nat.lib().fp_body_create(
            self.ctx.handle, nat.dptr(nodes), grid.nnodes, nat.iptr(grid.cell_ptr), nat.iptr(grid.cell_idx),
            grid.ncells, 0, element.code, nat.iptr(shc), sh.shape[1], self.CPoffset,
            nat.CHARLEN[characteristiclength], self.kerneloffset, self.kernelcutoff, ctypes.byref(h)))
        self.handle = h
        if kutta is None:
            # the reference's Kutta pattern (pi, pi+1, pj, pj-1) holds on structured triangulated grids
            # (src/FLOWPanel_postprocess.jl:1075-1097); on quads and on unstructured meshes -- where the
            # reference warns that pattern is wrong (:1077-1080) -- the two cells sharing each shedding
            # edge, pi and pj, are averaged
            nv = grid.nvertices()
            structured = getattr(grid, "dims", None) is not None
            kutta = "reference" if (structured and len(nv) and np.all(nv == 3)) else "pair"
        if kutta not in ("reference", "pair", "none"):
            raise ValueError("kutta must be 'reference', 'pair' or 'none'")
        self.kutta = kutta

    # --- reference-style properties
    @property
    def ncells(self):
        return self.grid.ncells

    @property
    def nnodes(self):
        return self.grid.nnodes

    @property
    def nsheddings(self):
        return self.shedding.shape[1]

    @property
    def bodies(self):
        return [self]

    def check_solved(self):
        return self._solved

    def add_field(self, name, data):
        self.fields[name] = data

    def get_field(self, name):
        return self.fields[name]

    def __del__(self):
        try:
            if getattr(self, "handle", None) and nat._LIB is not None:
                nat._LIB.fp_body_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class NonLiftingBody(AbstractBody):
    """NonLiftingBody{ConstantSource}(grid) (src/FLOWPanel_nonliftingbody.jl:32-130) or
    NonLiftingBody{ConstantDoublet}(grid) (src/FLOWPanel_nonliftingbody.jl:258-352): a body
    without wake; the doublet body assembles U_constant_doublet . n (a vortex ring without
    shedding edges) and solves G mu = -Uinf . n."""

    def __init__(self, grid, element=ConstantSource, **kw):
        if element not in (ConstantSource, ConstantDoublet):
            raise NotImplementedError("NonLiftingBody supports ConstantSource and ConstantDoublet elements")
        super().__init__(grid, element, None, **kw)


class RigidWakeBody(AbstractBody):
    """RigidWakeBody{VortexRing, N}(grid, shedding).

    ``formulation=1`` is the square no-flow-through system; ``formulation=2`` the
    watertight least-squares one (``RigidWakeBody{VortexRing, 2}``,
    src/FLOWPanel_liftingbody.jl:236-316): ``solve`` then prescribes the strengths
    ``elprescribe`` (default ``[(0, 0.0)]``, 0-based) and solves the normal equations.
    """

    def __init__(self, grid, shedding=None, element=VortexRing, formulation=1, **kw):
        if element is not VortexRing:
            raise NotImplementedError("RigidWakeBody supports VortexRing elements")
        if formulation not in (1, 2):
            raise ValueError("formulation must be 1 or 2")
        self.formulation = formulation
        super().__init__(grid, element, shedding, **kw)


class MultiBody:
    """Column-block composition of several bodies (MultiBody(bodies, names))."""

    def __init__(self, bodies, names=None):
        self.bodies = list(bodies)
        self.names = list(names) if names is not None else ["body%d" % i for i in range(len(self.bodies))]
        if len(self.names) != len(self.bodies):
            raise ValueError("need one name per body")
        ctxs = {id(b.ctx) for b in self.bodies}
        if len(ctxs) != 1:
            raise ValueError("all bodies must share one context")
        self.ctx = self.bodies[0].ctx
        self.fields = {}

    @property
    def ncells(self):
        return sum(b.ncells for b in self.bodies)

    @property
    def nsheddings(self):
        return sum(b.nsheddings for b in self.bodies)

    @property
    def strength(self):
        return np.concatenate([b.strength for b in self.bodies])

    def check_solved(self):
        return all(b.check_solved() for b in self.bodies)

    def add_field(self, name, data):
        self.fields[name] = data

    def get_field(self, name):
        return self.fields[name]


def generate_revolution_liftbody(bodytype, points, n_theta, axis_angle=270.0, **bodyoptargs):
    """generate_revolution_liftbody(bodytype, points, NDIVS_theta; bodyoptargs)
    (src/FLOWPanel_abstractliftingbody.jl:124-167): the contour ``points`` (k x 2:
    axial position, radius; closed, first point at the trailing edge) revolved
    about the x axis, split into triangles, every azimuthal strip shedding at the
    trailing edge.  ``bodyoptargs`` go to ``bodytype`` (e.g. ``formulation=2`` for
    the watertight least-squares RigidWakeBody{VortexRing, 2})."""
    from .geometry import revolution_liftbody, triangulate
    grid, sh = revolution_liftbody(points, n_theta, axis_angle)
    grid, sh = triangulate(grid, sh)
    return bodytype(grid, sh, **bodyoptargs)


def _components(body):
    return body.bodies


def calc_normals(body):
    """3 x N normals (src/FLOWPanel_abstractbody.jl:738-776)."""
    return _geometry(body)[0]


def calc_controlpoints(body):
    """3 x N control points (src/FLOWPanel_abstractbody.jl:635-711)."""
    return _geometry(body)[1]


def calc_areas(body):
    """N panel areas (src/FLOWPanel_abstractbody.jl:905-927)."""
    return _geometry(body)[2]


def _geometry(body):
    outs = []
    for b in _components(body):
        n = np.zeros((3, b.ncells), order="F")
        c = np.zeros((3, b.ncells), order="F")
        a = np.zeros(b.ncells)
        b.ctx.check(nat.lib().fp_body_geometry(b.ctx.handle, b.handle, nat.dptr(n), nat.dptr(c), nat.dptr(a)))
        outs.append((n, c, a))
    return (np.hstack([o[0] for o in outs]), np.hstack([o[1] for o in outs]),
            np.concatenate([o[2] for o in outs]))
