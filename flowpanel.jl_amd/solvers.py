"""``solve`` and the solver-function protocol.

``solve(body, Uinfs, Das, Dbs; solver)`` mirrors src/FLOWPanel_liftingbody.jl:172-216,
src/FLOWPanel_nonliftingbody.jl:93-130 and src/FLOWPanel_multibody.jl:312-364:
the AIC is assembled, the no-flow-through RHS formed and the system solved on
the device; the strengths stay on the device and a host copy is stored in the
body's fields (``Gamma``/``sigma``, ``Uinf``, ``Da``, ``Db``).

The solver functions keep the reference's protocol
``solver(y, A, b; optargs...) -> y`` (src/FLOWPanel_linearsolver.jl:24-157) for
caller matrices, and double as selectors of the device path when passed to
``solve``:

* ``solve_ludiv`` / ``solve_backslash`` -- FP64 LU (getrf + getrs)
* ``solve_gmres``                       -- FP64 GMRES (memory 50, atol = rtol = sqrt(eps))
* ``solve_lu32_ir``                     -- FP32 LU on MFMA + FP64 iterative refinement
* ``solve_gmres_ir``                    -- FP32 block-Jacobi GMRES inside FP64 refinement
"""
import ctypes

import numpy as np

from . import _native as nat

last_stats = {}


def _solve_mat(code, y, A, b, **optargs):
    A = np.asarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = A.shape[0]
    if A.ndim != 2 or A.shape[1] != n:
        raise ValueError("A must be square, got %s" % (A.shape,))
    if b.shape[0] != n:
        raise ValueError("b has %d rows, expected %d" % (b.shape[0], n))
    ctx = nat.context()
    Af = nat.f64(A)
    h = ctypes.c_void_p()
    ctx.check(nat.lib().fp_mat_upload(ctx.handle, n, nat.dptr(Af), ctypes.byref(h)))
    try:
        k = 1 if b.ndim == 1 else b.shape[1]
        bf = nat.f64(b.reshape(n, k))
        x = np.zeros((n, k), order="F")
        o = nat.opts(code, **optargs)
        s = nat.SolveStats()
        ctx.check(nat.lib().fp_solve_mat(ctx.handle, h, nat.dptr(bf), k, nat.dptr(x), ctypes.byref(o),
                                         ctypes.byref(s)))
        last_stats.clear()
        last_stats.update(s.as_dict())
    finally:
        nat.lib().fp_mat_destroy(h)
    x = x.reshape(b.shape)
    if y is None:
        return x
    y[...] = x
    return y


def solve_ludiv(y, A, b, **optargs):
    """solve_ludiv!(y, A, b): LU factorisation + substitution (FP64)."""
    return _solve_mat(nat.FP_SOLVER_LU, y, A, b, **optargs)


def solve_backslash(y, A, b, **optargs):
    """solve_backslash!(y, A, b): y = A \\ b (dense, square: LU with partial pivoting)."""
    return _solve_mat(nat.FP_SOLVER_LU, y, A, b, **optargs)


def solve_gmres(y, A, b, atol=None, rtol=None, memory=50, **optargs):
    """solve_gmres!(y, A, b): restarted GMRES in FP64 (defaults of src/FLOWPanel_linearsolver.jl:136-139)."""
    tol = rtol if rtol is not None else (atol if atol is not None else None)
    return _solve_mat(nat.FP_SOLVER_GMRES, y, A, b, tol=tol, restart=memory, **optargs)


def solve_lu32_ir(y, A, b, **optargs):
    """FP32 LU (MFMA) with FP64 iterative refinement of r = b - A x."""
    return _solve_mat(nat.FP_SOLVER_LU32_IR, y, A, b, **optargs)


def solve_gmres_ir(y, A, b, **optargs):
    """Block-Jacobi GMRES on the FP32 matrix inside FP64 iterative refinement."""
    return _solve_mat(nat.FP_SOLVER_GMRES_IR, y, A, b, **optargs)


_CODES = {solve_ludiv: nat.FP_SOLVER_LU, solve_backslash: nat.FP_SOLVER_LU, solve_gmres: nat.FP_SOLVER_GMRES,
          solve_lu32_ir: nat.FP_SOLVER_LU32_IR, solve_gmres_ir: nat.FP_SOLVER_GMRES_IR}


def set_freestream(body, Uinfs, Das=None, Dbs=None):
    """Upload Uinfs (3 x N) and Das/Dbs (3 x ns) of every component."""
    Uinfs = np.asarray(Uinfs, dtype=np.float64)
    if Uinfs.shape != (3, body.ncells):
        raise nat.FPError(nat.FP_ERR_SHAPE, "Invalid Uinfs; expected size (3, %d), got %s"
                          % (body.ncells, Uinfs.shape))
    ns = body.nsheddings
    if ns:
        for nm, D in (("Das", Das), ("Dbs", Dbs)):
            if D is None or np.asarray(D).shape != (3, ns):
                raise nat.FPError(nat.FP_ERR_SHAPE, "Invalid %s; expected size (3, %d), got %s"
                                  % (nm, ns, None if D is None else np.asarray(D).shape))
        Das = np.asarray(Das, dtype=np.float64)
        Dbs = np.asarray(Dbs, dtype=np.float64)
    col = sh = 0
    for b in body.bodies:
        U = nat.f64(Uinfs[:, col:col + b.ncells])
        da = nat.f64(Das[:, sh:sh + b.nsheddings]) if b.nsheddings else None
        db = nat.f64(Dbs[:, sh:sh + b.nsheddings]) if b.nsheddings else None
        b.ctx.check(nat.lib().fp_body_set_freestream(b.ctx.handle, b.handle, nat.dptr(U), nat.dptr(da),
                                                     nat.dptr(db)))
        b.add_field("Uinf", U)
        if b.nsheddings:
            b.add_field("Da", da)
            b.add_field("Db", db)
        col += b.ncells
        sh += b.nsheddings


def calc_elprescribe(body, indices=(0,), values=(0.0,)):
    """calc_elprescribe (src/FLOWPanel_multibody.jl:565-599): for every watertight
    (formulation 2) component, prescribe the strengths of its local ``indices``
    to ``values``; returns global 0-based (index, value) pairs."""
    out, off = [], 0
    for b in body.bodies:
        if getattr(b, "formulation", 1) == 2:
            out += [(off + int(i), float(v)) for i, v in zip(indices, values)]
        off += b.ncells
    return out


def solve(body, Uinfs, Das=None, Dbs=None, solver=solve_ludiv, solver_optargs=None, elprescribe=None, **unused):
    """solve(body, Uinfs[, Das, Dbs]; solver, solver_optargs) on the device.

    Watertight bodies (``RigidWakeBody(..., formulation=2)``) are solved by the
    least-squares formulation with ``elprescribe`` (default ``calc_elprescribe``).
    Returns the solver statistics (refinement steps, GMRES iterations, FP64
    relative residual, device milliseconds of assembly/factorisation/solve).
    """
    if solver not in _CODES:
        raise ValueError("solver must be one of %s" % [f.__name__ for f in _CODES])
    set_freestream(body, Uinfs, Das, Dbs)
    kw = dict(solver_optargs or {})
    if solver is solve_gmres:
        if "rtol" in kw or "atol" in kw:
            kw["tol"] = kw.pop("rtol", kw.pop("atol", None))
        if "memory" in kw:
            kw["restart"] = kw.pop("memory")
    o = nat.opts(_CODES[solver], **kw)
    s = nat.SolveStats()
    ctx = body.ctx
    comps = body.bodies
    lsq = any(getattr(b, "formulation", 1) == 2 for b in comps)
    if lsq or elprescribe is not None:
        pres = calc_elprescribe(body) if elprescribe is None else [(int(i), float(v)) for i, v in elprescribe]
        pres.sort()
        idx = np.array([p[0] for p in pres], dtype=np.int64)
        val = np.array([p[1] for p in pres], dtype=np.float64)
        ctx.check(nat.lib().fp_solve_bodies_lsq(ctx.handle, nat.handles(comps), len(comps), nat.iptr(idx),
                                                nat.dptr(val), len(pres), 0, ctypes.byref(o), ctypes.byref(s)))
    else:
        ctx.check(nat.lib().fp_solve_bodies(ctx.handle, nat.handles(comps), len(comps), ctypes.byref(o),
                                            ctypes.byref(s)))
    for b in comps:
        g = np.zeros(b.ncells)
        ctx.check(nat.lib().fp_body_get_strength(ctx.handle, b.handle, nat.dptr(g)))
        b.strength = g
        b._solved = True
        b.add_field(b.element.field, g)
    stats = s.as_dict()
    last_stats.clear()
    last_stats.update(stats)
    return stats


def solve_freestreams(body, Uinfs, Das=None, Dbs=None, solver=solve_ludiv, solver_optargs=None):
    """Strengths for k freestreams under the fixed-wake convention.

    ``Uinfs`` is 3 x N x k (or a list of k arrays 3 x N); the wake directions
    ``Das``/``Dbs`` (3 x ns) are shared by every case, so the AIC is assembled
    and factorised (or preconditioned) once and the k freestreams are solved as
    k right-hand sides.  Returns (strengths N x k, stats).  The body keeps the
    wake directions; its stored strengths are left unchanged.
    """
    if solver not in _CODES:
        raise ValueError("solver must be one of %s" % [f.__name__ for f in _CODES])
    U = np.asarray(Uinfs, dtype=np.float64)
    if U.ndim == 3 and U.shape[0] != 3 and U.shape[1] == 3:
        U = np.transpose(U, (1, 2, 0))  # k x 3 x N -> 3 x N x k
    if U.ndim != 3 or U.shape[:2] != (3, body.ncells):
        raise nat.FPError(nat.FP_ERR_SHAPE, "Invalid Uinfs; expected size (3, %d, k), got %s"
                          % (body.ncells, U.shape))
    k = U.shape[2]
    set_freestream(body, U[:, :, 0], Das, Dbs)
    Uk = np.ascontiguousarray(np.transpose(U, (2, 1, 0)))  # k x N x 3: column-major 3 x N x k
    G = np.zeros((body.ncells, k), order="F")
    o = nat.opts(_CODES[solver], **(solver_optargs or {}))
    s = nat.SolveStats()
    ctx = body.ctx
    comps = body.bodies
    ctx.check(nat.lib().fp_solve_bodies_multi(ctx.handle, nat.handles(comps), len(comps), nat.dptr(Uk), k,
                                              ctypes.byref(o), nat.dptr(G), ctypes.byref(s)))
    stats = s.as_dict()
    last_stats.clear()
    last_stats.update(stats)
    return G, stats


class BodyBatch:
    """Independent bodies (each its own system) solved together.

    Construction sizes the device buffers, autotunes the batched assembly and
    captures one hipGraph (persistent work-queue assembly of every body, then
    per body its wake columns, LU and refinement sweeps on parallel streams).
    ``solve()`` replays it; set each body's freestream first (``set_freestream``
    or ``solve``).  Solvers: ``solve_lu32_ir`` (FP32 AIC + LU, FP64 residuals;
    ``residual_mode=0`` recomputes them matrix-free) or ``solve_ludiv``.
    """

    def __init__(self, bodies, solver=solve_lu32_ir, solver_optargs=None):
        if solver not in (solve_lu32_ir, solve_ludiv, solve_backslash):
            raise ValueError("a body batch runs solve_lu32_ir or solve_ludiv")
        self.bodies = list(bodies)
        self.ctx = self.bodies[0].ctx
        o = nat.opts(_CODES[solver], **(solver_optargs or {}))
        h = ctypes.c_void_p()
        self.ctx.check(nat.lib().fp_batch_create(self.ctx.handle, nat.handles(self.bodies), len(self.bodies),
                                                 ctypes.byref(o), ctypes.byref(h)))
        self.handle = h

    def solve(self):
        s = nat.BatchStats()
        self.ctx.check(nat.lib().fp_batch_solve(self.ctx.handle, self.handle, ctypes.byref(s)))
        for b in self.bodies:
            g = np.zeros(b.ncells)
            self.ctx.check(nat.lib().fp_body_get_strength(self.ctx.handle, b.handle, nat.dptr(g)))
            b.strength = g
            b._solved = True
            b.add_field(b.element.field, g)
        return s.as_dict()

    def autotune_table(self):
        """[(targets per tile, workgroups, ms)] of the assembly candidates timed at construction."""
        out = []
        for k in range(9):
            tm, wg, ms = ctypes.c_int(), ctypes.c_int(), np.zeros(1)
            if nat.lib().fp_batch_autotune_get(self.handle, k, ctypes.byref(tm), ctypes.byref(wg), nat.dptr(ms)):
                break
            out.append((tm.value, wg.value, float(ms[0])))
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None) and nat._LIB is not None:
                nat._LIB.fp_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def assemble(body, precision="f64"):
    """Device AIC of the body (column-major numpy copy), _G_U! of the reference."""
    ctx = body.ctx
    comps = body.bodies
    h = ctypes.c_void_p()
    prec = nat.FP_F64 if precision == "f64" else nat.FP_F32
    ctx.check(nat.lib().fp_assemble(ctx.handle, nat.handles(comps), len(comps), prec, ctypes.byref(h)))
    try:
        n = nat.lib().fp_mat_size(h)
        G = np.zeros((n, n), order="F")
        ctx.check(nat.lib().fp_mat_download(ctx.handle, h, nat.dptr(G)))
    finally:
        nat.lib().fp_mat_destroy(h)
    return G


def calc_bc_noflowthrough(Us, normals):
    """RHS_i = -U_i . n_i (src/FLOWPanel_linearsolver.jl:321-328)."""
    Us = np.asarray(Us, dtype=np.float64)
    normals = np.asarray(normals, dtype=np.float64)
    if Us.shape != normals.shape:
        raise ValueError("Us and normals must have the same size")
    return -np.einsum("ij,ij->j", Us, normals)


def block_inverses(A, block):
    """The FP32 block-Jacobi inverses the Krylov solvers apply to A (blocks of
    ``block`` rows along the index order, the last one padded with the identity):
    array (nblocks, b, b), b = min(block, n)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    b = min(int(block), n)
    nb = (n + b - 1) // b
    ctx = nat.context()
    Af = nat.f64(A)
    h = ctypes.c_void_p()
    ctx.check(nat.lib().fp_mat_upload(ctx.handle, n, nat.dptr(Af), ctypes.byref(h)))
    try:
        out = np.zeros((nb, b, b), dtype=np.float32)
        ctx.check(nat.lib().fp_mat_block_inverses(ctx.handle, h, int(block),
                                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    finally:
        nat.lib().fp_mat_destroy(h)
    return out


def getrf(A, precision="f64"):
    """Device LU factorisation of a caller matrix: (LU, ipiv) in LAPACK form (1-based ipiv)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    ctx = nat.context()
    Af = nat.f64(A)
    h = ctypes.c_void_p()
    ctx.check(nat.lib().fp_mat_upload(ctx.handle, n, nat.dptr(Af), ctypes.byref(h)))
    try:
        LU = np.zeros((n, n), order="F")
        ipiv = np.zeros(n, dtype=np.int64)
        ctx.check(nat.lib().fp_mat_getrf(ctx.handle, h, nat.FP_F64 if precision == "f64" else nat.FP_F32,
                                         nat.dptr(LU), nat.iptr(ipiv)))
    finally:
        nat.lib().fp_mat_destroy(h)
    return LU, ipiv
