"""ctypes binding of libflowpanel_hip.so (the C ABI of include/flowpanel_hip.h).

The library is the only compute path: if it is missing or no GPU is visible,
every call raises; nothing falls back to a CPU implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libflowpanel_hip.so")

FP_OK = 0
FP_ERR_ARG = 1
FP_ERR_SHAPE = 2
FP_ERR_NONUNIT_DIRECTION = 3
FP_ERR_BAD_DIRECTIONS = 4
FP_ERR_UNSOLVED = 5
FP_ERR_SINGULAR = 6
FP_ERR_NOT_CONVERGED = 7
FP_ERR_DEVICE = 8
FP_ERR_NOMEM = 9
FP_ERR_UNSUPPORTED = 10

FP_ELEM_SOURCE = 0
FP_ELEM_VORTEXRING = 1
FP_F64 = 0
FP_F32 = 1
FP_SOLVER_LU = 0
FP_SOLVER_LU32_IR = 1
FP_SOLVER_GMRES_IR = 2
FP_SOLVER_GMRES = 3
FP_RESID_MATRIX_FREE = 0
FP_RESID_STORED = 1
FP_ORDER_INDEX = 0
FP_ORDER_SPATIAL = 1
FP_KUTTA_NONE = 0
FP_KUTTA_REFERENCE = 1
FP_KUTTA_PAIR = 2
FP_SHARPTE_OFF = 0
FP_SHARPTE_PAIRS = 1
FP_SHARPTE_CELLS = 2
CHARLEN = {"unitary": 0, "bbox": 1, "maxdist": 2, "sqrtarea": 3}

EXPORTS = [
    "fp_solve_opts_default", "fp_ctx_create", "fp_ctx_destroy", "fp_ctx_last_error", "fp_ctx_synchronize", "fp_ctx_trim",
    "fp_body_create", "fp_body_destroy", "fp_body_ncells", "fp_body_nsheddings", "fp_body_geometry",
    "fp_body_set_freestream", "fp_body_set_strength", "fp_body_get_strength", "fp_solve_bodies",
    "fp_assemble", "fp_mat_upload", "fp_mat_size", "fp_mat_block_inverses", "fp_mat_download", "fp_mat_destroy",
    "fp_solve_mat",
    "fp_mat_getrf", "fp_uind", "fp_phi", "fp_calcfield_U", "fp_calcfield_cp", "fp_ctx_profile",
    "fp_ctx_profile_only", "fp_ctx_profile_get", "fp_ctx_copy_calibrate", "fp_body_nnodes", "fp_body_set_nodes",
    "fp_liftdrag_opts_default", "fp_lift_drag", "fp_lift_drag_adjoint", "fp_solve_bodies_multi",
    "fp_batch_create", "fp_batch_solve", "fp_batch_destroy", "fp_batch_autotune_get", "fp_solve_bodies_lsq",
    "fp_calcfield_F", "fp_calcfield_sectionalforce", "fp_calcfield_ugradmu", "fp_calcfield_ugradmu_node", "fp_calcfield_ugradmu_combined", "fp_uind_noalloc",
    "fp_assembly_blocks", "fp_assembly_pairs", "fp_field_trails",
]


class SolveOpts(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
                ("restart", ctypes.c_int), ("max_iters", ctypes.c_int), ("max_refine", ctypes.c_int),
                ("precond_block", ctypes.c_int), ("coarse_block", ctypes.c_int), ("residual_mode", ctypes.c_int),
                ("precond_order", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("refinements", ctypes.c_int), ("iterations", ctypes.c_int), ("residual", ctypes.c_double),
                ("ms_assemble", ctypes.c_double), ("ms_factor", ctypes.c_double), ("ms_solve", ctypes.c_double),
                ("info", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LiftDragOpts(ctypes.Structure):
    _fields_ = [("probe_offset", ctypes.c_double), ("probe_length", ctypes.c_double),
                ("probe_charlen", ctypes.c_int), ("kutta_mode", ctypes.c_int),
                ("span_dir", ctypes.c_double * 3), ("ref_area", ctypes.c_double), ("solve", SolveOpts)]


class LiftDragStats(ctypes.Structure):
    _fields_ = [("CL", ctypes.c_double), ("CDi", ctypes.c_double), ("ms_forward", ctypes.c_double),
                ("ms_adjoint", ctypes.c_double), ("ms_transpose_solve", ctypes.c_double),
                ("iterations_adjoint", ctypes.c_int), ("forward", SolveStats)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "forward"}
        d["forward"] = self.forward.as_dict()
        return d


class BatchStats(ctypes.Structure):
    _fields_ = [("ms_graph", ctypes.c_double), ("ms_capture", ctypes.c_double), ("ms_extra", ctypes.c_double),
                ("residual", ctypes.c_double), ("panels", ctypes.c_int64), ("nbodies", ctypes.c_int),
                ("sweeps_in_graph", ctypes.c_int), ("extra_sweeps", ctypes.c_int), ("graph_nodes", ctypes.c_int),
                ("recaptured", ctypes.c_int), ("tile_targets", ctypes.c_int), ("queue_workgroups", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FPError(RuntimeError):
    """Error raised by the HIP library; ``code`` is the FP_ERR_* status."""

    def __init__(self, code, msg):
        super().__init__("[flowpanel_hip %d] %s" % (code, msg))
        self.code = code


_LIB = None
_CTX = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError("libflowpanel_hip.so is not built (run flowpanel.jl_amd/build.py); "
                           "there is no CPU fallback")
    L = ctypes.CDLL(LIBPATH)
    i64, dbl, cint = ctypes.c_int64, ctypes.c_double, ctypes.c_int
    pp = ctypes.POINTER(_vp)
    L.fp_solve_opts_default.argtypes = [ctypes.POINTER(SolveOpts), cint]
    L.fp_solve_opts_default.restype = None
    L.fp_ctx_create.argtypes = [cint, pp]
    L.fp_ctx_destroy.argtypes = [_vp]
    L.fp_ctx_destroy.restype = None
    L.fp_ctx_last_error.argtypes = [_vp]
    L.fp_ctx_last_error.restype = ctypes.c_char_p
    L.fp_ctx_synchronize.argtypes = [_vp]
    L.fp_ctx_trim.argtypes = [_vp]
    L.fp_body_create.argtypes = [_vp, _d, i64, _i64, _i64, i64, cint, cint, _i64, i64, dbl, cint, dbl, dbl, pp]
    L.fp_body_destroy.argtypes = [_vp]
    L.fp_body_destroy.restype = None
    L.fp_body_ncells.argtypes = [_vp]
    L.fp_body_ncells.restype = i64
    L.fp_body_nsheddings.argtypes = [_vp]
    L.fp_body_nsheddings.restype = i64
    L.fp_body_geometry.argtypes = [_vp, _vp, _d, _d, _d]
    L.fp_body_set_freestream.argtypes = [_vp, _vp, _d, _d, _d]
    L.fp_body_set_strength.argtypes = [_vp, _vp, _d]
    L.fp_body_get_strength.argtypes = [_vp, _vp, _d]
    L.fp_solve_bodies.argtypes = [_vp, pp, cint, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)]
    L.fp_assemble.argtypes = [_vp, pp, cint, cint, pp]
    L.fp_mat_upload.argtypes = [_vp, i64, _d, pp]
    L.fp_mat_size.argtypes = [_vp]
    L.fp_mat_size.restype = i64
    L.fp_mat_block_inverses.argtypes = [_vp, _vp, cint, ctypes.POINTER(ctypes.c_float)]
    L.fp_mat_block_inverses.restype = cint
    L.fp_mat_download.argtypes = [_vp, _vp, _d]
    L.fp_mat_destroy.argtypes = [_vp]
    L.fp_mat_destroy.restype = None
    L.fp_solve_mat.argtypes = [_vp, _vp, _d, i64, _d, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)]
    L.fp_mat_getrf.argtypes = [_vp, _vp, cint, _d, _i64]
    L.fp_uind.argtypes = [_vp, pp, cint, _d, i64, _d]
    L.fp_phi.argtypes = [_vp, pp, cint, _d, i64, _d]
    L.fp_calcfield_U.argtypes = [_vp, pp, cint, _d]
    L.fp_calcfield_cp.argtypes = [_vp, _d, i64, dbl, _i64, i64, cint, cint, _d]
    L.fp_ctx_profile.argtypes = [_vp, cint]
    L.fp_ctx_profile_only.argtypes = [_vp, ctypes.c_char_p]
    L.fp_ctx_profile_get.argtypes = [_vp, ctypes.c_char_p, _i64, _d]
    L.fp_ctx_copy_calibrate.argtypes = [_vp, i64]
    L.fp_solve_bodies_multi.argtypes = [_vp, pp, cint, _d, i64, ctypes.POINTER(SolveOpts), _d,
                                        ctypes.POINTER(SolveStats)]
    L.fp_batch_create.argtypes = [_vp, pp, cint, ctypes.POINTER(SolveOpts), pp]
    L.fp_batch_solve.argtypes = [_vp, _vp, ctypes.POINTER(BatchStats)]
    L.fp_batch_destroy.argtypes = [_vp]
    L.fp_batch_destroy.restype = None
    L.fp_batch_autotune_get.argtypes = [_vp, cint, ctypes.POINTER(cint), ctypes.POINTER(cint), _d]
    L.fp_solve_bodies_lsq.argtypes = [_vp, pp, cint, _i64, _d, i64, cint, ctypes.POINTER(SolveOpts),
                                      ctypes.POINTER(SolveStats)]
    L.fp_calcfield_ugradmu.argtypes = [_vp, pp, cint, cint, _d]
    L.fp_calcfield_F.argtypes = [_vp, pp, cint, _d, dbl, dbl, cint, _d, _d, _d, _d, _d]
    L.fp_calcfield_sectionalforce.argtypes = [_vp, _vp, _d, _d, i64, i64, _d, _d, _d]
    L.fp_calcfield_ugradmu_node.argtypes = [_vp, pp, cint, cint, _d, _d]
    L.fp_calcfield_ugradmu_combined.argtypes = [_vp, pp, cint, cint, ctypes.c_double, cint, _d]
    L.fp_uind_noalloc.argtypes = [_vp, pp, cint, _vp, i64, _vp]
    L.fp_assembly_blocks.argtypes = [_d, i64, _i64, _i64, i64, cint, ctypes.POINTER(ctypes.c_int32), i64, _i64]
    L.fp_assembly_blocks.restype = i64
    L.fp_assembly_pairs.argtypes = [_d, i64, _i64, _i64, i64, cint, ctypes.POINTER(ctypes.c_int32), i64, _i64]
    L.fp_assembly_pairs.restype = i64
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.fp_field_trails.argtypes = [_i64, _i64, i64, i64, cint, i32p, i32p, i32p, i64, i64, _i64]
    L.fp_field_trails.restype = i64
    L.fp_body_nnodes.argtypes = [_vp]
    L.fp_body_nnodes.restype = i64
    L.fp_body_set_nodes.argtypes = [_vp, _vp, _d]
    L.fp_liftdrag_opts_default.argtypes = [ctypes.POINTER(LiftDragOpts), cint]
    L.fp_liftdrag_opts_default.restype = None
    L.fp_lift_drag.argtypes = [_vp, pp, cint, _d, ctypes.POINTER(LiftDragOpts), _d, _d,
                               ctypes.POINTER(LiftDragStats)]
    L.fp_lift_drag_adjoint.argtypes = [_vp, pp, cint, _d, ctypes.POINTER(LiftDragOpts), _d, _d, _d, _d,
                                       ctypes.POINTER(LiftDragStats)]
    _LIB = L
    return L


class Context:
    """One HIP device + stream (fp_ctx)."""

    def __init__(self, device=0):
        h = _vp()
        st = lib().fp_ctx_create(device, ctypes.byref(h))
        if st != FP_OK:
            raise FPError(st, "fp_ctx_create(%d) failed: no usable GPU" % device)
        self.handle = h

    def check(self, st):
        if st != FP_OK:
            msg = lib().fp_ctx_last_error(self.handle)
            raise FPError(st, msg.decode() if msg else "")

    def synchronize(self):
        self.check(lib().fp_ctx_synchronize(self.handle))

    def trim(self):
        """Free the large device buffers the context keeps for re-use by the next call of the same size."""
        self.check(lib().fp_ctx_trim(self.handle))

    def profile(self, enable=True, only=None):
        """Start (clearing) or stop per-kernel device timing; `only` restricts it to one kernel class."""
        if enable and only:
            self.check(lib().fp_ctx_profile_only(self.handle, only.encode()))
        else:
            self.check(lib().fp_ctx_profile(self.handle, int(bool(enable))))

    def profile_get(self, kernel):
        """(launches, total device ms) recorded for one kernel class."""
        n = np.zeros(1, dtype=np.int64)
        ms = np.zeros(1)
        self.check(lib().fp_ctx_profile_get(self.handle, kernel.encode(), iptr(n), dptr(ms)))
        return int(n[0]), float(ms[0])

    def __del__(self):
        try:
            if self.handle and _LIB is not None:
                _LIB.fp_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def context():
    """Process-wide default context on device 0."""
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


def dptr(a):
    return a.ctypes.data_as(_d) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_i64) if a is not None else None


def f64(a):
    """Column-major contiguous float64 (a 3 x n array becomes x,y,z per point)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def handles(objs):
    arr = (_vp * len(objs))(*[o.handle for o in objs])
    return arr


def opts(solver, **kw):
    o = SolveOpts()
    lib().fp_solve_opts_default(ctypes.byref(o), solver)
    for k, v in kw.items():
        if v is None:
            continue
        if not hasattr(o, k):
            raise TypeError("unknown solver option %r" % k)
        if k == "precond_order" and isinstance(v, str):
            v = {"index": FP_ORDER_INDEX, "spatial": FP_ORDER_SPATIAL}[v]
        setattr(o, k, v)
    return o
