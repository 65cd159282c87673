"""fp_uind_noalloc (Uind! with no device allocation) against fp_uind, through the C ABI.

The caller owns the device arrays (allocated here with hipMalloc through
ctypes); the library takes its staging and partial sums from its per-GPU
scratch array.  Two calls in a row share that array, the second with fewer
targets (and so a different split of the panels); both must equal fp_uind
on the same targets, which runs the same kernels: bit-exact.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceArray:
    """A float64 device buffer owned by the test (hipMalloc/hipFree through ctypes)."""

    def __init__(self, hip, host):
        self.hip = hip
        self.host = np.asfortranarray(host, dtype=np.float64)
        self.nbytes = self.host.nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes),
                             H2D) == 0

    def download(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes),
                                  D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _noalloc(fp, body, d_targets, nt, d_out):
    from flowpanel_amd import _native as nat
    ctx = body.ctx
    ctx.check(nat.lib().fp_uind_noalloc(ctx.handle, nat.handles([body]), 1, d_targets, nt, d_out))
    ctx.synchronize()


def test_uind_noalloc_matches_uind(fp, geo, hip):
    g, sh = geo.loft_wing(n_chord=24, n_span=16, span=6.0, sweep_deg=20.0)
    body = fp.RigidWakeBody(g, sh)
    V = fp.direction(4.0)
    U = np.tile(V[:, None], (1, body.ncells))
    D = np.tile(V[:, None], (1, body.nsheddings))
    fp.solve(body, U, D, D)
    rng = np.random.default_rng(21)
    for nt in (5000, 700):      # second call smaller: other split count, same scratch array
        T = np.asfortranarray(rng.uniform([-1, -4, -1], [3, 4, 1], size=(nt, 3)).T)
        pre = rng.standard_normal((3, nt))            # += into what the caller holds
        ref = np.asfortranarray(pre.copy())
        fp.Uind(body, T, ref)
        dt, do = DeviceArray(hip, T), DeviceArray(hip, pre)
        try:
            _noalloc(fp, body, dt.ptr, nt, do.ptr)
            out = do.download()
        finally:
            dt.free()
            do.free()
        assert np.array_equal(out, ref), (nt, np.abs(out - ref).max())


def test_uind_noalloc_rejects_what_it_cannot_take(fp, geo, hip):
    body = fp.NonLiftingBody(geo.sphere(n_meridional=6, n_azimuthal=8))
    fp.solve(body, np.tile([[1.0], [0.0], [0.0]], (1, body.ncells)))
    from flowpanel_amd import _native as nat
    ctx = body.ctx
    T = np.zeros((3, 4), order="F")
    dt, do = DeviceArray(hip, T), DeviceArray(hip, T)
    try:
        lim = 1 << 20
        st = nat.lib().fp_uind_noalloc(ctx.handle, nat.handles([body]), 1, dt.ptr, lim + 1, do.ptr)
        assert st == nat.FP_ERR_ARG
        assert b"targets" in nat.lib().fp_ctx_last_error(ctx.handle)
        # host memory instead of device memory is refused before any launch
        st = nat.lib().fp_uind_noalloc(ctx.handle, nat.handles([body]), 1, T.ctypes.data_as(ctypes.c_void_p), 4,
                                       do.ptr)
        assert st == nat.FP_ERR_ARG
    finally:
        dt.free()
        do.free()
