"""CPU checks of the reverse-mode (adjoint) element and geometry functions.

The library's reverse functions (flowpanel.jl_amd/csrc/fp_adjoint.h) and its
geometry function (fp_geometry.h) are compiled for the host from the very
headers the HIP kernels include (tests/native/adjoint_host.cpp) and checked
against central finite differences of their own forward; the forward itself is
tied to the oracle's element kernels (ring, horseshoe) and geometry, so the
reverse is the derivative of the reference's functions.  A dropped adjoint
term, a wrong sign or a swapped vertex fails the FD comparison.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "adjoint_host.cpp")
PD = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def ah(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("adjhost") / "adjoint_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out, SRC], check=True)
    lib = ctypes.CDLL(out)
    D, I = ctypes.c_double, ctypes.c_int
    sig = {"ah_seg": [PD, PD, PD, D, D, PD], "ah_seg_rev": [PD, PD, PD, D, D, PD, PD, PD, PD],
           "ah_horse": [PD, PD, PD, PD, D, PD, D, D, PD], "ah_horse_rev": [PD, PD, PD, PD, D, PD, D, D, PD, PD],
           "ah_ring4": [PD, PD, D, D, PD], "ah_ring4_rev": [PD, PD, D, D, PD, PD, PD, PD],
           "ah_geom4": [PD, I, I, PD], "ah_geom4_dir": [PD, PD, I, I, PD]}
    for name, args in sig.items():
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = args
    return lib


def p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(PD)


def fd_grad(f, x, h=1e-6):
    """Central differences of the scalar f at x (flattened)."""
    x = np.array(x, dtype=np.float64)
    g = np.zeros(x.size)
    for k in range(x.size):
        xp, xm = x.copy().ravel(), x.copy().ravel()
        xp[k] += h
        xm[k] -= h
        g[k] = (f(xp.reshape(x.shape)) - f(xm.reshape(x.shape))) / (2 * h)
    return g.reshape(x.shape)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


OFF, CUT = 1e-8, 1e-14


def test_segment_reverse_matches_fd(ah):
    rng = np.random.default_rng(11)
    for _ in range(5):
        a, b, X, ub = rng.normal(size=(4, 3))

        def J(a_, b_, X_):
            u = np.zeros(3)
            ah.ah_seg(p(a_), p(b_), p(X_), OFF, CUT, p(u))
            return u.dot(ub)

        A, B, Xb = np.zeros(3), np.zeros(3), np.zeros(3)
        ah.ah_seg_rev(p(a), p(b), p(X), OFF, CUT, p(ub), p(A), p(B), p(Xb))
        assert rel(A, fd_grad(lambda v: J(v, b, X), a)) < 1e-7
        assert rel(B, fd_grad(lambda v: J(a, v, X), b)) < 1e-7
        assert rel(Xb, fd_grad(lambda v: J(a, b, v), X)) < 1e-7
        # translation invariance: the three adjoints sum to zero
        assert np.abs(A + B + Xb).max() < 1e-12 * np.abs(Xb).max()


def test_horseshoe_reverse_matches_fd(ah, orc):
    rng = np.random.default_rng(12)
    for _ in range(4):
        pa, pb, X, ub = rng.normal(size=(4, 3))
        da, db = rng.normal(size=(2, 3))
        da /= np.linalg.norm(da)
        db /= np.linalg.norm(db)
        s = 0.7
        # the host forward is the oracle's horseshoe
        u = np.zeros(3)
        ah.ah_horse(p(pa), p(pb), p(da), p(db), s, p(X), OFF, CUT, p(u))
        uo = np.zeros((3, 1), order="F")
        orc.U_semiinfinite_horseshoe(pa, pb, da, db, s, X.reshape(3, 1), uo)
        assert np.abs(u - uo[:, 0]).max() < 1e-13 * np.abs(u).max()

        def J(v):
            uu = np.zeros(3)
            ah.ah_horse(p(v[0]), p(v[1]), p(v[2]), p(v[3]), s, p(v[4]), OFF, CUT, p(uu))
            return uu.dot(ub)

        x0 = np.stack([pa, pb, da, db, X])
        bars = np.zeros(15)
        ah.ah_horse_rev(p(pa), p(pb), p(da), p(db), s, p(X), OFF, CUT, p(ub), p(bars))
        assert rel(bars.reshape(5, 3), fd_grad(J, x0)) < 1e-7


def test_ring_reverse_matches_fd(ah, orc):
    rng = np.random.default_rng(13)
    for nv in (4, 3):
        for _ in range(3):
            P = rng.normal(size=(4, 3)) * 0.5 + np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
            if nv == 3:
                P[3] = P[0]  # a triangle repeats its first vertex (zero closing edge)
            X = rng.normal(size=3) + [0.5, 0.5, 0.3]
            w = rng.normal(size=(2, 3))
            # forward = the oracle's vortex ring
            u = np.zeros(3)
            ah.ah_ring4(p(P), p(X), OFF, CUT, p(u))
            uo = np.zeros((3, 1), order="F")
            orc.U_vortexring(np.asfortranarray(P[:nv].T), np.arange(nv), 1.0, X.reshape(3, 1), uo)
            assert np.abs(u - uo[:, 0]).max() < 1e-13 * np.abs(u).max()
            g, Xb, pb = np.zeros(2), np.zeros(6), np.zeros(24)
            ah.ah_ring4_rev(p(P), p(X), OFF, CUT, p(w), p(g), p(Xb), p(pb))
            Xb, pb = Xb.reshape(2, 3), pb.reshape(2, 4, 3)
            for m in range(2):
                assert abs(g[m] - u.dot(w[m])) < 1e-13 * np.abs(u).max() * np.abs(w[m]).max()

                def Jx(v):
                    uu = np.zeros(3)
                    ah.ah_ring4(p(P), p(v), OFF, CUT, p(uu))
                    return uu.dot(w[m])

                def Jp(v):
                    v = v.copy()
                    if nv == 3:
                        v[3] = v[0]  # the padded slot is the first vertex: they move together
                    uu = np.zeros(3)
                    ah.ah_ring4(p(v), p(X), OFF, CUT, p(uu))
                    return uu.dot(w[m])

                assert rel(Xb[m], fd_grad(Jx, X)) < 1e-7
                fd = fd_grad(Jp, P)
                got = pb[m].copy()
                if nv == 3:
                    got[0] += got[3]
                    got[3] = 0
                    fd[3] = 0
                assert rel(got, fd) < 1e-7


@pytest.mark.parametrize("charlen", [0, 1, 2, 3])
@pytest.mark.parametrize("nv", [3, 4])
def test_geometry_directional_derivative(ah, orc, charlen, nv):
    rng = np.random.default_rng(14 + charlen + 10 * nv)
    P = np.array([[0, 0, 0], [1, 0.1, 0], [1.1, 0.9, 0.2], [0.1, 1, 0.1]], float) + rng.normal(size=(4, 3)) * 0.05
    if nv == 3:
        P[3] = P[0]
    out = np.zeros(8)
    ah.ah_geom4(p(P), nv, charlen, p(out))
    # forward = the oracle's geometry (normal, control point without offset, area)
    nodes = np.asfortranarray(P[:nv].T)
    n, cp, area = orc.geometry(nodes, np.array([0, nv]), np.arange(nv), cpoffset=0.0, charlen=charlen)
    assert np.abs(out[:3] - n[:, 0]).max() < 1e-14
    assert np.abs(out[3:6] - cp[:, 0]).max() < 1e-14
    assert abs(out[6] - area[0]) < 1e-14
    for _ in range(3):
        dp = rng.normal(size=(4, 3))
        if nv == 3:
            dp[3] = dp[0]
        d = np.zeros(8)
        ah.ah_geom4_dir(p(P), p(dp), nv, charlen, p(d))
        h = 1e-6
        op, om = np.zeros(8), np.zeros(8)
        ah.ah_geom4(p(P + h * dp), nv, charlen, p(op))
        ah.ah_geom4(p(P - h * dp), nv, charlen, p(om))
        fd = (op - om) / (2 * h)
        assert np.abs(d - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())
