"""Pins of the oracle's element kernels (CPU only).

Each check follows one of the reference's own element tests
(test/runtests_elements.jl, test/runtests_semiinfiniteelements.jl) or a closed
form the mathematics fixes, chosen so that a dropped term, a wrong sign or a
swapped index in the oracle fails it:

* U = grad(phi) for source and doublet panels (runtests_elements.jl:38-92), with
  a 4th-order central-difference Jacobian instead of ForwardDiff,
* div U = 0 (runtests_elements.jl:156-215),
* dphi/dz(+-0) = +-sigma/2, phi(inf) = 0 (runtests_elements.jl:254-282),
* phi_doublet(+-0) = -+mu/2, phi(inf) = 0 (runtests_elements.jl:284-312),
* the semi-infinite doublet/horseshoe counterparts (runtests_semiinfiniteelements.jl),
* brute-force quadrature of 1/r and n.(x-y)/|x-y|^3 over a panel,
* Biot-Savart line quadrature of a vortex ring, point-source far field, the
  long-finite-segment limit of the semi-infinite vortex, and the frame.
"""
import numpy as np
import pytest


def rot(a, b, c):
    """Rotation by angles (deg) about x, then y, then z (any fixed rotation serves)."""
    a, b, c = np.deg2rad([a, b, c])
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


@pytest.fixture(scope="module")
def pentagon():
    # the 5-vertex planar panel of test/runtests_elements.jl:41-55, rotated and translated
    P = np.array([[0, 0, 0], [0, 1, 0], [0.4, 1.2, 0], [1, 0.8, 0], [0.8, 0, 0]], float)
    R = rot(30, 10, 40)
    nodes = (R.T @ P.T) + 1.0
    return np.asfortranarray(nodes), np.arange(5)


def jac(f, X, h=1e-3):
    """4th-order central-difference Jacobian of f: R^3 -> R^m at X."""
    cols = []
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        cols.append((-f(X + 2 * e) + 8 * f(X + e) - 8 * f(X - e) + f(X - 2 * e)) / (12 * h))
    return np.stack(cols, axis=1)


def test_panel_frame_orthonormal(orc, pentagon):
    nodes, panel = pentagon
    t, o, n = orc.panel_frame(nodes, panel)
    M = np.stack([t, o, n])
    assert np.abs(M @ M.T - np.eye(3)).max() < 10 * np.finfo(float).eps
    assert np.abs(np.cross(t, o) - n).max() < 10 * np.finfo(float).eps
    # n is the right-hand normal of (p2-p1, p3-p1)
    c = np.cross(nodes[:, 1] - nodes[:, 0], nodes[:, 2] - nodes[:, 0])
    assert np.abs(n - c / np.linalg.norm(c)).max() < 1e-15


@pytest.mark.parametrize("elem", ["source", "doublet"])
def test_U_is_grad_phi(orc, pentagon, elem):
    nodes, panel = pentagon
    rng = np.random.default_rng(11)
    for _ in range(4):
        X = rng.random(3)

        def phi(Y):
            out = np.zeros(1)
            if elem == "source":
                orc.phi_constant_source(nodes, panel, 1.0, Y.reshape(3, 1), out, offset=0.0)
            else:
                orc.phi_constant_doublet(nodes, panel, 1.0, Y.reshape(3, 1), out)
            return out

        U = np.zeros((3, 1), order="F")
        if elem == "source":
            orc.U_constant_source(nodes, panel, 1.0, X.reshape(3, 1), U, offset=0.0)
        else:
            orc.U_vortexring(nodes, panel, 1.0, X.reshape(3, 1), U, cutoff=0.0, offset=0.0)
        g = jac(phi, X)[0]
        assert np.abs(g - U[:, 0]).max() <= 1e-9 * np.abs(U).max()


@pytest.mark.parametrize("elem", ["source", "doublet"])
def test_div_free(orc, pentagon, elem):
    nodes, panel = pentagon
    rng = np.random.default_rng(5)
    for _ in range(4):
        X = rng.random(3)

        def U(Y):
            out = np.zeros((3, 1), order="F")
            if elem == "source":
                orc.U_constant_source(nodes, panel, 1.0, Y.reshape(3, 1), out, offset=0.0)
            else:
                orc.U_vortexring(nodes, panel, 1.0, Y.reshape(3, 1), out, cutoff=0.0, offset=0.0)
            return out[:, 0]

        J = jac(U, X)
        assert abs(np.trace(J)) <= 1e-9 * np.abs(J).max()
        # irrotational too
        curl = np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])
        assert np.abs(curl).max() <= 1e-8 * np.abs(J).max()


def unit_square():
    return np.asfortranarray(np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], float).T), np.arange(4)


@pytest.mark.parametrize("hsign", [1.0, -1.0])
def test_source_jump_and_far_field(orc, hsign):
    nodes, panel = unit_square()
    off = np.array([0.231, -0.417, 0.0])
    h = hsign * 1e-8
    T = np.asfortranarray(np.stack([off, off + [0, 0, hsign * 1e10], off + [0, 0, h]], axis=1))
    phis = np.zeros(3)
    orc.phi_constant_source(nodes, panel, 1.0, T, phis, offset=0.0)
    assert abs(phis[1]) < 1e-5
    assert abs((phis[2] - phis[0]) / h - np.sign(h) * 0.5) < 1e-6
    # the velocity jump: normal velocity +-sigma/2 just above/below
    U = np.zeros((3, 1), order="F")
    orc.U_constant_source(nodes, panel, 1.0, T[:, 2:3], U, offset=0.0)
    assert abs(U[2, 0] - np.sign(h) * 0.5) < 1e-6


@pytest.mark.parametrize("hsign", [1.0, -1.0])
def test_doublet_jump_and_far_field(orc, hsign):
    nodes, panel = unit_square()
    off = np.array([-0.35, 0.52, 0.0])
    h = hsign * 1e-8
    T = np.asfortranarray(np.stack([off, off + [0, 0, hsign * 1e10], off + [0, 0, h]], axis=1))
    phis = np.zeros(3)
    orc.phi_constant_doublet(nodes, panel, 1.0, T, phis)
    assert abs(phis[1]) < 1e-7
    assert abs(phis[2] - (-np.sign(h)) * 0.5) < 1e-6


def test_source_potential_quadrature(orc):
    from scipy import integrate
    nodes, panel = unit_square()
    X = np.array([0.3, 1.7, 0.9])
    out = np.zeros(1)
    orc.phi_constant_source(nodes, panel, 1.0, X.reshape(3, 1), out, offset=0.0)
    val, _ = integrate.dblquad(lambda y, x: 1.0 / np.linalg.norm(X - [x, y, 0]), -1, 1, -1, 1,
                               epsabs=1e-13, epsrel=1e-12)
    assert abs(out[0] - (-val / (4 * np.pi))) < 1e-10


def test_doublet_potential_quadrature(orc):
    from scipy import integrate
    nodes, panel = unit_square()
    X = np.array([-0.4, 1.6, -0.7])
    out = np.zeros(1)
    orc.phi_constant_doublet(nodes, panel, 1.0, X.reshape(3, 1), out)

    def f(y, x):
        d = X - np.array([x, y, 0.0])
        return d[2] / np.linalg.norm(d) ** 3

    val, _ = integrate.dblquad(f, -1, 1, -1, 1, epsabs=1e-13, epsrel=1e-12)
    assert abs(out[0] - (-val / (4 * np.pi))) < 1e-10


def test_vortexring_biot_savart_quadrature(orc, pentagon):
    nodes, panel = pentagon
    X = np.array([0.3, -0.2, 0.8])
    U = np.zeros((3, 1), order="F")
    orc.U_vortexring(nodes, panel, 1.0, X.reshape(3, 1), U, cutoff=0.0, offset=0.0)
    # Gauss-Legendre along each edge: u = G/(4 pi) int dl x (X - l) / |X - l|^3
    s, w = np.polynomial.legendre.leggauss(200)
    ref = np.zeros(3)
    nv = len(panel)
    for k in range(nv):
        a, b = nodes[:, k], nodes[:, (k + 1) % nv]
        for sk, wk in zip(s, w):
            p = a + (sk + 1) / 2 * (b - a)
            r = X - p
            ref += wk / 2 * np.cross(b - a, r) / np.linalg.norm(r) ** 3
    ref /= 4 * np.pi
    assert np.abs(U[:, 0] - ref).max() < 1e-12 * max(1, np.abs(ref).max())


def test_point_source_far_field(orc):
    nodes, panel = unit_square()
    X = np.array([30.0, -40.0, 120.0])
    U = np.zeros((3, 1), order="F")
    orc.U_constant_source(nodes, panel, 1.0, X.reshape(3, 1), U, offset=0.0)
    r = np.linalg.norm(X)
    ref = 4.0 / (4 * np.pi) * X / r ** 3  # sigma * area / (4 pi r^2) along r
    assert np.abs(U[:, 0] - ref).max() < 1e-4 * np.abs(ref).max()


def test_semiinfinite_vortex_long_segment_limit(orc):
    p = np.array([0.2, -0.3, 0.1])
    d = np.array([1.0, 2.0, -0.5])
    d /= np.linalg.norm(d)
    X = np.asfortranarray(np.array([[0.5, 0.7, -0.4], [-1.0, 0.3, 0.9]]).T)
    U = np.zeros((3, 2), order="F")
    orc.U_semiinfinite_vortex(p, d, 1.0, X, U, cutoff=0.0, offset=0.0)
    L = 1e7
    R = np.zeros((3, 2), order="F")
    orc.U_boundvortex(p, p + L * d, 1.0, X, R, cutoff=0.0, offset=0.0)
    assert np.abs(U - R).max() < 1e-6 * np.abs(R).max()


@pytest.fixture(scope="module")
def te_panel():
    # trailing edge of test/runtests_semiinfiniteelements.jl:31-51, rotated and translated
    R = rot(30, 10, 40)
    pa = R.T @ np.array([-2.0, -1.0, 0.0]) + 1.0
    pb = R.T @ np.array([0.0, 1.0, 0.0]) + 1.0
    d = R.T @ np.array([1.0, 0.0, 0.0])
    return pa, pb, d


def test_semiinfinite_U_is_grad_phi(orc, te_panel):
    pa, pb, d = te_panel
    rng = np.random.default_rng(2)
    for _ in range(4):
        X = rng.random(3)

        def phi(Y):
            out = np.zeros(1)
            orc.phi_semiinfinite_doublet(pa, pb, d, 1.0, Y.reshape(3, 1), out)
            return out

        U = np.zeros((3, 1), order="F")
        orc.U_semiinfinite_horseshoe(pa, pb, d, d, 1.0, X.reshape(3, 1), U, cutoff=0.0, offset=0.0)
        g = jac(phi, X)[0]
        assert np.abs(g - U[:, 0]).max() <= 1e-9 * np.abs(U).max()

        def Uf(Y):
            out = np.zeros((3, 1), order="F")
            orc.U_semiinfinite_horseshoe(pa, pb, d, d, 1.0, Y.reshape(3, 1), out, cutoff=0.0, offset=0.0)
            return out[:, 0]

        J = jac(Uf, X)
        assert abs(np.trace(J)) <= 1e-9 * np.abs(J).max()


@pytest.mark.parametrize("hsign", [1.0, -1.0])
def test_semiinfinite_doublet_jump(orc, hsign):
    pa, pb, d = np.array([-2.0, -1.0, 0.0]), np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0])
    h = hsign * 1e-8
    T = np.asfortranarray(np.array([[0, 0, hsign * 1e14], [0, 0, h]]).T)
    out = np.zeros(2)
    orc.phi_semiinfinite_doublet(pa, pb, d, 1.0, T, out)
    assert abs(out[0]) < 1e-8
    assert abs(out[1] - np.sign(h) * 0.5) < 1e-8


def test_nonplanar_semiinfinite_reduces_to_planar(orc, te_panel):
    pa, pb, d = te_panel
    X = np.asfortranarray(np.array([[0.3, 0.2, 0.1], [2.0, -1.0, 0.5]]).T)
    a = np.zeros(2)
    b = np.zeros(2)
    orc.phi_semiinfinite_doublet(pa, pb, d, 1.0, X, a)
    orc.phi_semiinfinite_doublet2(pa, pb, d, d, 1.0, X, b)
    assert np.abs(a - b).max() < 1e-15


def test_nonunit_direction_is_an_error(orc):
    with pytest.raises(orc.OracleError):
        orc.U_semiinfinite_vortex(np.zeros(3), np.array([1.0, 0, 1e-3]), 1.0, np.ones((3, 1)), np.zeros((3, 1)))


def test_offset_cutoff_regularisation(orc):
    # target on the segment line: the cutoff zeroes the self-induced velocity
    U = np.zeros((3, 1), order="F")
    orc.U_boundvortex(np.zeros(3), np.array([1.0, 0, 0]), 1.0, np.array([[2.0], [0.0], [0.0]]), U)
    assert np.all(U == 0)


@pytest.fixture(scope="module")
def te_nonplanar(te_panel):
    # the same trailing edge with db turned 0.3 rad out of the da-plane: the
    # non-planar semi-infinite sheet (src/FLOWPanel_elements.jl:1110-1244)
    R = rot(30, 10, 40)
    pa, pb, da = te_panel
    db = R.T @ np.array([np.cos(0.3), 0.0, np.sin(0.3)])
    return pa, pb, da / np.linalg.norm(da), db / np.linalg.norm(db)


def _phi2(orc, pa, pb, da, db):
    def f(Y):
        out = np.zeros(1)
        orc.phi_semiinfinite_doublet2(pa, pb, da, db, 1.0, Y.reshape(3, 1), out)
        return out
    return f


def test_nonplanar_semiinfinite_U_is_grad_phi(orc, te_nonplanar):
    # A constant doublet sheet equals the vortex loop on its boundary, so the
    # gradient of the non-planar sheet (planar strip along da plus the triangle
    # (pb, pb + L da, pb + L db)) is the horseshoe velocity with legs da and db
    # (U_semiinfinite_doublet == U_semiinfinite_horseshoe, src/FLOWPanel_elements.jl:948).
    pa, pb, da, db = te_nonplanar
    assert np.dot(da, db) > 0
    rng = np.random.default_rng(5)
    phi = _phi2(orc, pa, pb, da, db)
    for _ in range(4):
        X = 2 * rng.random(3)
        U = np.zeros((3, 1), order="F")
        orc.U_semiinfinite_horseshoe(pa, pb, da, db, 1.0, X.reshape(3, 1), U, cutoff=0.0, offset=0.0)
        g = jac(phi, X, h=1e-4)[0]
        assert np.abs(g - U[:, 0]).max() <= 1e-9 * np.abs(U).max()


def test_nonplanar_semiinfinite_check_has_teeth(orc, te_nonplanar):
    # A planted mistake, built from the oracle on the CPU: the correction triangle
    # with reversed orientation, phi_planar - triangle = 2 phi_planar - phi2.  The
    # gradient check above must reject it.
    pa, pb, da, db = te_nonplanar
    X = np.array([0.7, 1.3, 0.4])
    good = _phi2(orc, pa, pb, da, db)

    def planar(Y):
        out = np.zeros(1)
        orc.phi_semiinfinite_doublet(pa, pb, da, 1.0, Y.reshape(3, 1), out)
        return out

    U = np.zeros((3, 1), order="F")
    orc.U_semiinfinite_horseshoe(pa, pb, da, db, 1.0, X.reshape(3, 1), U, cutoff=0.0, offset=0.0)
    bad = jac(lambda Y: 2 * planar(Y) - good(Y), X, h=1e-4)[0]
    assert np.abs(bad - U[:, 0]).max() > 1e-3 * np.abs(U).max()
    # and a planar-only sheet (the triangle dropped) fails as well
    drop = jac(planar, X, h=1e-4)[0]
    assert np.abs(drop - U[:, 0]).max() > 1e-3 * np.abs(U).max()


@pytest.mark.parametrize("hsign", [1.0, -1.0])
def test_nonplanar_semiinfinite_jump_on_the_triangle(orc, hsign):
    # phi(+-0) = -+mu/2 across the correction triangle, far from the strip:
    # da = x, db in the x-z plane, the TE along y; a point above the triangle
    # (pb, pb + L da, pb + L db) at distance h from it.
    pa, pb = np.array([0.0, -1.0, 0.0]), np.array([0.0, 1.0, 0.0])
    da = np.array([1.0, 0.0, 0.0])
    db = np.array([np.cos(0.4), 0.0, np.sin(0.4)])
    # the triangle's unit normal (da x db) is -y: a point on the triangle's plane
    # (y = 1) between the legs, then offset along -y by h
    nz = np.cross(da, db)
    nz /= np.linalg.norm(nz)
    P = pb + 5.0 * (np.cos(0.2) * np.array([1.0, 0, 0]) + np.sin(0.2) * np.array([0, 0, 1.0]))
    h = hsign * 1e-8
    T = (P + h * nz).reshape(3, 1)
    out = np.zeros(1)
    orc.phi_semiinfinite_doublet2(pa, pb, da, db, 1.0, np.asfortranarray(T), out)
    # the planar strip along da lies at y in [-1, 1], z = 0: its solid angle at P
    # (distance ~ 5 sin(0.2) off its plane, outside it) is smooth across the
    # triangle; the jump of the sheet's potential across the triangle is mu
    out2 = np.zeros(1)
    T2 = (P - h * nz).reshape(3, 1)
    orc.phi_semiinfinite_doublet2(pa, pb, da, db, 1.0, np.asfortranarray(T2), out2)
    assert abs((out[0] - out2[0]) - np.sign(h) * 1.0) < 1e-6
