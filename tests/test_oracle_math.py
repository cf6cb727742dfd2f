"""Pins of the oracle that come from the mathematics, independent of the
reference's own outputs: closed-form Legendre values, exactness of the
quadrature rules, the operator matrices against an independent high-order
quadrature, Cholesky against numpy, the reduced (zero-on-edge-1) basis,
conservation, RK4 order, and the analytic standing wave the reference's own
compare.f checks against (src/compare.f:81-87, 152-160)."""
import math

import numpy as np
import numpy.polynomial.legendre as L
import pytest

from conftest import oracle_for, quad_rule
from spoc_amd import configs

RULES = [(6, 4), (12, 6), (25, 10), (33, 12), (42, 14)]  # points, exact degree


def duffy(n=24):
    """Gauss-Legendre tensor rule on the reference triangle
    {xi, eta >= -1, xi + eta <= 0} through the collapsed map; exact for
    polynomials of total degree < 2n - 1."""
    s, w = L.leggauss(n)
    S, T = np.meshgrid(s, s, indexing="ij")
    W = np.outer(w, w) * (1 - T) / 2
    xi = (1 + S) * (1 - T) / 2 - 1
    return xi.ravel(), T.ravel(), W.ravel()


def test_legendre_closed_form():
    O = oracle_for(3, 12, 4)
    x = np.linspace(-1, 1, 13)
    for n in range(0, 9):
        c = np.zeros(n + 1)
        c[n] = 1.0
        ref = L.legval(x, c)
        dref = L.legval(x, L.legder(c))
        got = np.array([O.lib.orc_po(n, float(v)) for v in x])
        dgot = np.array([O.lib.orc_dpo(n, float(v)) for v in x])
        assert np.allclose(got, ref, atol=1e-14), n
        assert np.allclose(dgot, dref, atol=1e-12), n


@pytest.mark.parametrize("ng,deg", RULES)
def test_triangle_rules_exact_on_monomials(ng, deg):
    w, x, y, _, _ = quad_rule(ng, 3)
    X, Y, W = duffy()
    assert abs(w.sum() - 2.0) < 1e-13
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = np.sum(W * X ** a * Y ** b)
            assert abs(np.sum(w * x ** a * y ** b) - exact) < 2e-13, (a, b)
    # and not beyond (the rule degree is what the point count promises)
    errs = [abs(np.sum(w * x ** a * y ** (deg + 1 - a)) - np.sum(W * X ** a * Y ** (deg + 1 - a)))
            for a in range(deg + 2)]
    assert max(errs) > 1e-10


@pytest.mark.parametrize("ngf", range(3, 10))
def test_gauss_legendre_edge_rules(ngf):
    _, _, _, wf, tf = quad_rule(12, ngf)
    t, w = L.leggauss(ngf)
    assert np.allclose(np.sort(tf), t, atol=1e-15) and np.allclose(wf[np.argsort(tf)], w, atol=1e-15)


def _basis(nc, x, y):
    out = []
    for a in range(nc + 1):
        for b in range(nc + 1 - a):
            ca = np.zeros(a + 1)
            ca[a] = 1
            cb = np.zeros(b + 1)
            cb[b] = 1
            out.append((L.legval(x, ca) * L.legval(y, cb),
                        L.legval(x, L.legder(ca)) * L.legval(y, cb),
                        L.legval(x, ca) * L.legval(y, L.legder(cb))))
    return out


@pytest.mark.parametrize("nc,ng,ngf", [(3, 12, 4), (5, 25, 6), (6, 33, 7), (7, 42, 8)])
def test_operator_matrices_against_independent_quadrature(nc, ng, ngf):
    O = oracle_for(nc, ng, ngf)
    t = O.tables()
    nm = O.nm
    X, Y, W = duffy()
    B = _basis(nc, X, Y)
    M = np.array([[np.sum(W * bi[0] * bj[0]) for bj in B] for bi in B])
    GX = np.array([[np.sum(W * bi[0] * bj[1]) for bj in B] for bi in B])
    GY = np.array([[np.sum(W * bi[0] * bj[2]) for bj in B] for bi in B])
    scale = np.abs(M).max()
    assert np.abs(t["ammo"][:nm, :nm] - M).max() < 1e-12 * scale
    assert np.abs(t["agx"][:nm, :nm] - GX).max() < 1e-11 * scale
    assert np.abs(t["agy"][:nm, :nm] - GY).max() < 1e-11 * scale
    # Cholesky factor of the reference's DPOTRF: U^T U = M, numpy's factor
    U = np.triu(t["amm"][:nm, :nm])
    assert np.allclose(U.T @ U, M, rtol=0, atol=1e-12 * scale)
    # the same factorisation as numpy's of the same matrix (to the forward
    # error cond(U) * eps allows)
    A = t["ammo"][:nm, :nm]
    tol = 1e-16 * np.linalg.cond(A) * np.abs(U).max()  # first-order forward bound
    assert np.abs(U - np.linalg.cholesky(A).T).max() < max(tol, 1e-13)


@pytest.mark.parametrize("nc,ng,ngf", [(3, 12, 4), (6, 33, 7)])
def test_reduced_basis_vanishes_on_edge1_and_amml_is_its_gram(nc, ng, ngf):
    O = oracle_for(nc, ng, ngf)
    t = O.tables()
    nm, nl = O.nm, O.nm - nc - 1
    # R from reduc_gn applied to unit vectors; R^T from reduc_inv_gn
    import ctypes as C
    R = np.zeros((nl, nm))
    for k in range(nm):
        e = np.zeros(64)
        e[k] = 1.0
        r = np.zeros(64)
        O.lib.orc_reduc_gn(O._tab, e.ctypes.data_as(C.POINTER(C.c_double)),
                           r.ctypes.data_as(C.POINTER(C.c_double)))
        R[:, k] = r[:nl]
    for c in range(nl):
        r = np.zeros(64)
        r[c] = 1.0
        g = np.zeros(64)
        O.lib.orc_reduc_inv_gn(O._tab, g.ctypes.data_as(C.POINTER(C.c_double)),
                               r.ctypes.data_as(C.POINTER(C.c_double)))
        assert np.array_equal(g[:nm], R[c, :]), c
    # the reduced functions sum_k R[c,k] phi_k vanish on edge 1 (eta = -1)
    xs = np.linspace(-1, 1, 9)
    B = _basis(nc, xs, -np.ones_like(xs))
    V = np.array([b[0] for b in B])  # (nm, 9)
    assert np.abs(R @ V).max() < 1e-13
    M = t["ammo"][:nm, :nm]
    UL = np.triu(t["amml"][:nl, :nl])
    assert np.allclose(UL.T @ UL, R @ M @ R.T, atol=1e-12)


def test_linear_mass_conservation_and_rk4_order():
    O = oracle_for(3, 12, 4)
    t = O.tables()
    w, x, y, wf, tf = quad_rule(12, 4)
    t.update(nc=3, ng=12, ngf=4, nm=10, wi=w, xi=x, yi=y, wif=wf, tif=tf)
    c = configs.config1(t)
    m, p = c["mesh"], c["phys"]
    mass0 = O.energy(m, p, *c["state"])[2]
    out = O.rk4("lin", m, p, *c["state"], 20.0, 20)
    mass1 = O.energy(m, p, *out)[2]
    assert abs(mass1 - mass0) < 1e-11 * abs(mass0)
    # RK4 is fourth order in dt on the (linear) semi-discrete system
    T = 80.0  # dt = 20, 10, 5 s: inside the CFL limit of this mesh
    sols = {n: O.rk4("lin", m, p, *c["state"], T / n, n)[2] for n in (4, 8, 16, 256)}
    e1 = np.linalg.norm(sols[4] - sols[256])
    e2 = np.linalg.norm(sols[8] - sols[256])
    e3 = np.linalg.norm(sols[16] - sols[256])
    assert 10 < e1 / e2 < 20 and 13 < e2 / e3 < 19, (e1 / e2, e2 / e3)  # -> 16


def standing_wave_error(O, t, n):
    """Linear RHS against the analytic standing wave of src/compare.f:81-87:
    eta = cos(k x) cos(w t), u = (g/c) sin(k x) sin(w t), c = sqrt(g H) =
    sqrt(10), compared at w t = 0.2 pi in the L2 norm."""
    lx = 1000e3
    g0, H = 10.0, 1.0
    x, y, tri = configs.structured_mesh(n + 1, n + 1, 0.0, lx, 0.0, lx)
    k = 2 * math.pi / lx
    cs = configs.make_case(t, x, y, tri, g0=g0, H=H, eta=lambda X, Y: np.cos(k * X))
    c0 = math.sqrt(g0 * H)
    Tend = 0.2 * math.pi / (c0 * k)
    nst = 4 * n
    u, v, h = O.rk4("lin", cs["mesh"], cs["phys"], *cs["state"], Tend / nst, nst)
    href = configs.project(t, x, y, tri, lambda X, Y: math.cos(0.2 * math.pi) * np.cos(k * X))
    uref = configs.project(t, x, y, tri, lambda X, Y: math.sin(0.2 * math.pi) * g0 / c0 *
                           np.sin(k * X))
    nm = t["nm"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / cs["mesh"]["vect"][4]

    def l2(a):
        return math.sqrt(np.sum(det * np.einsum("ie,ij,je->e", a[:nm], M, a[:nm])))

    return l2(h - href) / l2(href), l2(u - uref) / l2(uref), l2(v) / l2(uref)


def test_standing_wave_known_answer_and_convergence():
    O = oracle_for(3, 12, 4)
    t = O.tables()
    w, x, y, wf, tf = quad_rule(12, 4)
    t.update(nc=3, ng=12, ngf=4, nm=10, wi=w, xi=x, yi=y, wif=wf, tif=tf)
    e8 = standing_wave_error(O, t, 4)
    e16 = standing_wave_error(O, t, 8)
    assert e16[0] < 2e-3 and e16[1] < 2e-3 and e16[2] < 1e-3, e16
    # p = 3 converges at order >= 3 in h
    assert e8[0] / e16[0] > 6 and e8[1] / e16[1] > 6, (e8, e16)
