"""The adaptive driver's re-projection onto a new mesh (spoc_interpol,
spoc_calh2; host C++) against golden vectors written by the reference
Fortran itself (interpol, interlin, calh2, src/adapt.f:1662-1824, 303-373),
bit for bit: onto a 4-split refinement with some elements kept, and onto a
larger mesh whose outside points take interlin's wall projection.  Plus what
the mathematics fixes: a polynomial of degree <= p given on every element
is carried onto any mesh inside the old one unchanged (to rounding), and a
field projected onto its own mesh is itself."""
import numpy as np
import pytest

import spoc_amd
from conftest import gold, host_tables
from spoc_amd import configs, mesh

CASES = [(3, 12, 4), (6, 33, 7)]


def _ff(x, y):  # the harness's test function, same expression
    return x * 1.0e-6 * y * 2.0e-6 + 3.0 * x * 1.0e-6 - y * y * 1.0e-12 + 7.0


def _case(nc, ng, ngf):
    g = gold(f"interp_p{nc}.npz")
    t = host_tables(nc, ng, ngf)
    c = configs.random_case(t, int(g["nx"]), int(g["ny"]), seed=int(g["seed"]), scale=0.01)
    nm = t["nm"]
    fields = [c["state"][0], c["state"][2], c["phys"]["hb"]]
    for f, u in enumerate(fields):
        assert np.array_equal(u[:nm], g[f"u{f}"])
    return t, c["mesh"], fields, g


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_interpol_onto_refinement_bit_exact(nc, ng, ngf):
    t, m, fields, g = _case(nc, ng, ngf)
    out = spoc_amd.interpol(t, m, g["xa"], g["ya"], g["ta"], g["tia"], fields)
    for f in range(3):
        assert np.array_equal(out[f][:t["nm"]], g[f"a{f}"]), f
    kept = np.nonzero(g["tia"])[0]
    assert len(kept) > 0
    for i in kept:
        assert np.array_equal(out[0][:, i], fields[0][:, g["tia"][i] - 1])


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_interpol_onto_larger_mesh_bit_exact(nc, ng, ngf):
    t, m, fields, g = _case(nc, ng, ngf)
    tib = np.zeros(g["tb"].shape[1], np.int32)
    out = spoc_amd.interpol(t, m, g["xb"], g["yb"], g["tb"], tib, fields)
    for f in range(3):
        assert np.array_equal(out[f][:t["nm"]], g[f"b{f}"]), f


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_calh2_bit_exact(nc, ng, ngf):
    t, m, fields, g = _case(nc, ng, ngf)
    h1 = spoc_amd.calh2(t, _ff, g["xa"], g["ya"], g["ta"], g["tia"], fields[2])
    assert np.array_equal(h1[:t["nm"]], g["calh2"])


def _poly_field(t, m, deg):
    """Modal coefficients of the global polynomial x^a y^b sum (degree deg)
    on every element, by calh2 (exact projection of a polynomial)."""
    s = 1.0e-6

    def f(x, y):
        return 1.0 + (x * s) ** deg - 0.5 * (x * s) * (y * s) ** (deg - 1) + 0.25 * (y * s)

    return spoc_amd.calh2(t, f, m["x"], m["y"], m["in"]), f


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_polynomial_carried_unchanged(nc, ng, ngf):
    t = host_tables(nc, ng, ngf)
    c = configs.random_case(t, 8, 8, seed=5, scale=0.01)
    m = c["mesh"]
    u, f = _poly_field(t, m, nc)
    # a new mesh strictly inside the old one (a shrunken jittered mesh)
    xn, yn, tn = configs.jittered_mesh(11, 10, 0.9e6, 0.85e6, 0.25, 17)
    xn, yn = xn + 0.05e6, yn + 0.07e6
    out = spoc_amd.interpol(t, m, xn, yn, tn, np.zeros(tn.shape[1], np.int32), [u])[0]
    ref = spoc_amd.calh2(t, f, xn, yn, tn)
    # compare the fields at the quadrature points: the modal coefficients of
    # this basis are ill-conditioned at high p (mass matrix cond ~1e8 at
    # p=6), the values are not
    nm, ng = t["nm"], t["ng"]
    P = t["pi2"][:ng, :nm]
    vo, vr = P @ out[:nm], P @ ref[:nm]
    err = np.abs(vo - vr).max() / np.abs(vr).max()
    assert err < 1e-10, err  # (rounding through the cond ~1e8 mass matrix at p=6)


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_projection_onto_own_mesh_is_identity(nc, ng, ngf):
    t = host_tables(nc, ng, ngf)
    c = configs.random_case(t, 7, 6, seed=8, scale=0.01)
    m = c["mesh"]
    u = c["state"][1]
    out = spoc_amd.interpol(t, m, m["x"], m["y"], m["in"], np.zeros(m["ne"], np.int32), [u])[0]
    nm, ng = t["nm"], t["ng"]
    P = t["pi2"][:ng, :nm]
    vo, vu = P @ out[:nm], P @ u[:nm]
    assert np.abs(vo - vu).max() < 1e-10 * np.abs(vu).max()


def _bad(arr, idx, val):
    a = np.array(arr, copy=True, order="F")
    a.flat[idx] = val
    return a


@pytest.mark.parametrize("key,idx,val", [("in", 5, 0), ("in", 7, 10 ** 6), ("itm", 3, 0),
                                         ("itm", 4, 10 ** 6), ("tseg", 0, 0), ("tseg", 1, -1),
                                         ("tseg", 3, 10 ** 6), ("pseg", 2, 0),
                                         ("pseg", 5, 10 ** 6)])
def test_interpol_rejects_malformed_old_mesh(key, idx, val):
    """Index arrays pointing outside their ranges are an error before any
    walk reads through them (the reference would read out of bounds)."""
    t = host_tables(3, 12, 4)
    c = configs.random_case(t, 5, 4, seed=3, scale=0.01)
    m = dict(c["mesh"])
    m[key] = _bad(m[key], idx, val)
    u = c["state"][0]
    with pytest.raises(spoc_amd.SpocError, match=key):
        spoc_amd.interpol(t, m, m["x"], m["y"], c["mesh"]["in"], np.zeros(m["ne"], np.int32), [u])


@pytest.mark.parametrize("what", ["innew", "tinter"])
def test_interpol_rejects_malformed_new_mesh(what):
    t = host_tables(3, 12, 4)
    c = configs.random_case(t, 5, 4, seed=3, scale=0.01)
    m = c["mesh"]
    tri = np.array(m["in"], order="F")
    tinter = np.zeros(m["ne"], np.int32)
    if what == "innew":
        tri[1, 2] = len(m["x"]) + 1
    else:
        tinter[4] = m["ne"] + 1
    with pytest.raises(spoc_amd.SpocError, match=what):
        spoc_amd.interpol(t, m, m["x"], m["y"], tri, tinter, [c["state"][0]])


def test_calh2_rejects_malformed_input():
    t = host_tables(3, 12, 4)
    c = configs.random_case(t, 5, 4, seed=3, scale=0.01)
    m = c["mesh"]
    tri = np.array(m["in"], order="F")
    tri[0, 1] = 0
    with pytest.raises(spoc_amd.SpocError, match="in"):
        spoc_amd.calh2(t, _ff, m["x"], m["y"], tri)
    h0 = c["phys"]["hb"]
    tinter = np.zeros(m["ne"], np.int32)
    tinter[2] = h0.shape[1] + 1  # past h0's elements
    with pytest.raises(spoc_amd.SpocError, match="tinter"):
        spoc_amd.calh2(t, _ff, m["x"], m["y"], m["in"], tinter, h0)
    tinter[2] = 1  # copies from h0, which is missing
    with pytest.raises(spoc_amd.SpocError, match="tinter"):
        spoc_amd.calh2(t, _ff, m["x"], m["y"], m["in"], tinter, None)
