"""Mesh refinement and coarsening (spoc_amr_*, spoc_amd/csrc/amr.cpp) against
the reference's refinement hierarchy (src/adapt.f:380-1258): the flags of
tests/golden/amr_cycles.npz replayed through the library give the
reference's meshes, tabsol and tinter bit for bit, cycle after cycle; every
mesh is conforming (the pointer routines accept it), counter-clockwise,
covers the same area, and differs by at most one level across a face."""
import os

import numpy as np
import pytest

import spoc_amd
from spoc_amd import mesh

G = os.path.join(os.path.dirname(__file__), "golden", "amr_cycles.npz")


def _area(x, y, tri):
    a, b, c = tri - 1
    return 0.5 * ((x[b] - x[a]) * (y[c] - y[a]) - (x[c] - x[a]) * (y[b] - y[a]))


@pytest.mark.parametrize("name", ["c1", "jit"])
def test_amr_cycles_match_reference(name):
    g = np.load(G)
    x, y, tri = g[f"{name}_x"], g[f"{name}_y"], g[f"{name}_tri"]
    m = mesh.build(x, y, tri, 3)
    assert np.array_equal(m["tseg"], g[f"{name}_tseg"]) and np.array_equal(m["itm"], g[f"{name}_itm"])
    amr = spoc_amd.Amr(m)
    area0 = _area(x, y, tri).sum()
    ct = tri
    for cyc in range(int(g[f"{name}_cycles"])):
        k = f"{name}_{cyc}"
        nx, ny, nt = amr.adapt(g[k + "_fconv"], g[k + "_fderaff"])
        tabsol = amr.mesh()[3]
        assert np.array_equal(nx, g[k + "_x"]) and np.array_equal(ny, g[k + "_y"]), k
        assert np.array_equal(nt, g[k + "_tri"]), k
        assert np.array_equal(tabsol, g[k + "_tabsol"]), k
        assert np.array_equal(spoc_amd.inter_only_new(ct, nt), g[k + "_tinter"]), k
        # a valid mesh: the pointer routines (with the mesh check) accept it,
        # every triangle is CCW and the area is the original's
        a = _area(nx, ny, nt)
        assert (a > 0).all()
        assert abs(a.sum() - area0) <= 1e-9 * area0
        mm = mesh.build(nx, ny, nt, 3)
        lev = amr.levels()
        inner = mm["tseg"][1] > 0
        e1, e2 = mm["tseg"][0, inner] - 1, mm["tseg"][1, inner] - 1
        assert np.abs(lev[e1] - lev[e2]).max() <= 1
        ct = nt


def test_amr_refine_everything_is_a_4_split():
    """Refining every triangle of a mesh at level 0 gives four triangles per
    triangle through the edge midpoints and no green halves."""
    g = np.load(G)
    x, y, tri = g["jit_x"], g["jit_y"], g["jit_tri"]
    m = mesh.build(x, y, tri, 3)
    amr = spoc_amd.Amr(m)
    ne = tri.shape[1]
    nx, ny, nt = amr.adapt(np.ones(ne, np.int32))
    assert nt.shape[1] == 4 * ne
    assert len(nx) == len(x) + m["nf"]  # one new node per face
    assert (amr.levels() == 1).all()
    np.testing.assert_allclose(_area(nx, ny, nt).sum(), _area(x, y, tri).sum(), rtol=1e-12)


def test_amr_coarsening_back_to_the_start():
    """Refining one triangle and then flagging its four children for
    coarsening gives back the original triangles (the node numbering of the
    hierarchy drops the midpoints again)."""
    g = np.load(G)
    x, y, tri = g["c1_x"], g["c1_y"], g["c1_tri"]
    m = mesh.build(x, y, tri, 3)
    amr = spoc_amd.Amr(m)
    ne = tri.shape[1]
    f = np.zeros(ne, np.int32)
    f[200] = 1
    amr.adapt(f)
    _, _, t1, tabsol = amr.mesh()
    kids = np.arange(ne + 1, ne + 5)
    assert set(kids) <= set(tabsol.tolist())
    fder = np.zeros(t1.shape[1], np.int32)
    fder[kids - 1] = 1  # (read by hierarchy number, as the reference reads it)
    nx, ny, nt = amr.adapt(np.zeros(t1.shape[1], np.int32), fder)
    assert np.array_equal(nx, x) and np.array_equal(ny, y)
    assert np.array_equal(nt, tri)


def test_amr_rejects_bad_input():
    g = np.load(G)
    x, y, tri = g["c1_x"], g["c1_y"], g["c1_tri"]
    m = mesh.build(x, y, tri, 3)
    bad = dict(m)
    bad["itm"] = np.asfortranarray(np.minimum(m["itm"] * 0 + 10 ** 6, 10 ** 6))
    with pytest.raises(spoc_amd.SpocError):
        spoc_amd.Amr(bad)
    amr = spoc_amd.Amr(m)
    with pytest.raises(spoc_amd.SpocError):
        amr.adapt(np.zeros(5, np.int32))  # wrong length
    # refine a triangle twice, then ask to coarsen the first family while one
    # of its children is refined: refused, hierarchy unchanged
    ne = tri.shape[1]
    f = np.zeros(ne, np.int32)
    f[100] = 1
    amr.adapt(f)
    _, _, t1, tabsol = amr.mesh()
    e = int(np.where(tabsol == ne + 1)[0][0])
    f = np.zeros(t1.shape[1], np.int32)
    f[e] = 1
    amr.adapt(f)
    _, _, t2, tabsol2 = amr.mesh()
    fder = np.zeros(t2.shape[1], np.int32)
    fder[ne:ne + 4] = 1  # the first family's numbers, one of them refined
    with pytest.raises(spoc_amd.SpocError):
        amr.adapt(np.zeros(t2.shape[1], np.int32), fder)
    assert np.array_equal(amr.mesh()[2], t2)
    # and the hierarchy still adapts
    amr.adapt(np.zeros(t2.shape[1], np.int32))
    assert np.array_equal(amr.mesh()[2], t2)
