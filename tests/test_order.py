"""The library's internal element numbering (spoc_mesh_order, host side):
a permutation inside each ensemble member that puts face neighbours close
in index, whatever the caller's numbering.  (That results do not depend on
it is a GPU test: test_gpu_order.py.)"""
import numpy as np

import spoc_amd
from conftest import host_tables
from spoc_amd import configs
from spoc_amd import mesh as M


def _spread(m, perm):
    """Mean |internal index distance| between face neighbours, and the share
    of neighbour pairs in the same block of 16."""
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    t = np.asarray(m["tseg"]) - 1
    inner = t[1] >= 0
    a, b = inv[t[0, inner]], inv[t[1, inner]]
    return float(np.abs(a - b).mean()), float(np.mean(a // 16 == b // 16))


def _shuffled(m, seed):
    perm = np.random.default_rng(seed).permutation(m["ne"])
    return M.build(m["x"], m["y"], np.asfortranarray(m["in"][:, perm]), 3, 1e-4, 2e-11)


def test_order_is_a_permutation_with_locality():
    t = host_tables(3, 12, 4)
    x, y, tri = configs.jittered_mesh(61, 31, 4000e3, 2000e3, 0.3, 7)
    zm = M.build(x, y, tri, 3, 1e-4, 2e-11)  # the generator's own Z order
    sm = _shuffled(zm, 1)
    ident = np.arange(sm["ne"])
    d_shuf, blk_shuf = _spread(sm, ident)
    d_gen, blk_gen = _spread(zm, ident)
    for mode in ("hilbert", "zcurve"):
        p = spoc_amd.mesh_order(sm, mode)
        assert np.array_equal(np.sort(p), ident)
        d_lib, blk_lib = _spread(sm, p)
        # the shuffle puts almost no neighbour in the same block of 16; the
        # curves put about two thirds there (the generator's lattice-aligned
        # Z order three quarters)
        assert blk_shuf < 0.01 and blk_lib > 0.6 and blk_lib > 0.8 * blk_gen, (mode, blk_lib)
        assert d_shuf > 20 * d_lib and d_lib < 2 * d_gen, (mode, d_shuf, d_lib, d_gen)


def test_order_of_row_numbered_structured_mesh():
    """A structured mesh numbered row by row (src/rmal.f): neighbours one
    row apart are 2 (nx - 1) indices apart in the caller's numbering."""
    x, y, tri = configs.structured_mesh(65, 33, -2e6, 2e6, -1e6, 1e6)
    m = M.build(x, y, tri, 3, 1e-4, 2e-11)
    d_in, blk_in = _spread(m, np.arange(m["ne"]))
    for mode in ("hilbert", "zcurve"):
        d_lib, blk_lib = _spread(m, spoc_amd.mesh_order(m, mode))
        # rows: 0.63 of the neighbour pairs in one block of 16, curves 0.76
        assert d_lib < d_in and blk_lib > blk_in + 0.1, (mode, d_in, d_lib, blk_in, blk_lib)


def test_order_stays_inside_members():
    t = host_tables(3, 12, 4)
    mem = [configs.random_case(t, nx, ny, seed=s)["mesh"] for nx, ny, s in
           ((9, 7, 1), (12, 10, 2), (7, 7, 3))]
    p = spoc_amd.mesh_order([_shuffled(m, 5 + i) for i, m in enumerate(mem)])
    e0 = 0
    for m in mem:
        part = p[e0:e0 + m["ne"]]
        assert np.array_equal(np.sort(part), np.arange(e0, e0 + m["ne"]))
        e0 += m["ne"]
