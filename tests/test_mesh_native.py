"""The native mesh builder (spoc_mesh_build, spoc_mesh_changemaille; host C++)
against the test-side Python restatement (tests/mesh_ref.py) on meshes
larger and more varied than the reference goldens hold (the goldens pin
both through tests/test_oracle_golden.py): every integer array equal and
every geometry word bit-identical; and the rejections of check_mesh."""
import numpy as np
import pytest

import mesh_ref as R
from spoc_amd import configs, mesh


def _raw_tris(nx, ny, seed):
    """Delaunay-free jittered triangulation before the CM step: rmal's
    triangles with node positions jittered (corner triangles included)."""
    x, y, tri = mesh.rmal(nx, ny, 0.0, 1.0e6, 0.0, 7.0e5)
    rng = np.random.default_rng(seed)
    inner = (x > 0) & (x < 1.0e6) & (y > 0) & (y < 7.0e5)
    x = x + inner * rng.uniform(-0.2, 0.2, len(x)) * 1.0e6 / (nx - 1)
    y = y + inner * rng.uniform(-0.2, 0.2, len(y)) * 7.0e5 / (ny - 1)
    return x, y, tri


@pytest.mark.parametrize("nx,ny,seed", [(5, 4, 1), (17, 9, 2), (40, 31, 3), (64, 50, 4)])
def test_changemaille_matches_restatement(nx, ny, seed):
    x, y, tri = _raw_tris(nx, ny, seed)
    assert np.array_equal(mesh.changemaille(len(x), tri), R.changemaille(len(x), tri))


@pytest.mark.parametrize("nx,ny,seed,nc", [(5, 4, 1, 2), (17, 9, 2, 3), (40, 31, 3, 6),
                                            (64, 50, 4, 7)])
def test_build_matches_restatement(nx, ny, seed, nc):
    x, y, tri = _raw_tris(nx, ny, seed)
    tri = R.changemaille(len(x), tri)
    got = mesh.build(x, y, tri, nc, 1e-4, 2e-11)
    ref = R.build(x, y, tri, nc, 1e-4, 2e-11)
    assert got["nf"] == ref["nf"]
    for k in ("itm", "its", "tseg", "pseg", "ssit", "norm", "vect", "normt", "bdl", "fc", "deta"):
        assert np.array_equal(got[k], ref[k]), k
    # every face has two distinct elements or is a wall face on edge 1
    t = got["tseg"]
    assert (t[0] > 0).all() and (t[0] != t[1]).all()
    walls = t[1] == 0
    assert (got["ssit"][0, walls] == 1).all()


def test_jittered_config_mesh_matches_restatement():
    x, y, tri = configs.jittered_mesh(30, 20, 2.0e6, 1.0e6, 0.3, 11)
    got = mesh.build(x, y, tri, 4, 1e-4, 2e-11)
    ref = R.build(x, y, tri, 4, 1e-4, 2e-11)
    for k in ("itm", "its", "tseg", "pseg", "ssit", "norm", "vect", "normt", "bdl", "fc"):
        assert np.array_equal(got[k], ref[k]), k


def test_two_wall_faces_rejected_unless_unchecked():
    x, y, tri = mesh.rmal(4, 4, 0.0, 1.0, 0.0, 1.0)  # corner triangles: two wall faces
    with pytest.raises(ValueError):
        mesh.build(x, y, tri, 3)
    m = mesh.build(x, y, tri, 3, check=False)
    ref = R.build(x, y, tri, 3, check=False)
    for k in ("itm", "tseg", "ssit", "its"):
        assert np.array_equal(m[k], ref[k]), k


def test_bad_node_index_rejected():
    x, y, tri = mesh.rmal(4, 4, 0.0, 1.0, 0.0, 1.0)
    tri = np.array(tri, order="F")
    tri[1, 3] = 99
    with pytest.raises(ValueError):
        mesh.build(x, y, tri, 3, check=False)
    with pytest.raises(ValueError):
        mesh.changemaille(len(x), tri)
