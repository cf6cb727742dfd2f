"""The driver's files without Fortran I/O (spoc_state_file_*,
spoc_mesh_file_*; host side): byte-identical to a second writer, exact
round trips, list-directed input as Fortran writes it, rejection of
malformed files, and the reference's own fixed-mesh driver (run on the
CPU) reading the files the library wrote and writing a temp.bin the library
reads back, equal to the oracle's RK4 from the same inputs bit for bit."""
import os
import shutil
import tempfile

import numpy as np
import pytest

import fortran_case as FC
import spoc_amd
from conftest import big_stack, host_tables, oracle_for
from spoc_amd import configs


def _rand_state(nm, ne, seed):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal((45, ne))) for _ in range(6)]


def test_state_file_round_trip_and_bytes(tmp_path):
    nm, ne = 15, 37
    arrs = _rand_state(nm, ne, 1)
    p1, p2 = str(tmp_path / "a.bin"), str(tmp_path / "b.bin")
    spoc_amd.write_state_file(p1, nm, 1234.5, *arrs)
    FC.write_init_python(p2, nm, 1234.5, arrs)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    t, *back = spoc_amd.read_state_file(p1, nm, ne)
    assert t == 1234.5
    for a, b in zip(arrs, back):
        assert np.array_equal(a[:nm], b[:nm]) and not b[nm:].any()


def test_state_file_rejects_wrong_sizes(tmp_path):
    nm, ne = 10, 20
    p = str(tmp_path / "s.bin")
    spoc_amd.write_state_file(p, nm, 0.0, *_rand_state(nm, ne, 2))
    with pytest.raises(spoc_amd.SpocError):
        spoc_amd.read_state_file(p, nm, ne + 1)
    with open(p, "r+b") as f:  # break the trailing marker of the 2nd record
        f.seek(4 + 8 + 4 + 4 + nm * ne * 8)
        f.write(b"\x00\x00\x00\x00")
    with pytest.raises(spoc_amd.SpocError):
        spoc_amd.read_state_file(p, nm, ne)
    with pytest.raises(spoc_amd.SpocError):
        spoc_amd.read_state_file(str(tmp_path / "missing.bin"), nm, ne)


def test_mesh_file_round_trip_exact(tmp_path):
    x, y, tri = configs.jittered_mesh(12, 9, 1.0e6, 7.0e5, 0.3, 3)
    p = str(tmp_path / "cm.dat")
    spoc_amd.write_mesh_file(p, x, y, tri, 2.0, 0.5)
    xs, ys, ts = spoc_amd.read_mesh_file(p, 2.0, 0.5)
    assert np.array_equal(ts, tri)
    # x / scx * scx is not always x; the reference read path multiplies the
    # file's numbers, which are x / scx rounded to 17 digits (exact doubles)
    assert np.array_equal(xs, (x / 2.0) * 2.0) and np.array_equal(ys, (y / 0.5) * 0.5)


def test_mesh_file_reads_fortran_list_directed(tmp_path):
    p = tmp_path / "cm.dat"
    p.write_text("           4           2\n"
                 "           1   0.0000000000000000D+00 ,  0.0000000000000000E+00\n"
                 "  2 1.5D3 0.\n 3, 0.0 , 2.5e+03\n4 1500.0 2500.0\n"
                 "           1           1           2           3\n"
                 "           2           2           4           3\n")
    x, y, tri = spoc_amd.read_mesh_file(str(p), 1.0, 2.0)
    assert list(x) == [0.0, 1500.0, 0.0, 1500.0]
    assert list(y) == [0.0, 0.0, 5000.0, 5000.0]
    assert tri.T.tolist() == [[1, 2, 3], [2, 4, 3]]


def test_reference_driver_reads_and_writes_our_files():
    exe = FC.driver("fs", "ref")
    if not os.path.exists(exe):
        pytest.skip("reference drivers not built (make -C oracle drivers)")
    nc, ng, ngf = 3, 12, 4
    t = host_tables(nc, ng, ngf)
    c = configs.random_case(t, 7, 6, seed=77, scale=0.01)
    m, p = c["mesh"], c["phys"]
    d = tempfile.mkdtemp(prefix="files_")
    try:
        FC.write_case(d, nc=nc, ng=ng, ngf=ngf, x=m["x"], y=m["y"], tri=m["in"],
                      state=c["state"], hb=p["hb"], w0=p["w0"], z0=p["z0"], dt=10.0, nite=6,
                      per2=3, mu0=p["mu0"], mb0=p["mb0"], f0=1e-4, beta=2e-11, g0=p["g0"])
        r = FC.run(exe, d, 120)
        assert r.returncode == 0, r.stderr[-2000:]
        nm, ne = t["nm"], m["ne"]
        time, u, v, h, w, z, hb = spoc_amd.read_state_file(os.path.join(d, "temp.bin"), nm, ne)
        assert time == 60.0
        O = oracle_for(nc, ng, ngf)
        q = dict(p)
        q["fc"] = m["fc"]
        ref = big_stack(O.rk4, "fs", m, q, *c["state"], 10.0, 6)
        for a, b in zip((u, v, h), ref):
            assert np.array_equal(a[:nm], b[:nm])
        for a, b in zip((w, z, hb), (p["w0"], p["z0"], p["hb"])):
            assert np.array_equal(a[:nm], b[:nm])
    finally:
        shutil.rmtree(d, ignore_errors=True)
