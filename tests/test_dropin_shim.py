"""The Fortran cal_var shim's re-upload logic, on the CPU: the shim object
(spoc_amd/fortran/spoc_calvar_fs.o) is linked against a stand-in of the
C-ABI (tests/dropin/abi_stub.c) that records which operands reach
spoc_mesh_set and spoc_fields_set.  The adaptive driver may remesh to a mesh
with the same element and face counts (src/shwwnl_adap.f:420-432) and
re-project every field; the shim must notice each change, including those a
checksum of the connectivity cannot see."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, host_tables
from spoc_amd import configs
from spoc_amd import mesh as M

OBJ = os.path.join(ROOT, "spoc_amd", "fortran", "spoc_calvar_fs.o")
STUB = os.path.join(ROOT, "tests", "dropin", "abi_stub.c")


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    fc = shutil.which("flang") or shutil.which("amdflang")
    if not fc or not os.path.exists(OBJ) or not shutil.which("gcc"):
        pytest.skip("flang or the shim object is missing (python -m spoc_amd.build)")
    d = tmp_path_factory.mktemp("shim")
    so = str(d / "libshim_stub.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-c", STUB, "-o", str(d / "stub.o")], check=True)
    subprocess.run([fc, "-shared", "-fPIC", OBJ, str(d / "stub.o"), "-o", so], check=True)
    lib = C.CDLL(so)
    lib.stub_itm.restype = C.POINTER(C.c_int)
    lib.stub_vect.restype = C.POINTER(C.c_double)
    lib.stub_w0.restype = C.POINTER(C.c_double)
    t = host_tables(3, 12, 4)
    const = (C.c_int * 4).in_dll(lib, "const_")
    const[0], const[1], const[2], const[3] = 3, t["nm"], 12, 4
    return lib


def _counts(lib):
    c = (C.c_int * 4)()
    lib.stub_counts(c)
    return dict(ctx=c[0], mesh=c[1], fields=c[2], rhs=c[3])


def _call(lib, m, phys, state):
    """cal_var(nf, ne, u0, v0, h0, u1, v1, h1, tseg, pseg, ssit, itm, its, norm,
    nbdl, pbdl, nint, pint, bdl, cip, vect, normt, mu0, mb0, fc, beta, w0, z0,
    g0, hb, time), every argument by reference."""
    ne, nf = m["ne"], m["nf"]
    d = lambda a: np.asfortranarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
    keep = []

    def arr(a, dt):
        a = np.asfortranarray(a, dt)
        keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    out = [np.zeros((45, ne), order="F") for _ in range(3)]
    i = lambda v: C.byref(C.c_int(v))
    r = lambda v: C.byref(C.c_double(v))
    dummy = np.zeros(8, np.int32)
    lib.cal_var_(i(nf), i(ne), arr(state[0], np.float64), arr(state[1], np.float64),
                 arr(state[2], np.float64), arr(out[0], np.float64), arr(out[1], np.float64),
                 arr(out[2], np.float64), arr(m["tseg"], np.int32), arr(m["pseg"], np.int32),
                 arr(m["ssit"], np.int32), arr(m["itm"], np.int32), arr(m["its"], np.int32),
                 arr(m["norm"], np.float64), i(0), arr(dummy, np.int32), i(0),
                 arr(dummy, np.int32), arr(m["bdl"], np.int32), arr(dummy, np.int32),
                 arr(m["vect"], np.float64), arr(m["normt"], np.float64), r(phys["mu0"]),
                 r(phys["mb0"]), arr(phys["fc"], np.float64), r(0.0), arr(phys["w0"], np.float64),
                 arr(phys["z0"], np.float64), r(phys["g0"]), arr(phys["hb"], np.float64), r(0.0))


def _old_checksum(m):
    """The linear connectivity sum the shim used before: it cannot see
    geometry, and it is blind to many renumberings."""
    itm, its, tseg = (np.asarray(m[k], np.int64) for k in ("itm", "its", "tseg"))
    e = np.arange(1, m["ne"] + 1)
    return int((itm[0] * 3 + itm[1] * 5 + itm[2] * 7 + e * (its[0] + 2)).sum() +
               (tseg[0] * 11 + tseg[1] * 13).sum())


def test_shim_reuploads_on_every_change(shim):
    t = host_tables(3, 12, 4)
    c = configs.random_case(t, 6, 5, seed=11, scale=0.01)
    m, phys, st = c["mesh"], c["phys"], c["state"]
    _call(shim, m, phys, st)
    assert _counts(shim) == dict(ctx=1, mesh=1, fields=1, rhs=1)
    _call(shim, m, phys, st)  # nothing changed: no upload
    assert _counts(shim) == dict(ctx=1, mesh=1, fields=1, rhs=2)

    # same connectivity, new node positions: the old checksum is unchanged
    x2, y2 = m["x"] * 1.01 + 3.0e3, m["y"] * 0.99
    m2 = M.build(x2, y2, m["in"], 3, 1e-4, 2e-11)
    assert _old_checksum(m2) == _old_checksum(m)
    assert not np.array_equal(m2["vect"], m["vect"])
    _call(shim, m2, phys, st)
    n = _counts(shim)
    assert (n["mesh"], n["fields"]) == (2, 2)  # a new mesh takes the fields again
    got = np.ctypeslib.as_array(shim.stub_vect(), (5 * m["ne"],))
    assert np.array_equal(got, np.asarray(m2["vect"]).ravel(order="F"))

    # one wind coefficient of one element changes (hb(1,1) does not)
    p2 = dict(phys)
    p2["w0"] = np.array(phys["w0"], order="F")
    p2["w0"][2, 7] += 1e-6
    _call(shim, m2, p2, st)
    n = _counts(shim)
    assert (n["mesh"], n["fields"]) == (2, 3)
    got = np.ctypeslib.as_array(shim.stub_w0(), (45 * m["ne"],)).reshape(45, -1, order="F")
    assert got[2, 7] == p2["w0"][2, 7]

    # the elements renumbered (same ne and nf): connectivity differs
    perm = np.random.default_rng(3).permutation(m["ne"])
    m3 = M.build(x2, y2, np.asfortranarray(m2["in"][:, perm]), 3, 1e-4, 2e-11)
    assert (m3["ne"], m3["nf"]) == (m2["ne"], m2["nf"])
    _call(shim, m3, p2, st)
    n = _counts(shim)
    assert (n["mesh"], n["fields"], n["rhs"]) == (3, 4, 5)
    got = np.ctypeslib.as_array(shim.stub_itm(), (3 * m["ne"],))
    assert np.array_equal(got, np.asarray(m3["itm"]).ravel(order="F"))
