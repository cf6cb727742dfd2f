"""C-ABI library checks that need no GPU: exports, host table builder."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import spoc_amd
from conftest import ROOT, gold, oracle_for, quad_rule


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "spoc_abi.h")).read()
    names = sorted(set(re.findall(r"^\w[\w\s\*]*?\b(spoc_\w+)\(", hdr, re.M)))
    assert len(names) >= 15
    lib = spoc_amd.load_library()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("nc,ng,ngf", [(3, 12, 4), (4, 25, 5), (5, 25, 6), (6, 33, 7)])
def test_host_tables_match_reference_bit_for_bit(nc, ng, ngf):
    g = gold(f"tables_p{nc}_{ng}.npz")
    t = spoc_amd.build_tables(nc, *quad_rule(ng, ngf))
    nm, nl = t["nm"], t["nm"] - nc - 1
    for k in ("amm", "agx", "agy", "ammo"):
        assert np.array_equal(t[k][:nm, :nm], g[k][:nm, :nm]), k
    assert np.array_equal(t["amml"][:nl, :nl], g["amml"][:nl, :nl])
    assert np.array_equal(t["pi2"][:ng, :nm], g["pi2"][:ng, :nm])
    assert np.array_equal(t["pico"][:ngf, :nm], g["pico"][:ngf, :nm])


def test_host_tables_42_point_rule_match_oracle():
    t = spoc_amd.build_tables(7, *quad_rule(42, 8))
    o = oracle_for(7, 42, 8).tables()
    nm = 36
    for k in ("amm", "agx", "agy", "ammo"):
        assert np.array_equal(t[k][:nm, :nm], o[k][:nm, :nm]), k


def test_singular_rule_is_an_error_not_nan():
    # p=4 has 15 modes; the 12-point rule cannot make the mass matrix SPD
    # (the reference prints "pb avec matrice" and runs on to NaN).
    with pytest.raises(spoc_amd.SpocError, match="not positive definite"):
        spoc_amd.build_tables(4, *quad_rule(12, 5))


def test_unsupported_shape_rejected_before_device_use():
    t = spoc_amd.build_tables(3, *quad_rule(25, 4))
    with pytest.raises(spoc_amd.SpocError, match="unsupported"):
        spoc_amd.Context(t)


def test_adaptive_driver_links_the_dropin_cal_conv():
    """The reference's adaptive driver (src/shwwnl_adap.f) built with the
    drop-in objects (oracle/Makefile, target adaptive): its cal_conv is the
    ISO_C_BINDING shim, which calls spoc_cal_conv from libspoc_amd.so, and
    adapt.f's own cal_conv is present only under its renamed symbol."""
    import shutil
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "shwwnl_adap_fs_spoc")
    if not os.path.exists(exe) or not shutil.which("nm"):
        pytest.skip("adaptive driver not built (make -C oracle adaptive)")
    syms = subprocess.run(["nm", exe], capture_output=True, text=True, check=True).stdout
    defined = {ln.split()[-1] for ln in syms.splitlines() if " T " in ln}
    undefined = {ln.split()[-1] for ln in syms.splitlines() if ln.strip().startswith("U ")}
    assert {"cal_conv_", "cal_conv_reference_", "cal_var_", "checkconv_"} <= defined
    assert "spoc_cal_conv" in undefined and "spoc_rhs" in undefined


_ENOMEM_SCRIPT = r'''
import ctypes as C, resource, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import spoc_amd
from spoc_amd import configs, mesh
lib = spoc_amd.load_library()
x, y, tri = configs.structured_mesh(401, 401, 0.0, 1e6, 0.0, 1e6)   # 320,000 triangles
tri = np.asfortranarray(tri, dtype=np.int32)
nn, ne = len(x), tri.shape[1]
i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
f64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
z = lambda *s, t=np.float64: np.zeros(s, t, order="F")
out = [z(3, ne, t=np.int32), z(3, ne, t=np.int32), z(2, 3 * ne, t=np.int32),
       z(2, 3 * ne, t=np.int32), z(2, 3 * ne, t=np.int32), z(2, 3 * ne), z(5, ne), z(2, ne),
       z(ne, t=np.int32), z(mesh.LDM, ne), z(ne)]
nf = C.c_int()
args = [nn, ne, i32(tri), f64(np.ascontiguousarray(x)), f64(np.ascontiguousarray(y)), 3,
        C.c_double(1e-4), C.c_double(0.0), mesh.LDM, 1, C.byref(nf)]
args += [i32(a) if a.dtype == np.int32 else f64(a) for a in out]
vm = int([l for l in open("/proc/self/status") if l.startswith("VmSize")][0].split()[1]) * 1024
resource.setrlimit(resource.RLIMIT_AS, (vm + (8 << 20), resource.RLIM_INFINITY))
rc = lib.spoc_mesh_build(*args)
resource.setrlimit(resource.RLIMIT_AS, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
print(rc, lib.spoc_last_error().decode())
'''


def test_host_allocation_failure_is_an_error_code():
    """An entry point whose host allocations fail returns SPOC_ENOMEM with a
    message and the caller's process goes on (no C++ exception crosses the
    C-ABI).  The mesh builder's maps for 320,000 triangles are asked for
    under an address-space limit 8 MiB above what the process holds."""
    import subprocess
    import sys
    if os.environ.get("SPOC_AMD_LIB"):
        pytest.skip("the sanitizer build's allocator reports running out and stops instead")
    r = subprocess.run([sys.executable, "-c", _ENOMEM_SCRIPT, ROOT], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    rc, msg = r.stdout.strip().split(" ", 1)
    assert int(rc) == spoc_amd.ENOMEM and "memory" in msg, r.stdout
