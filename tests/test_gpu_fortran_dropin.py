"""The drop-in, end to end: the reference's own fixed-mesh driver (program
shwater, src/shwwnl.f, built from its sources by `make -C oracle drivers`)
linked with its own calvar*.o, against the same driver objects linked with
the ISO_C_BINDING cal_var of spoc_amd/fortran and libspoc_amd.so.  Same run
directory inputs; the FP64 path must reproduce every output file (final
state temp.bin, energy series oc.energ, graphics *.gra) byte for byte.  The
mixed-precision path (SPOC_PRECISION=fp32) is held to a relative mass-norm
bound instead."""
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

import fortran_case as FC
from conftest import host_tables
from spoc_amd import configs

pytestmark = pytest.mark.gpu

CASES = {
    # variant: (p, ng, ngf, steps, dt)
    "lin": (3, 12, 4, 10, None),
    "fs": (6, 33, 7, 10, 10.0),
    "ns": (4, 25, 5, 10, 10.0),
    "inv": (5, 25, 6, 10, 10.0),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_dirs(variant):
    nc, ng, ngf, nite, dt = CASES[variant]
    t = host_tables(nc, ng, ngf)
    if variant == "lin":
        c = configs.config1(t)
        dt = c["dt"]
    else:
        c = configs.random_case(t, 9, 9, seed=300 + nc, scale=0.01)
        if variant == "inv":
            c["phys"]["mu0"] = 0.0
    m, p = c["mesh"], c["phys"]
    root = tempfile.mkdtemp(prefix=f"dropin_{variant}_")
    dirs = {}
    for kind in ("ref", "spoc", "spoc32"):
        d = os.path.join(root, kind)
        FC.write_case(d, nc=nc, ng=ng, ngf=ngf, x=m["x"], y=m["y"], tri=m["in"],
                      state=c["state"], hb=p["hb"], w0=p["w0"], z0=p["z0"], dt=dt, nite=nite,
                      per2=5, mu0=p["mu0"], mb0=p["mb0"], f0=1e-4, beta=2e-11, g0=p["g0"])
        dirs[kind] = d
    return root, dirs, t, m


def outputs(d):
    return sorted(f for f in os.listdir(d) if f.endswith((".bin", ".gra", ".energ"))
                  and f != "init.bin")


@pytest.mark.parametrize("variant", ["lin", "fs", "ns", "inv"])
def test_reference_driver_with_dropin_cal_var(variant):
    for kind in ("ref", "spoc"):
        if not os.path.exists(FC.driver(variant, kind)):
            pytest.skip("drivers not built (make -C oracle drivers)")
    root, dirs, t, m = make_dirs(variant)
    try:
        say("drop-in driver", variant, "ne", m["ne"])
        r = FC.run(FC.driver(variant, "ref"), dirs["ref"], 120)
        assert r.returncode == 0, r.stderr[-2000:]
        r = FC.run(FC.driver(variant, "spoc"), dirs["spoc"], 120, env={"SPOC_PRECISION": "fp64"})
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        files = outputs(dirs["ref"])
        assert "temp.bin" in files and "oc.energ" in files, files
        assert outputs(dirs["spoc"]) == files
        for f in files:
            a = open(os.path.join(dirs["ref"], f), "rb").read()
            b = open(os.path.join(dirs["spoc"], f), "rb").read()
            assert a == b, f"{variant}: {f} differs"
        # mixed precision: same driver, SPOC_PRECISION=fp32
        r = FC.run(FC.driver(variant, "spoc"), dirs["spoc32"], 120, env={"SPOC_PRECISION": "fp32"})
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        nm = t["nm"]
        ref = FC.read_temp(os.path.join(dirs["ref"], "temp.bin"), nm, m["ne"])
        got = FC.read_temp(os.path.join(dirs["spoc32"], "temp.bin"), nm, m["ne"])
        M = t["ammo"][:nm, :nm]
        det = 1.0 / m["vect"][4]
        for f in range(1, 4):
            d = got[f] - ref[f]
            num = np.sum(det * np.einsum("ie,ij,je->e", d, M, d))
            den = np.sum(det * np.einsum("ie,ij,je->e", ref[f], M, ref[f]))
            err = float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))
            say("  fp32 drop-in field", f, "rel err", err)
            assert err < 1e-7, (variant, f, err)
    finally:
        shutil.rmtree(root, ignore_errors=True)
