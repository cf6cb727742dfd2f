"""spoc_energy_noalloc, the entry point that allocates nothing (its partial
integrals live in a fixed device array), against spoc_energy and
spoc_ensemble_energy, the entry points it varies: bit-identical integrals on
two calls in a row, the second on a smaller mesh, and for an ensemble."""
import ctypes as C
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import host_tables
from spoc_amd import configs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DevBuf:
    """n doubles of device memory through the HIP runtime (the test
    allocates; the entry point under test does not).  The HIP functions are
    looked up through libspoc_amd.so's own dependencies: a process that has
    imported torch can hold a second HIP runtime under the same name."""

    def __init__(self, n):
        self.hip = spoc_amd.load_library()
        for f in ("hipMalloc", "hipMemset", "hipMemcpy", "hipFree", "hipDeviceSynchronize"):
            getattr(self.hip, f).restype = C.c_int
        self.n = n
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0

    def fill_nan(self):  # all-ones bytes: a NaN in every double
        assert self.hip.hipMemset(self.p, 0xFF, C.c_size_t(8 * self.n)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def get(self):
        h = np.zeros(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n),
                                  2) == 0  # device to host
        return h

    def free(self):
        self.hip.hipFree(self.p)


def _setup(ctx, c):
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])


def test_energy_noalloc_equals_energy():
    t = host_tables(6, 33, 7)
    ctx = spoc_amd.Context(t)
    out = DevBuf(24)
    for nx, ny, seed in ((21, 13, 5), (9, 7, 6)):  # the second mesh is smaller
        c = configs.random_case(t, nx, ny, seed=seed, scale=0.01)
        say("energy_noalloc ne", c["mesh"]["ne"])
        ctx.set_mesh(c["mesh"])
        _setup(ctx, c)
        ctx.upload(*c["state"])
        ctx.rk4("fs", 1.0, 3, precision="fp64")
        ref = np.array(ctx.energy())
        out.fill_nan()
        ctx.energy_noalloc(out.p.value)
        ctx.sync()
        h = out.get()
        assert np.array_equal(h[:3], ref), (h[:3], ref)
        assert np.isnan(h[3:]).all()  # nothing written past the output
    # an ensemble: one (ke, pe, mass) per member
    mem = [configs.random_case(t, nx, ny, seed=s, scale=0.01)
           for nx, ny, s in ((8, 7, 1), (11, 9, 2), (7, 7, 3))]
    say("ensemble", [m["mesh"]["ne"] for m in mem])
    ctx.set_ensemble_mesh([m["mesh"] for m in mem])
    P = [m["phys"] for m in mem]
    cat = spoc_amd.Context.concat
    ctx.set_fields(cat([p["hb"] for p in P]), P[0]["g0"], cat([p["fc"] for p in P]),
                   cat([p["w0"] for p in P]), cat([p["z0"] for p in P]), P[0]["mu0"],
                   P[0]["mb0"])
    ctx.upload(*(cat([m["state"][f] for m in mem]) for f in range(3)))
    ctx.rk4("fs", 1.0, 2, precision="fp32")
    ke, pe, mass = ctx.ensemble_energy()
    out.fill_nan()
    ctx.energy_noalloc(out.p.value)
    ctx.sync()
    got = out.get()[:9].reshape(3, 3)
    assert np.array_equal(got[:, 0], ke) and np.array_equal(got[:, 1], pe)
    assert np.array_equal(got[:, 2], mass)
    ctx.close()
    out.free()
