"""The device path at the shapes and physics of the benchmark
configurations, on reduced meshes:

- configurations 2 (fs, p=6), 4 (inv, p=5) and 5 (fs, p=4, 25-point rule):
  10 RK4 steps of the FP64 path equal the reference Fortran's own run
  (golden vectors written by it) bit for bit, energies included; the mixed
  path is within 1e-8 (relative L2, mass norm);
- configuration 3 (ns, p=7, 42-point rule, beyond the reference's 33-point
  arrays): 5 RK4 steps of the FP64 path equal the oracle bit for bit; the
  mixed path is held to the metric's 1e-6 (measured: u 2.5e-9, v 2.9e-9,
  eta 3.6e-8: the seamounts' depth variation and the eddies' large velocity
  make the FP32 part of the mass flux a larger share of eta's tendency than
  in the other configurations);
- configuration 5 as an ensemble: six members of 300-1,500 triangles in one
  union mesh; FP64 equals the oracle member by member, the mixed path is
  within 1e-8 of it, and every member equals its solo run on both paths."""
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import gold, host_tables, oracle_for
from spoc_amd import configs

pytestmark = pytest.mark.gpu
_ctx = {}

CFG = {
    "c2": ((6, 33, 7), lambda tb: configs.config2(tb, nx=25, ny=25), "fs"),
    "c4": ((5, 25, 6), lambda tb: configs.config4(tb, nx=25, ny=13), "inv"),
    "c5": ((4, 25, 5), lambda tb: configs.config5(tb, nmem=1, ntri=(1100, 1100))[0], "fs"),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ctx_for(shape):
    if shape not in _ctx:
        t = host_tables(*shape)
        _ctx[shape] = (t, spoc_amd.Context(t))
    return _ctx[shape]


def set_fields(ctx, p):
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])


def mrel(t, vect, a, b):
    nm = t["nm"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / vect[4]
    d = a[:nm] - b[:nm]
    num = np.sum(det * np.einsum("ie,ij,je->e", d, M, d))
    den = np.sum(det * np.einsum("ie,ij,je->e", b[:nm], M, b[:nm]))
    return float(np.sqrt(num / den))


@pytest.mark.parametrize("name", sorted(CFG))
def test_configuration_rk4_vs_reference_goldens(name):
    shape, gen, variant = CFG[name]
    t, ctx = ctx_for(shape)
    g = gold("rk4_cfg.npz")
    c = gen(t)
    nm = t["nm"]
    assert c["dt"] == float(g[f"{name}_dt"]) and c["mesh"]["ne"] == int(g[f"{name}_ne"])
    for k, a in zip("uvh", c["state"]):
        assert np.array_equal(a[:nm], g[f"{name}_{k}0"]), k
    say("config", name, variant, shape, "ne", c["mesh"]["ne"], "dt", c["dt"])
    ctx.set_mesh(c["mesh"])
    set_fields(ctx, c["phys"])
    ctx.upload(*c["state"])
    e0 = ctx.energy()
    ctx.rk4(variant, c["dt"], 10, precision="fp64")
    got = ctx.download()
    e1 = ctx.energy()
    for k, a in zip("uvh", got):
        assert np.array_equal(a[:nm], g[f"{name}_{k}"]), k
    assert np.array_equal(np.array([e0, e1]), g[f"{name}_energy"])
    ctx.upload(*c["state"])
    ctx.rk4(variant, c["dt"], 10, precision="fp32")
    got32 = ctx.download()
    errs = [mrel(t, c["mesh"]["vect"], a, g[f"{name}_{k}"]) for k, a in zip("uvh", got32)]
    say("  mixed path rel L2 u v h:", errs)
    assert max(errs) < 1e-8, errs


def test_configuration3_ns_p7_vs_oracle():
    shape = (7, 42, 8)
    t, ctx = ctx_for(shape)
    O = oracle_for(*shape)
    c = configs.config3(t, nx=25, ny=13)
    assert c["variant"] == "ns"
    say("config3 ne", c["mesh"]["ne"], "dt", c["dt"])
    ctx.set_mesh(c["mesh"])
    set_fields(ctx, c["phys"])
    ref = O.rk4("ns", c["mesh"], c["phys"], *c["state"], c["dt"], 5)
    ctx.upload(*c["state"])
    e0 = ctx.energy()
    ctx.rk4("ns", c["dt"], 5, precision="fp64")
    got = ctx.download()
    e1 = ctx.energy()
    nm = t["nm"]
    for k, a, b in zip("uvh", got, ref):
        assert np.array_equal(a[:nm], b[:nm]), k
    assert e0 == O.energy(c["mesh"], c["phys"], *c["state"])
    assert e1 == O.energy(c["mesh"], c["phys"], *ref)
    ctx.upload(*c["state"])
    ctx.rk4("ns", c["dt"], 5, precision="fp32")
    errs = [mrel(t, c["mesh"]["vect"], a, b) for a, b in zip(ctx.download(), ref)]
    say("  mixed path rel L2 u v h:", errs)
    assert max(errs) < 1e-6, errs


def test_configuration5_ensemble_members():
    shape = (4, 25, 5)
    t, ctx = ctx_for(shape)
    O = oracle_for(*shape)
    mem = configs.config5(t, nmem=6, ntri=(300, 1500))
    sizes = [c["mesh"]["ne"] for c in mem]
    dt, n, nm = mem[0]["dt"], 5, t["nm"]
    say("config5 ensemble sizes", sizes, "dt", dt)
    assert min(sizes) >= 250 and max(sizes) <= 1700 and len(set(sizes)) > 1
    ref = [O.rk4("fs", c["mesh"], c["phys"], *c["state"], dt, n) for c in mem]
    solo = {}
    for prec in ("fp64", "fp32"):
        solo[prec] = []
        for c in mem:
            ctx.set_mesh(c["mesh"])
            set_fields(ctx, c["phys"])
            ctx.upload(*c["state"])
            ctx.rk4("fs", dt, n, precision=prec)
            solo[prec].append(ctx.download())
    ctx.set_ensemble_mesh([c["mesh"] for c in mem])
    P = [c["phys"] for c in mem]
    cat = spoc_amd.Context.concat
    ctx.set_fields(cat([p["hb"] for p in P]), P[0]["g0"], cat([p["fc"] for p in P]),
                   cat([p["w0"] for p in P]), cat([p["z0"] for p in P]), P[0]["mu0"],
                   P[0]["mb0"])
    for prec in ("fp64", "fp32"):
        ctx.upload(*(cat([c["state"][f] for c in mem]) for f in range(3)))
        ctx.rk4("fs", dt, n, precision=prec)
        got = ctx.download()
        e0 = 0
        for m, c in enumerate(mem):
            ne = c["mesh"]["ne"]
            for f in range(3):
                a = got[f][:nm, e0:e0 + ne]
                assert np.array_equal(a, solo[prec][m][f][:nm]), (prec, m, f)
                if prec == "fp64":
                    assert np.array_equal(a, ref[m][f][:nm]), (m, f)
                else:
                    err = mrel(t, c["mesh"]["vect"], a, ref[m][f])
                    assert err < 1e-8, (m, f, err)
            e0 += ne
