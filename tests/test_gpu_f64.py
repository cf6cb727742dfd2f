"""Plain FP64 GPU path against the oracle: bit-exact (the kernels reproduce
the reference's operation order), through the C-ABI."""
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import big_stack, host_tables, oracle_for
from spoc_amd import configs

pytestmark = pytest.mark.gpu
SHAPES = [(2, 6, 3), (3, 12, 4), (4, 25, 5), (5, 25, 6), (6, 33, 7), (7, 42, 8)]
_ctx_cache = {}


def ctx_for(nc, ng, ngf):
    key = (nc, ng, ngf)
    if key not in _ctx_cache:
        t = host_tables(nc, ng, ngf)
        _ctx_cache[key] = (t, spoc_amd.Context(t))
    return _ctx_cache[key]


def setup_case(ctx, c):
    ctx.set_mesh(c["mesh"])
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nc,ng,ngf", SHAPES)
@pytest.mark.parametrize("variant", ["lin", "fs", "ns", "inv"])
def test_cal_var_fp64_bit_exact(nc, ng, ngf, variant):
    t, ctx = ctx_for(nc, ng, ngf)
    O = oracle_for(nc, ng, ngf)
    nx = 17 if nc <= 3 else 12
    c = configs.random_case(t, nx, nx - 2, seed=100 + nc)
    say("case", variant, nc, ng, ngf, "ne", c["mesh"]["ne"])
    setup_case(ctx, c)
    got = ctx.cal_var(variant, *c["state"], precision="fp64")
    ref = O.cal_var(variant, c["mesh"], c["phys"], *c["state"])
    nm = t["nm"]
    for k, a, b in zip("uvh", got, ref):
        assert np.array_equal(a[:nm], b[:nm]), (k, float(np.max(np.abs(a[:nm] - b[:nm]))))
        assert np.all(a[nm:] == 0.0)


@pytest.mark.parametrize("variant", ["lin", "fs", "ns", "inv"])
def test_rk4_fp64_bit_exact_config1(variant):
    t, ctx = ctx_for(3, 12, 4)
    O = oracle_for(3, 12, 4)
    c = configs.config1(t)
    c["phys"]["mu0"] = 100.0
    say("rk4 config1", variant)
    setup_case(ctx, c)
    ctx.upload(*c["state"])
    e0 = ctx.energy()
    ctx.rk4(variant, c["dt"], 20)
    got = ctx.download()
    e1 = ctx.energy()
    ref = O.rk4(variant, c["mesh"], c["phys"], *c["state"], c["dt"], 20)
    for k, a, b in zip("uvh", got, ref):
        assert np.array_equal(a[:10], b[:10]), k
    assert e0 == O.energy(c["mesh"], c["phys"], *c["state"])
    assert e1 == O.energy(c["mesh"], c["phys"], *ref)


def test_rk4_fp64_p6_multi_tile_with_walls():
    t, ctx = ctx_for(6, 33, 7)
    O = oracle_for(6, 33, 7)
    c = configs.random_case(t, 21, 13, seed=7, scale=0.01)
    say("rk4 p6 fs ne", c["mesh"]["ne"])
    setup_case(ctx, c)
    ctx.upload(*c["state"])
    ctx.rk4("fs", 1.0, 3)
    got = ctx.download()
    ref = O.rk4("fs", c["mesh"], c["phys"], *c["state"], 1.0, 3)
    for k, a, b in zip("uvh", got, ref):
        assert np.array_equal(a[:28], b[:28]), k
