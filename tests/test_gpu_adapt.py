"""The device refinement indicator (spoc_checkconv, spoc_cal_conv) against
the reference's own outputs (golden vectors) and the oracle, bit for bit:
indicators, per-element flags and the three decisions; on random states,
on the state left by device RK4 steps of both arithmetic paths, and per
member of an ensemble (each member equals its own single-mesh run)."""
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import gold, host_tables, oracle_for
from spoc_amd import configs

pytestmark = pytest.mark.gpu
_ctx = {}


def ctx_for(nc, ng, ngf):
    if (nc, ng, ngf) not in _ctx:
        t = host_tables(nc, ng, ngf)
        _ctx[(nc, ng, ngf)] = (t, spoc_amd.Context(t))
    return _ctx[(nc, ng, ngf)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [(3, 12, 4), (6, 33, 7)]


@pytest.mark.parametrize("nc,ng,ngf", CASES)
@pytest.mark.parametrize("which", ["mid", "rollback", "flat"])
def test_cal_conv_matches_reference_goldens(nc, ng, ngf, which):
    g = gold(f"calconv_p{nc}.npz")
    t, ctx = ctx_for(nc, ng, ngf)
    c = configs.random_case(t, int(g["nx"]), int(g["ny"]), seed=int(g["seed"]), scale=0.01)
    nm = t["nm"]
    st = [np.asfortranarray(a) for a in c["state"]]
    for k, a in zip(("u0", "v0", "h0"), st):
        assert np.array_equal(a[:nm], g[k]), k
    if which == "flat":
        st[2] = np.zeros_like(st[2])
    say("cal_conv goldens", nc, which, c["mesh"]["ne"])
    ctx.set_mesh(c["mesh"])
    ctx.upload(*st)
    f1, f2, f3 = g[f"{which}_fac"]
    r = ctx.cal_conv(f1, f2, f3)
    assert np.array_equal(r["fconv"], g[f"{which}_fconv"])
    assert np.array_equal(r["fderaff"], g[f"{which}_fderaff"])
    assert [r["conv"], r["deraff"], r["timedir"]] == list(g[f"{which}_flags"])
    if which == "flat":
        assert np.array_equal(r["amax"][:2], g["amax"][:2])
        assert not r["amax"][2].any()
    else:
        assert np.array_equal(r["amax"], g["amax"])
        for f in range(3):
            assert np.array_equal(ctx.checkconv(f), g["amax"][f])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cal_conv_after_device_steps_matches_oracle(prec):
    nc, ng, ngf = 6, 33, 7
    t, ctx = ctx_for(nc, ng, ngf)
    O = oracle_for(nc, ng, ngf)
    c = configs.random_case(t, 15, 13, seed=202, scale=0.01)
    m = c["mesh"]
    say("cal_conv after steps", prec, m["ne"])
    ctx.set_mesh(m)
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
    ctx.upload(*c["state"])
    ctx.rk4("fs", 2.0, 5, precision=prec)
    u, v, h = ctx.download()
    ref = O.cal_conv(m, u, v, h, 1.5, 0.8, 40.0)
    got = ctx.cal_conv(1.5, 0.8, 40.0)
    assert np.array_equal(got["amax"], ref["amax"])
    for k in ("fconv", "fderaff"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("conv", "deraff", "timedir", "supmax"):
        assert got[k] == ref[k], k


def test_ensemble_members_equal_single_runs():
    nc, ng, ngf = 4, 25, 5
    t, ctx = ctx_for(nc, ng, ngf)
    mem = [configs.random_case(t, nx, ny, seed=60 + i, scale=0.01)
           for i, (nx, ny) in enumerate([(9, 7), (11, 9), (8, 8)])]
    # member 1's state is scaled so that the members' normalisations differ
    states = [[np.asfortranarray(a * (8.0 if i == 1 else 1.0)) for a in c["state"]]
              for i, c in enumerate(mem)]
    fac = (1.2, 0.9, 30.0)
    single = []
    for c, st in zip(mem, states):
        ctx.set_mesh(c["mesh"])
        ctx.upload(*st)
        single.append(ctx.cal_conv(*fac))
    ctx.set_ensemble_mesh([c["mesh"] for c in mem])
    cat = spoc_amd.Context.concat
    ctx.upload(*(cat([st[f] for st in states]) for f in range(3)))
    say("ensemble cal_conv", [c["mesh"]["ne"] for c in mem])
    r = ctx.cal_conv(*fac)
    e0 = 0
    for i, c in enumerate(mem):
        ne = c["mesh"]["ne"]
        s = single[i]
        assert np.array_equal(r["amax"][:, e0:e0 + ne], s["amax"]), i
        for k in ("fconv", "fderaff"):
            assert np.array_equal(r[k][e0:e0 + ne], s[k]), (i, k)
        for k in ("conv", "deraff", "timedir", "supmax"):
            assert r[k][i] == s[k], (i, k)
        e0 += ne
