"""Persistent linear RK4 on the mixed path (spoc_set_persistent; no reference
counterpart): every step of a spoc_rk4_steps call in one launch whose waves
keep their element blocks and, between stages, wait for their neighbouring
blocks.  Held to bit-for-bit equality with the launch-per-stage path it
replaces (the reference's loop, src/shwwnl.f:233-284, one right-hand side
per stage), against that path in every compiled form and workgroup size,
across calls of different lengths and time steps, up to the largest mesh it
takes at p = 3 (1,536 elements; 1,024 at p = 4, 768 at p = 5, none above,
where the per-stage launches are faster) and just past it (where it hands
over to the per-stage path)."""
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import host_tables
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


def set_fields(ctx, p):
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])


def calls(ctx, c, persist):
    ctx.set_persistent(persist)
    ctx.upload(*c["state"])
    ctx.rk4("lin", 2.0, 1, precision="fp32")
    ctx.rk4("lin", 2.0, 5, precision="fp32")
    ctx.rk4("lin", 1.5, 3, precision="fp32")
    out = ctx.download()
    ctx.set_persistent(True)
    return out


def same(a, b):
    for f in range(3):
        assert np.array_equal(a[f], b[f]), (f, np.abs(a[f] - b[f]).max())


@pytest.mark.parametrize("nc,ng,ngf", [(3, 12, 4), (5, 25, 6), (2, 6, 3)])
def test_persistent_equals_per_stage_launches(nc, ng, ngf):
    t, ctx = ctx_for(nc, ng, ngf)
    c = configs.random_case(t, 12, 10, seed=31 + nc, scale=0.01)
    ctx.set_mesh(c["mesh"])
    set_fields(ctx, c["phys"])
    for form in range(3):
        for wg in (2, 4, 8):
            say("persistent", nc, ng, ngf, "form", form, "waves", wg)
            ctx.set_tiles("lin", wg + 16 * form, 0, autotune=False)
            same(calls(ctx, c, False), calls(ctx, c, True))
    ctx.set_tiles("lin", 0, 0, autotune=True)


@pytest.mark.parametrize("n", [28, 29])
def test_persistent_at_the_mesh_limit(n):
    """p = 3: 1,458 elements (92 blocks): persistent; 1,568: the per-stage
    path."""
    t, ctx = ctx_for(3, 12, 4)
    c = configs.random_case(t, n, n, seed=5, scale=0.01)
    say("persistent limit", c["mesh"]["ne"])
    ctx.set_mesh(c["mesh"])
    set_fields(ctx, c["phys"])
    same(calls(ctx, c, False), calls(ctx, c, True))


def test_persistent_config1_long_run():
    """The 512-element workload: 300 steps in one call, against the
    per-stage path (graphs of 8 steps)."""
    t, ctx = ctx_for(3, 12, 4)
    cfg = configs.config1(t)
    ctx.set_mesh(cfg["mesh"])
    set_fields(ctx, cfg["phys"])
    say("persistent config1", cfg["mesh"]["ne"])
    out = []
    for persist in (False, True):
        ctx.set_persistent(persist)
        ctx.upload(*cfg["state"])
        ctx.rk4("lin", cfg["dt"], 300, precision="fp32")
        out.append(ctx.download())
    ctx.set_persistent(True)
    same(out[0], out[1])


def test_persistent_two_contexts_at_once():
    """Two contexts (two streams) running persistent launches from two host
    threads at the same time: the launches on one device run one after the
    other, and each context gets the bits it gets alone."""
    import threading
    t, _ = ctx_for(3, 12, 4)
    cases = [configs.random_case(t, 20 + 4 * i, 18, seed=60 + i, scale=0.01) for i in range(2)]
    ctxs = [spoc_amd.Context(t) for _ in range(2)]
    alone = []
    for ctx, c in zip(ctxs, cases):
        ctx.set_mesh(c["mesh"])
        set_fields(ctx, c["phys"])
        ctx.upload(*c["state"])
        ctx.rk4("lin", 2.0, 200, precision="fp32")
        alone.append(ctx.download())
        ctx.upload(*c["state"])
    say("persistent two contexts", [c["mesh"]["ne"] for c in cases])
    errors = []

    def run(ctx):
        try:
            ctx.rk4("lin", 2.0, 200, precision="fp32")
        except Exception as ex:  # reported below, in the main thread
            errors.append(ex)

    th = [threading.Thread(target=run, args=(ctx,)) for ctx in ctxs]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    for ctx, ref in zip(ctxs, alone):
        same(ctx.download(), ref)
    for ctx in ctxs:
        ctx.close()


def test_persistent_ensemble_union():
    """Members of uneven size (40 to 646 triangles, 1,192 in all) as one
    union mesh: the persistent launch advances them all, each member equal
    bit for bit to the per-stage launches of the union and to its own run
    alone."""
    t, ctx = ctx_for(3, 12, 4)
    sizes = [(6, 5), (16, 12), (12, 9), (20, 18)]
    mem = [configs.random_case(t, nx, ny, seed=80 + i, scale=0.01)
           for i, (nx, ny) in enumerate(sizes)]
    say("persistent ensemble", [c["mesh"]["ne"] for c in mem])
    alone = []
    for c in mem:
        ctx.set_mesh(c["mesh"])
        set_fields(ctx, c["phys"])
        ctx.upload(*c["state"])
        ctx.rk4("lin", 2.0, 40, precision="fp32")
        alone.append(ctx.download())
    ctx.set_ensemble_mesh([c["mesh"] for c in mem])
    P = [c["phys"] for c in mem]
    cat = spoc_amd.Context.concat
    ctx.set_fields(cat([p["hb"] for p in P]), P[0]["g0"], cat([p["fc"] for p in P]),
                   cat([p["w0"] for p in P]), cat([p["z0"] for p in P]), P[0]["mu0"],
                   P[0]["mb0"])
    state = tuple(cat([c["state"][f] for c in mem]) for f in range(3))
    out = []
    for persist in (False, True):
        ctx.set_persistent(persist)
        ctx.upload(*state)
        ctx.rk4("lin", 2.0, 40, precision="fp32")
        out.append(ctx.download())
    ctx.set_persistent(True)
    same(out[0], out[1])
    e0 = 0
    for m, c in enumerate(mem):
        ne = c["mesh"]["ne"]
        for f in range(3):
            a, b = out[1][f][:t["nm"], e0:e0 + ne], alone[m][f][:t["nm"]]
            assert np.array_equal(a, b), (m, f)
        e0 += ne


def test_persistent_failure_path_forced_from_the_host():
    """With its abort word raised before the launch (spoc_set_persistent 2),
    every wave leaves at its first wait that is not yet satisfied and the
    call returns SPOC_EHIP with its message, not a hang; the next call, with
    the word left alone, runs normally."""
    t, ctx = ctx_for(3, 12, 4)
    c = configs.random_case(t, 12, 10, seed=9, scale=0.01)
    ctx.set_mesh(c["mesh"])
    set_fields(ctx, c["phys"])
    ctx.set_persistent(False)
    ctx.upload(*c["state"])
    ctx.rk4("lin", 2.0, 5, precision="fp32")
    ref = ctx.download()
    say("persistent forced abort")
    ctx.set_persistent(2)
    ctx.upload(*c["state"])
    with pytest.raises(spoc_amd.SpocError) as ei:
        ctx.rk4("lin", 2.0, 5, precision="fp32")
    assert ei.value.code == -2 and "neighbouring block" in str(ei.value), str(ei.value)
    ctx.set_persistent(True)
    ctx.upload(*c["state"])
    ctx.rk4("lin", 2.0, 5, precision="fp32")
    same(ctx.download(), ref)
    with pytest.raises(spoc_amd.SpocError):
        ctx.set_persistent(3)
