"""Runtime features of the device path with no reference counterpart, held
to exact equality with the plain runs they replace:

- ensemble: members run as one union mesh give, member by member, the same
  bits as each member run alone (every kernel treats an element column
  independently), for both arithmetic paths; per-member energies likewise;
- graph replay of the RK4 step gives the same bits as eager launches,
  across replays separated by eager work, after a change of dt, and with
  several steps per graph."""
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


def members(t, seed0):
    sizes = [(9, 7), (11, 9), (8, 8)]
    return [configs.random_case(t, nx, ny, seed=seed0 + i, scale=0.01)
            for i, (nx, ny) in enumerate(sizes)]


SHAPES = [(3, 12, 4), (6, 33, 7)]


@pytest.mark.parametrize("nc,ng,ngf", SHAPES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", ["lin", "fs", "ns"])
def test_ensemble_members_equal_single_runs(nc, ng, ngf, prec, variant):
    t, ctx = ctx_for(nc, ng, ngf)
    mem = members(t, 40 + nc)
    say("ensemble", variant, prec, nc, [c["mesh"]["ne"] for c in mem])
    dt, n = 2.0, 4
    single = []
    for c in mem:
        ctx.set_mesh(c["mesh"])
        set_fields(ctx, c["phys"])
        ctx.upload(*c["state"])
        ctx.rk4(variant, dt, n, precision=prec)
        single.append(ctx.download())
    ctx.set_ensemble_mesh([c["mesh"] for c in mem])
    P = [c["phys"] for c in mem]
    cat = spoc_amd.Context.concat
    ctx.set_fields(cat([p["hb"] for p in P]), P[0]["g0"], cat([p["fc"] for p in P]),
                   cat([p["w0"] for p in P]), cat([p["z0"] for p in P]), P[0]["mu0"],
                   P[0]["mb0"])
    ctx.upload(*(cat([c["state"][f] for c in mem]) for f in range(3)))
    ctx.rk4(variant, dt, n, precision=prec)
    got = ctx.download()
    e0 = 0
    for m, c in enumerate(mem):
        ne = c["mesh"]["ne"]
        for f in range(3):
            a, b = got[f][:t["nm"], e0:e0 + ne], single[m][f][:t["nm"]]
            assert np.array_equal(a, b), (m, f, np.abs(a - b).max())
        e0 += ne


def test_ensemble_energy_equals_single_energies():
    t, ctx = ctx_for(3, 12, 4)
    mem = members(t, 70)
    ref = []
    for c in mem:
        ctx.set_mesh(c["mesh"])
        set_fields(ctx, c["phys"])
        ctx.upload(*c["state"])
        ctx.rk4("fs", 2.0, 3, precision="fp64")
        ref.append(ctx.energy())
    ctx.set_ensemble_mesh([c["mesh"] for c in mem])
    P = [c["phys"] for c in mem]
    cat = spoc_amd.Context.concat
    ctx.set_fields(cat([p["hb"] for p in P]), P[0]["g0"], cat([p["fc"] for p in P]),
                   cat([p["w0"] for p in P]), cat([p["z0"] for p in P]), P[0]["mu0"],
                   P[0]["mb0"])
    ctx.upload(*(cat([c["state"][f] for c in mem]) for f in range(3)))
    ctx.rk4("fs", 2.0, 3, precision="fp64")
    ke, pe, mass = ctx.ensemble_energy()
    for m in range(len(mem)):
        assert (ke[m], pe[m], mass[m]) == ref[m], (m, (ke[m], pe[m], mass[m]), ref[m])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", ["lin", "fs"])
def test_graph_replay_equals_eager(prec, variant):
    t, ctx = ctx_for(3, 12, 4)
    c = configs.random_case(t, 10, 8, seed=91, scale=0.01)
    ctx.set_mesh(c["mesh"])
    set_fields(ctx, c["phys"])
    say("graph", variant, prec)

    def run(graph):
        ctx.set_graph(graph)
        ctx.set_persistent(False)  # (the linear steps would not use the graph)
        ctx.upload(*c["state"])
        ctx.rk4(variant, 2.0, 3, precision=prec)   # graph captured, replayed 3x
        ctx.rk4(variant, 2.0, 1, precision=prec)   # eager step between replays
        ctx.rk4(variant, 2.0, 4, precision=prec)   # cached graph replayed again
        ctx.rk4(variant, 1.5, 2, precision=prec)   # new dt: captured again
        ctx.rk4(variant, 1.5, 19, precision=prec)  # 8-step graph twice, 3 eager
        ctx.rk4(variant, 1.5, 8, precision=prec)   # the 8-step graph again
        return ctx.download()

    eager, graph = run(False), run(True)
    ctx.set_graph(True)
    ctx.set_persistent(True)
    for f in range(3):
        assert np.array_equal(eager[f], graph[f]), (f, np.abs(eager[f] - graph[f]).max())


@pytest.mark.parametrize("variant", ["lin", "fs", "ns"])
def test_tiles_do_not_change_results(variant):
    """The autotuner's choices (2, 4 or 8 waves of 16 elements per
    workgroup, each kernel in 3 compiled forms: register budget and the E+G
    split) only change how element blocks map to waves and workgroups."""
    t, ctx = ctx_for(6, 33, 7)
    c = configs.random_case(t, 14, 12, seed=77, scale=0.01)
    ctx.set_mesh(c["mesh"])
    set_fields(ctx, c["phys"])
    tiles = ctx.autotune(variant)
    say("autotuned tiles", variant, tiles)
    assert tiles[0] & 15 in (2, 4, 8) and tiles[0] >> 4 in (0, 1, 2)
    assert tiles[1] == 0 if variant == "lin" else (tiles[1] & 15 in (2, 4, 8) and
                                                   tiles[1] >> 4 in (0, 1, 2))
    res = []
    for t0, t1 in [(2, 2), (4, 4), (8, 8), (2, 8), (8, 4), (4 + 16, 4 + 16), (8 + 32, 2 + 32),
                   (2 + 16, 8 + 32), (4 + 32, 4 + 16)]:
        ctx.set_tiles(variant, t0, t1)
        ctx.upload(*c["state"])
        ctx.rk4(variant, 2.0, 3, precision="fp32")
        res.append(ctx.download())
    ctx.set_tiles(variant, 0, 0)
    for r in res[1:]:
        for f in range(3):
            assert np.array_equal(r[f], res[0][f])
