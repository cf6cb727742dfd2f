"""The library's internal element numbering (spoc_set_ordering) changes no
result: on a mesh whose caller numbering is shuffled, every output that
crosses the C-ABI (RK4 states of both paths, single right-hand sides,
energies, refinement indicators and decisions, the adjoint and its
bathymetry part) is the same bit for bit with the caller's numbering kept
(SPOC_ORDER_INPUT) and with the Hilbert and Z-order renumberings; likewise
for an ensemble, whose members are renumbered separately."""
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import host_tables
from spoc_amd import configs
from spoc_amd import mesh as M

pytestmark = pytest.mark.gpu
MODES = ("input", "hilbert", "zcurve", "auto")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def shuffled_case(t, nx, ny, seed):
    """random_case with the caller's element numbering shuffled (fields and
    state permuted with it: each is element-local)."""
    c = configs.random_case(t, nx, ny, seed=seed, scale=0.01)
    m = c["mesh"]
    perm = np.random.default_rng(seed + 1).permutation(m["ne"])
    m2 = M.build(m["x"], m["y"], np.asfortranarray(m["in"][:, perm]), t["nc"], 1e-4, 2e-11)
    p = {k: (np.asfortranarray(v[:, perm]) if isinstance(v, np.ndarray) and v.ndim == 2 else v)
         for k, v in c["phys"].items()}
    st = tuple(np.asfortranarray(a[:, perm]) for a in c["state"])
    return dict(mesh=m2, phys=p, state=st)


def run_all(ctx, c, variant):
    p = c["phys"]
    ctx.set_mesh(c["mesh"])
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
    out = {}
    for prec in ("fp64", "fp32"):
        out[f"rhs_{prec}"] = ctx.cal_var(variant, *c["state"], precision=prec)
        ctx.upload(*c["state"])
        ctx.rk4(variant, 2.0, 10, precision=prec)  # a graph of 8 steps and 2 eager
        out[f"rk4_{prec}"] = ctx.download()
        out[f"energy_{prec}"] = ctx.energy()
        r = ctx.cal_conv(0.05, 0.005, 10.0)
        out[f"calconv_{prec}"] = (r["fconv"], r["fderaff"], r["amax"], r["supmax"], r["conv"],
                                  r["deraff"], r["timedir"])
        out[f"checkconv_{prec}"] = ctx.checkconv(2)
    lam = [a * 1e-3 for a in c["state"]]
    out["adjoint_rhs"] = ctx.cal_var_adjoint(variant, *lam, state=c["state"],
                                             with_hb=variant != "lin")
    ctx.upload(*c["state"])
    ctx.adjoint_upload(*lam)
    ctx.rk4_adjoint(variant, 2.0, 3)
    out["adjoint_rk4"] = ctx.adjoint_download()
    if variant != "lin":
        out["adjoint_hb"] = ctx.adjoint_hb_download()
    return out


def same(a, b, path=""):
    if isinstance(a, (tuple, list)):
        for i, (x, y) in enumerate(zip(a, b)):
            same(x, y, f"{path}[{i}]")
    elif isinstance(a, np.ndarray):
        assert np.array_equal(a, b), (path, np.abs(a - b).max())
    else:
        assert a == b, (path, a, b)


@pytest.mark.parametrize("nc,ng,ngf", [(3, 12, 4), (6, 33, 7)])
@pytest.mark.parametrize("variant", ["lin", "fs", "ns"])
def test_results_independent_of_internal_order(nc, ng, ngf, variant):
    t = host_tables(nc, ng, ngf)
    ctx = spoc_amd.Context(t)
    c = shuffled_case(t, 14, 11, 60 + nc)
    say("order", variant, nc, "ne", c["mesh"]["ne"])
    res = {}
    for mode in MODES:
        ctx.set_ordering(mode)
        res[mode] = run_all(ctx, c, variant)
    ctx.close()
    for mode in MODES[1:]:
        for k in res["input"]:
            same(res[mode][k], res["input"][k], f"{mode}:{k}")


def test_ensemble_members_renumbered_separately():
    t = host_tables(4, 25, 5)
    ctx = spoc_amd.Context(t)
    mem = [shuffled_case(t, nx, ny, s) for nx, ny, s in ((10, 8, 1), (13, 9, 2), (9, 9, 3))]
    P = [c["phys"] for c in mem]
    cat = spoc_amd.Context.concat
    res = {}
    for mode in MODES:
        ctx.set_ordering(mode)
        ctx.set_ensemble_mesh([c["mesh"] for c in mem])
        ctx.set_fields(cat([p["hb"] for p in P]), P[0]["g0"], cat([p["fc"] for p in P]),
                       cat([p["w0"] for p in P]), cat([p["z0"] for p in P]), P[0]["mu0"],
                       P[0]["mb0"])
        r = {}
        for prec in ("fp64", "fp32"):
            ctx.upload(*(cat([c["state"][f] for c in mem]) for f in range(3)))
            ctx.rk4("fs", 2.0, 5, precision=prec)
            r[prec] = ctx.download()
            r[f"energy_{prec}"] = ctx.ensemble_energy()
            cc = ctx.cal_conv(0.05, 0.005, 10.0)
            r[f"cc_{prec}"] = (cc["fconv"], cc["fderaff"], cc["amax"], cc["supmax"])
        res[mode] = r
    ctx.close()
    for mode in MODES[1:]:
        for k in res["input"]:
            same(res[mode][k], res["input"][k], f"{mode}:{k}")
