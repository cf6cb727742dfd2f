"""Mixed-precision MFMA path against the FP64 oracle.  Tolerances are
relative L2 in the mass-matrix norm, sqrt(sum_e det_e d^T M d / sum_e det_e
x^T M x).

- Nonlinear variants: FP32 quadrature (unit roundoff 6e-8) of the
  nonlinear, viscous, Coriolis and forcing terms on top of an FP64 linear
  core; one right-hand side measures ~1.4e-8, the bound is 1e-6.  Five RK4
  steps at p=6 measure ~7e-10, the bound is 1e-8.
- Linear variant: all FP64 (MFMA), in the mass-orthonormal basis, so the
  error is FP64 rounding amplified by the basis change (cond(U) grows with
  p): 9e-16 at p=2 up to 4e-9 at p=7 for one right-hand side, bound 1e-7;
  100 RK4 steps of config 1 (p=3) measure 1.5e-14, bound 1e-12."""
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import host_tables, oracle_for
from spoc_amd import configs

pytestmark = pytest.mark.gpu
SHAPES = [(2, 6, 3), (3, 12, 4), (4, 25, 5), (5, 25, 6), (6, 33, 7), (7, 42, 8)]
_ctx = {}


def ctx_for(nc, ng, ngf):
    if (nc, ng, ngf) not in _ctx:
        t = host_tables(nc, ng, ngf)
        _ctx[(nc, ng, ngf)] = (t, spoc_amd.Context(t))
    return _ctx[(nc, ng, ngf)]


def setup_case(ctx, c):
    ctx.set_mesh(c["mesh"])
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])


def mrel(t, m, a, b):
    nm = t["nm"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / m["vect"][4]
    d = a[:nm] - b[:nm]
    num = np.sum(det * np.einsum("ie,ij,je->e", d, M, d))
    den = np.sum(det * np.einsum("ie,ij,je->e", b[:nm], M, b[:nm]))
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [(v, "fp32") for v in ("lin", "fs", "ns", "inv")]


@pytest.mark.parametrize("nc,ng,ngf", SHAPES)
@pytest.mark.parametrize("variant,prec", CASES)
def test_cal_var_fp32_close_to_oracle(nc, ng, ngf, variant, prec):
    t, ctx = ctx_for(nc, ng, ngf)
    O = oracle_for(nc, ng, ngf)
    nx = 17 if nc <= 3 else 12
    c = configs.random_case(t, nx, nx - 2, seed=200 + nc, scale=0.05)
    say("f32 case", variant, prec, nc, ng, ngf, "ne", c["mesh"]["ne"])
    setup_case(ctx, c)
    got = ctx.cal_var(variant, *c["state"], precision=prec)
    ref = O.cal_var(variant, c["mesh"], c["phys"], *c["state"])
    errs = [mrel(t, c["mesh"], a, b) for a, b in zip(got, ref)]
    say("  M-norm rel err u v h:", errs)
    assert max(errs) < (1e-7 if variant == "lin" else 1e-6), errs


def test_rk4_fp32_linear_config1_eta_error():
    t, ctx = ctx_for(3, 12, 4)
    O = oracle_for(3, 12, 4)
    c = configs.config1(t)
    setup_case(ctx, c)
    ctx.upload(*c["state"])
    ctx.rk4("lin", c["dt"], 100, precision="fp32")
    got = ctx.download()
    ref = O.rk4("lin", c["mesh"], c["phys"], *c["state"], c["dt"], 100)
    e = [mrel(t, c["mesh"], a, b) for a, b in zip(got, ref)]
    say("config1 lin 100 steps fp32 rel err u v h:", e)
    assert max(e) < 1e-12, e


@pytest.mark.parametrize("variant,prec", CASES[1:])
def test_rk4_fp32_nonlinear_p6(variant, prec):
    t, ctx = ctx_for(6, 33, 7)
    O = oracle_for(6, 33, 7)
    c = configs.random_case(t, 13, 9, seed=31, scale=0.01)
    setup_case(ctx, c)
    ctx.upload(*c["state"])
    say("f32 rk4 p6", variant, prec)
    ctx.rk4(variant, 0.5, 5, precision=prec)
    got = ctx.download()
    ref = O.rk4(variant, c["mesh"], c["phys"], *c["state"], 0.5, 5)
    e = [mrel(t, c["mesh"], a, b) for a, b in zip(got, ref)]
    say(variant, prec, "p6 5 steps rel err u v h:", e)
    assert max(e) < 1e-8, e
