"""Failure detection of the device path: the driver's testnan
(src/shwwnl.f:426-445, called at src/shwwnl.f:297) stops the run when the
kinetic energy prints as NaN.  spoc_rk4_steps reports SPOC_ENAN when any
coefficient of the new state is NaN or infinite, spoc_energy when the
kinetic energy is NaN; the state stays downloadable.  The non-finite value
is planted in the input on the host; every kernel runs to its normal end.

Also: a long run queues thousands of dispatches, which spoc_rk4_steps
bounds to 512 ahead of the host; the results equal an unbounded eager
run's."""
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


@pytest.fixture(scope="module")
def case():
    t = host_tables(3, 12, 4)
    ctx = spoc_amd.Context(t)
    c = configs.random_case(t, 8, 7, seed=21, scale=0.01)
    ctx.set_mesh(c["mesh"])
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
    return t, ctx, c


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", ["lin", "fs"])
def test_nan_in_state_is_reported(case, prec, variant):
    t, ctx, c = case
    say("nan", variant, prec)
    ctx.upload(*c["state"])
    ctx.rk4(variant, 2.0, 3, precision=prec)  # finite: no error
    assert np.isfinite(ctx.energy()[0])
    h = np.array(c["state"][2], order="F")
    h[0, 5] = np.nan
    ctx.upload(c["state"][0], c["state"][1], h)
    with pytest.raises(spoc_amd.SpocError) as ei:
        ctx.rk4(variant, 2.0, 3, precision=prec)
    assert ei.value.code == spoc_amd.ENAN and "NaN" in str(ei.value)
    u, v, hh = ctx.download()  # the state is still there
    assert not np.isfinite(hh[:t["nm"]]).all()
    with pytest.raises(spoc_amd.SpocError) as ei:
        ctx.energy()
    assert ei.value.code == spoc_amd.ENAN


def test_infinity_in_state_is_reported(case):
    t, ctx, c = case
    u = np.array(c["state"][0], order="F")
    u[2, 9] = np.inf
    ctx.upload(u, c["state"][1], c["state"][2])
    with pytest.raises(spoc_amd.SpocError) as ei:
        ctx.rk4("lin", 2.0, 1, precision="fp64")
    assert ei.value.code == spoc_amd.ENAN


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_long_run_bounded_queue_equals_eager(case, prec):
    """300 steps: 4,800 FP64 (2,400 mixed) dispatches in one call, queued at
    most 512 ahead of the host, from graphs of 8 steps and eagerly."""
    t, ctx, c = case
    say("long run", prec)
    out = []
    for graph in (True, False):
        ctx.set_graph(graph)
        ctx.upload(*c["state"])
        ctx.rk4("fs", 1.0, 300, precision=prec)
        out.append(ctx.download())
    ctx.set_graph(True)
    for f in range(3):
        assert np.array_equal(out[0][f], out[1][f])
