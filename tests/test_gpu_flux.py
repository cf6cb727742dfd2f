"""The local Lax-Friedrichs (Rusanov) option of the FP32-MFMA path's face
flux (spoc_set_flux).  It has no reference counterpart, so it is pinned by
what the scheme must do: alpha = 0 is the central flux bit for bit; the
penalty conserves mass (it moves the same flux across both sides of a
face); it dissipates energy where the central flux keeps it; it leaves a
smooth solution nearly unchanged; graph replay equals eager launches; the
paths that support the central flux only refuse it."""
import sys

import numpy as np
import pytest

import spoc_amd
from conftest import host_tables
from spoc_amd import configs

pytestmark = pytest.mark.gpu
_ctx = {}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lin_case():
    if "lin" not in _ctx:
        t = host_tables(3, 12, 4)
        c = configs.config1(t)
        ctx = spoc_amd.Context(t)
        ctx.set_mesh(c["mesh"])
        p = c["phys"]
        ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
        _ctx["lin"] = (t, c, ctx)
    return _ctx["lin"]


def run(ctx, c, alpha, nsteps, chunks=1, variant="lin"):
    ctx.set_flux(alpha)
    ctx.upload(*c["state"])
    e = [ctx.energy()]
    for _ in range(chunks):
        ctx.rk4(variant, c["dt"], nsteps // chunks, precision="fp32")
        e.append(ctx.energy())
    ctx.set_flux(0.0)
    return ctx.download(), e


def test_zero_alpha_is_the_central_flux():
    t, c, ctx = lin_case()
    say("flux: alpha 0 vs default")
    ctx.upload(*c["state"])
    ctx.rk4("lin", c["dt"], 20, precision="fp32")
    ref = ctx.download()
    got, _ = run(ctx, c, 0.0, 20)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_rusanov_conserves_mass_and_dissipates_energy():
    t, c, ctx = lin_case()
    say("flux: mass and energy, linear p=3")
    _, ec = run(ctx, c, 0.0, 200, chunks=4)
    _, er = run(ctx, c, 1.0, 200, chunks=4)
    tot = lambda e: e[0] + e[1]
    m0 = abs(ec[0][2]) + 1.0
    for (k, p, m) in er:
        assert abs(m - er[0][2]) < 1e-12 * m0, (m, er[0][2])
    # Rusanov: energy strictly decreasing, and below the central flux's
    # (which keeps it up to the RK4 time error)
    er_t = [tot(e) for e in er]
    say("  energy central", [tot(e) for e in ec], "rusanov", er_t)
    assert all(b < a for a, b in zip(er_t, er_t[1:])), er_t
    assert tot(er[-1]) < tot(ec[-1])


def test_rusanov_changes_a_smooth_solution_little():
    t, c, ctx = lin_case()
    say("flux: smooth solution")
    (uc, vc, hc), _ = run(ctx, c, 0.0, 100)
    (ur, vr, hr), _ = run(ctx, c, 1.0, 100)
    nm = t["nm"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / c["mesh"]["vect"][4]

    def rel(a, b):
        d = a[:nm] - b[:nm]
        return float(np.sqrt(np.sum(det * np.einsum("ie,ij,je->e", d, M, d)) /
                             np.sum(det * np.einsum("ie,ij,je->e", b[:nm], M, b[:nm]))))

    eh = rel(hr, hc)
    assert 0 < eh < 2e-2, eh


def test_nonlinear_variant_and_graph_replay():
    t = host_tables(4, 25, 5)
    c = configs.random_case(t, 12, 10, seed=31, scale=0.01)
    ctx = spoc_amd.Context(t)
    ctx.set_mesh(c["mesh"])
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
    say("flux: fs p=4, graph vs eager")
    out = []
    for graph in (False, True):
        ctx.set_graph(graph)
        ctx.set_flux(0.5)
        ctx.upload(*c["state"])
        ctx.rk4("fs", 2.0, 6, precision="fp32")
        out.append(ctx.download())
    ctx.set_flux(0.0)
    ctx.set_graph(True)
    for a, b in zip(*out):
        assert np.array_equal(a, b)
        assert np.isfinite(a).all()


def test_split_eg_path_p6():
    """At p = 6 E+G runs two waves per element block (u, v and eta); the
    option there, on the nonlinear free-slip variant started from the
    smooth Gaussian of the linear basin: graph replay equals eager launches,
    mass is conserved as with the central flux (the penalty moves the same
    flux across both sides of a face, and none through walls), and the
    smooth solution changes little."""
    t = host_tables(6, 33, 7)
    c = configs.config1(t)
    ctx = spoc_amd.Context(t)
    ctx.set_mesh(c["mesh"])
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
    say("flux: fs p=6 (split E+G)")
    out, mass = [], []
    for alpha, graph in ((0.0, True), (1.0, False), (1.0, True)):
        ctx.set_graph(graph)
        ctx.set_flux(alpha)
        ctx.upload(*c["state"])
        m0 = ctx.energy()[2]
        ctx.rk4("fs", c["dt"], 40, precision="fp32")
        mass.append((m0, ctx.energy()[2]))
        out.append(ctx.download())
    ctx.set_flux(0.0)
    ctx.set_graph(True)
    for a, b in zip(out[1], out[2]):
        assert np.array_equal(a, b)
        assert np.isfinite(a).all()
    scale = abs(mass[0][0]) + 1.0
    drift_c = abs(mass[0][1] - mass[0][0])
    drift_r = abs(mass[1][1] - mass[1][0])
    say("  mass drift central", drift_c, "rusanov", drift_r)
    assert drift_r <= 10 * drift_c + 1e-12 * scale
    nm = t["nm"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / c["mesh"]["vect"][4]
    d = out[1][2][:nm] - out[0][2][:nm]
    ref = out[0][2][:nm]
    eh = float(np.sqrt(np.sum(det * np.einsum("ie,ij,je->e", d, M, d)) /
                       np.sum(det * np.einsum("ie,ij,je->e", ref, M, ref))))
    say("  eta change", eh)
    assert 0 < eh < 2e-2, eh


def test_central_only_paths_refuse_the_option():
    t, c, ctx = lin_case()
    ctx.set_flux(1.0)
    ctx.upload(*c["state"])
    try:
        with pytest.raises(spoc_amd.SpocError):
            ctx.rk4("lin", c["dt"], 1, precision="fp64")
        with pytest.raises(spoc_amd.SpocError):
            ctx.cal_var("lin", *c["state"], precision="fp64")
    finally:
        ctx.set_flux(0.0)
    with pytest.raises(spoc_amd.SpocError):
        ctx.set_flux(-1.0)
