"""Discrete adjoint (no reference counterpart), pinned by the identities that
define it rather than by an oracle:

- dot-product test for one right-hand side: <R v, w> = <v, R^T w>, with R v
  from the FP64 path (the reference's own arithmetic) and R^T w from the
  adjoint kernels;
- the same for N RK4 steps: <Phi^N v, w> = <v, (Phi^N)^T w>;
- a gradient: J(x0) = sum over elements inside a disc of the exact
  integral of eta^2 after N steps; the adjoint gives dJ/dx0, checked
  against central differences (exact up to rounding: J is quadratic in x0
  for the linear variant)."""
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


def setup(ctx, c):
    ctx.set_mesh(c["mesh"])
    p = c["phys"]
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])


def rand_state(rng, nm, ne, ld=45):
    out = []
    for _ in range(3):
        a = np.zeros((ld, ne), order="F")
        a[:nm] = rng.standard_normal((nm, ne))
        out.append(a)
    return out


def dot(a, b, nm):
    return sum(float(np.sum(x[:nm] * y[:nm])) for x, y in zip(a, b))


SHAPES = [(2, 6, 3), (3, 12, 4), (4, 25, 5), (5, 25, 6), (6, 33, 7), (7, 42, 8)]


@pytest.mark.parametrize("nc,ng,ngf", SHAPES)
def test_rhs_adjoint_dot_product(nc, ng, ngf):
    t, ctx = ctx_for(nc, ng, ngf)
    nx = 12 if nc <= 4 else 9
    c = configs.random_case(t, nx, nx - 2, seed=500 + nc)
    setup(ctx, c)
    nm, ne = t["nm"], c["mesh"]["ne"]
    rng = np.random.default_rng(nc)
    v, w = rand_state(rng, nm, ne), rand_state(rng, nm, ne)
    Rv = ctx.cal_var("lin", *v, precision="fp64")
    RTw = ctx.cal_var_adjoint("lin", *w)
    lhs, rhs = dot(Rv, w, nm), dot(v, RTw, nm)
    scale = np.sqrt(dot(Rv, Rv, nm) * dot(w, w, nm))
    say("adjoint dot p=%d" % nc, lhs, rhs, abs(lhs - rhs) / scale)
    assert abs(lhs - rhs) <= 1e-11 * scale, (lhs, rhs)


@pytest.mark.parametrize("nc,ng,ngf", [(3, 12, 4), (6, 33, 7)])
def test_rk4_adjoint_dot_product(nc, ng, ngf):
    t, ctx = ctx_for(nc, ng, ngf)
    c = configs.random_case(t, 11, 9, seed=600 + nc)
    setup(ctx, c)
    nm, ne = t["nm"], c["mesh"]["ne"]
    rng = np.random.default_rng(60 + nc)
    v, w = rand_state(rng, nm, ne), rand_state(rng, nm, ne)
    dt, n = 20.0, 12
    ctx.upload(*v)
    ctx.rk4("lin", dt, n, precision="fp64")
    Phv = ctx.download()
    ctx.adjoint_upload(*w)
    ctx.rk4_adjoint("lin", dt, n)
    PhTw = ctx.adjoint_download()
    lhs, rhs = dot(Phv, w, nm), dot(v, PhTw, nm)
    scale = np.sqrt(dot(Phv, Phv, nm) * dot(w, w, nm))
    say("rk4 adjoint dot p=%d" % nc, lhs, rhs, abs(lhs - rhs) / scale)
    assert abs(lhs - rhs) <= 1e-11 * scale, (lhs, rhs)


def test_gradient_of_disc_energy_matches_central_differences():
    t, ctx = ctx_for(3, 12, 4)
    c = configs.config1(t)
    setup(ctx, c)
    m = c["mesh"]
    nm, ne = t["nm"], m["ne"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / m["vect"][4]
    # centroids from the node arrays; the disc of radius 200 km at the origin
    tri = np.asarray(m["in"]) - 1
    cx, cy = m["x"][tri].mean(axis=0), m["y"][tri].mean(axis=0)
    inside = (cx ** 2 + cy ** 2) < 200e3 ** 2
    dt, n = c["dt"], 25

    def J(x0):
        ctx.upload(*x0)
        ctx.rk4("lin", dt, n, precision="fp64")
        h = ctx.download()[2][:nm]
        return float(np.sum(det[inside] * np.einsum("ie,ij,je->e", h[:, inside], M,
                                                    h[:, inside])))

    x0 = [np.asfortranarray(a) for a in c["state"]]
    ctx.upload(*x0)
    ctx.rk4("lin", dt, n, precision="fp64")
    h = ctx.download()[2]
    lam_h = np.zeros_like(h)
    lam_h[:nm, inside] = 2.0 * det[inside] * (M @ h[:nm, inside])
    zero = np.zeros_like(h)
    ctx.adjoint_upload(zero, zero, lam_h)
    ctx.rk4_adjoint("lin", dt, n)
    grad = ctx.adjoint_download()
    rng = np.random.default_rng(4001)
    for d in range(3):
        v = rand_state(rng, nm, ne)
        eps = 1e-3
        jp = J([a + eps * b for a, b in zip(x0, v)])
        jm = J([a - eps * b for a, b in zip(x0, v)])
        fd = (jp - jm) / (2 * eps)
        ad = dot(grad, v, nm)
        say("gradient direction", d, fd, ad, abs(fd - ad) / abs(fd))
        assert abs(fd - ad) <= 1e-8 * abs(fd), (fd, ad)


# -- nonlinear variants -------------------------------------------------------
# The Jacobian depends on the state: J_R(x) v is taken by central differences
# of the FP64 path (the reference's arithmetic), so the identities hold to
# the differencing error, well under the 1e-6 bound.

def fd_rhs(ctx, variant, x, v, eps):
    xp = [a + eps * b for a, b in zip(x, v)]
    xm = [a - eps * b for a, b in zip(x, v)]
    rp = ctx.cal_var(variant, *xp, precision="fp64")
    rm = ctx.cal_var(variant, *xm, precision="fp64")
    return [(a - b) / (2 * eps) for a, b in zip(rp, rm)]


@pytest.mark.parametrize("variant", ["fs", "ns", "inv"])
@pytest.mark.parametrize("nc,ng,ngf", [(3, 12, 4), (6, 33, 7)])
def test_nonlinear_rhs_adjoint_against_differences(variant, nc, ng, ngf):
    t, ctx = ctx_for(nc, ng, ngf)
    c = configs.random_case(t, 10, 8, seed=700 + nc, scale=0.05)
    if variant == "inv":
        c["phys"]["mu0"] = 0.0
    setup(ctx, c)
    nm, ne = t["nm"], c["mesh"]["ne"]
    rng = np.random.default_rng(70 + nc)
    x = [np.asfortranarray(a) for a in c["state"]]
    v, w = rand_state(rng, nm, ne), rand_state(rng, nm, ne)
    for a in v:
        a *= 0.05
    Jv = fd_rhs(ctx, variant, x, v, 1e-3)
    *JTw, hbbar = ctx.cal_var_adjoint(variant, *w, state=x, with_hb=True)
    lhs, rhs = dot(Jv, w, nm), dot(v, JTw, nm)
    scale = np.sqrt(dot(Jv, Jv, nm) * dot(w, w, nm))
    say("nonlinear adjoint", variant, "p=%d" % nc, lhs, rhs, abs(lhs - rhs) / scale)
    assert abs(lhs - rhs) <= 1e-7 * scale, (lhs, rhs)
    # bathymetry: d/d(eps) R(x; hb + eps z) against <z, hbbar>
    p = c["phys"]
    z = np.zeros_like(p["hb"])
    z[:nm] = rng.standard_normal((nm, ne)) * 1.0
    eps = 1e-2
    r = []
    for sgn in (1, -1):
        ctx.set_fields(p["hb"] + sgn * eps * z, p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"],
                       p["mb0"])
        r.append(ctx.cal_var(variant, *x, precision="fp64"))
    setup(ctx, c)
    dRz = [(a - b) / (2 * eps) for a, b in zip(*r)]
    lhs, rhs = dot(dRz, w, nm), float(np.sum(z[:nm] * hbbar[:nm]))
    scale = np.sqrt(dot(dRz, dRz, nm) * dot(w, w, nm))
    say("  hb adjoint", lhs, rhs, abs(lhs - rhs) / scale)
    assert abs(lhs - rhs) <= 1e-7 * scale, (lhs, rhs)


@pytest.mark.parametrize("variant", ["fs", "inv"])
def test_nonlinear_gradient_against_central_differences(variant):
    """J = exact integral of eta^2 over the elements inside a 200 km disc
    after n steps; gradients with respect to the initial eta and to the
    bathymetry from the checkpointed adjoint (interval 3 over 8 steps, so
    three segments), against central differences of the FP64 path along
    random directions."""
    t, ctx = ctx_for(5, 25, 6)
    c = configs.config4(t, nx=24, ny=13)
    if variant == "fs":
        c["phys"]["mu0"] = 100.0
    setup(ctx, c)
    m, p = c["mesh"], c["phys"]
    nm, ne = t["nm"], m["ne"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / m["vect"][4]
    tri = np.asarray(m["in"]) - 1
    cx, cy = m["x"][tri].mean(axis=0), m["y"][tri].mean(axis=0)
    inside = (cx ** 2 + cy ** 2) < 400e3 ** 2
    dt, n = c["dt"], 8
    x0 = [np.asfortranarray(a) for a in c["state"]]

    def J(x, hb=None):
        if hb is not None:
            ctx.set_fields(hb, p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
        ctx.upload(*x)
        ctx.rk4(variant, dt, n, precision="fp64")
        h = ctx.download()[2][:nm]
        if hb is not None:
            setup(ctx, c)
        return float(np.sum(det[inside] * np.einsum("ie,ij,je->e", h[:, inside], M,
                                                    h[:, inside])))

    ctx.upload(*x0)
    ctx.rk4(variant, dt, n, precision="fp64")
    h = ctx.download()[2]
    lam_h = np.zeros_like(h)
    lam_h[:nm, inside] = 2.0 * det[inside] * (M @ h[:nm, inside])
    zero = np.zeros_like(h)
    ctx.upload(*x0)
    ctx.adjoint_upload(zero, zero, lam_h)
    ctx.set_checkpoint(3)
    ctx.rk4_adjoint(variant, dt, n)
    grad = ctx.adjoint_download()
    ghb = ctx.adjoint_hb_download()
    # every stage input kept from the forward sweep (automatic on a small
    # mesh) gives the same numbers as recomputing them from checkpoints
    ctx.upload(*x0)
    ctx.adjoint_upload(zero, zero, lam_h)
    ctx.set_checkpoint(0)
    ctx.rk4_adjoint(variant, dt, n)
    for f in range(3):
        assert np.array_equal(ctx.adjoint_download()[f], grad[f]), f
    assert np.array_equal(ctx.adjoint_hb_download(), ghb)
    rng = np.random.default_rng(4001)
    for d in range(2):
        v = np.zeros_like(h)
        v[:nm] = rng.standard_normal((nm, ne)) * 0.01
        eps = 1e-2
        fd = (J([x0[0], x0[1], x0[2] + eps * v]) - J([x0[0], x0[1], x0[2] - eps * v])) / (2 * eps)
        ad = float(np.sum(grad[2][:nm] * v[:nm]))
        say("eta0 direction", d, fd, ad, abs(fd - ad) / abs(fd))
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
        z = np.zeros_like(h)
        z[:nm] = rng.standard_normal((nm, ne)) * 1.0
        eps = 1e-1
        fd = (J(x0, p["hb"] + eps * z) - J(x0, p["hb"] - eps * z)) / (2 * eps)
        ad = float(np.sum(ghb[:nm] * z[:nm]))
        say("hb direction", d, fd, ad, abs(fd - ad) / abs(fd))
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


def test_autograd_function_wraps_forward_and_adjoint():
    """Gradient-check harness: the RK4 run as a torch.autograd.Function
    (forward = spoc_rk4_steps, backward = spoc_rk4_adjoint_steps), with the
    cost written in torch; the gradient torch returns for the initial eta
    matches central differences."""
    import torch

    t, ctx = ctx_for(5, 25, 6)
    c = configs.config4(t, nx=20, ny=11)
    setup(ctx, c)
    m = c["mesh"]
    nm, ne = t["nm"], m["ne"]
    variant, dt, n = "inv", c["dt"], 6
    u0, v0, h0 = (np.asfortranarray(a) for a in c["state"])

    class RK4(torch.autograd.Function):
        @staticmethod
        def forward(fctx, eta0):
            h = np.asfortranarray(eta0.detach().numpy())
            fctx.save_for_backward(eta0)
            ctx.upload(u0, v0, h)
            ctx.rk4(variant, dt, n, precision="fp64")
            return torch.from_numpy(np.ascontiguousarray(ctx.download()[2]))

        @staticmethod
        def backward(fctx, g):
            (eta0,) = fctx.saved_tensors
            z = np.zeros_like(u0)
            ctx.upload(u0, v0, np.asfortranarray(eta0.detach().numpy()))
            ctx.adjoint_upload(z, z, np.asfortranarray(g.numpy()))
            ctx.rk4_adjoint(variant, dt, n)
            return torch.from_numpy(np.ascontiguousarray(ctx.adjoint_download()[2]))

    M = torch.from_numpy(np.ascontiguousarray(t["ammo"][:nm, :nm]))
    det = torch.from_numpy(1.0 / m["vect"][4])

    def cost(h):
        hh = h[:nm]
        return torch.sum(det * torch.einsum("ie,ij,je->e", hh, M, hh))

    eta0 = torch.from_numpy(np.ascontiguousarray(h0)).requires_grad_(True)
    J = cost(RK4.apply(eta0))
    J.backward()
    grad = eta0.grad.numpy()
    rng = np.random.default_rng(11)
    v = np.zeros_like(h0)
    v[:nm] = rng.standard_normal((nm, ne)) * 0.01
    eps = 1e-2
    with torch.no_grad():
        jp = float(cost(RK4.apply(torch.from_numpy(np.ascontiguousarray(h0 + eps * v)))))
        jm = float(cost(RK4.apply(torch.from_numpy(np.ascontiguousarray(h0 - eps * v)))))
    fd, ad = (jp - jm) / (2 * eps), float(np.sum(grad[:nm] * v[:nm]))
    say("autograd harness", fd, ad, abs(fd - ad) / abs(fd))
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
