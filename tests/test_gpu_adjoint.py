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
