"""Synthetic workloads: meshes, physics fields and initial states.

Meshes follow the reference's preparation chain (structured mesh as
src/rmal.f, corner fix and edge-1-on-boundary rotation as src/chmal.f, then
the driver's pointers); initial fields are L2-projected element by element
the way src/init.f:290-344 (calh) does.  Every random draw is seeded.
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as M
from . import nm_of

LDM = 45


def cfl_dt(x, y, tri, nc: int, hb1, g0: float, ra: float = 0.2) -> float:
    """Time step from the CFL rule and rounding ladder of src/cfl.f:4-44
    (with the safety factor ra)."""
    i1, i2, i3 = tri[0] - 1, tri[1] - 1, tri[2] - 1
    cc0 = np.sqrt(g0 * hb1)
    d = []
    for a, b in ((i1, i2), (i1, i3), (i2, i3)):
        L = np.sqrt((x[b] - x[a]) ** 2 + (y[b] - y[a]) ** 2)
        d.append(L / float(nc) * ra / cc0)
    cfl = float(min(6000.0, np.min(np.minimum(np.minimum(d[0], d[1]), d[2]))))
    if cfl >= 100:
        return float(10 * int(cfl / 10))
    if cfl >= 10:
        return float(2 * int(cfl / 2))
    if cfl >= 1:
        return float(int(cfl))
    return float(int(cfl * 10)) / 10.0


def cfl_dt_altitude(x, y, tri, nc: int, hb1, g0: float, ra: float = 0.2) -> float:
    """CFL time step with the smallest triangle altitude in place of the
    edge length of src/cfl.f:15-31 (same rounding ladder).  On jittered
    meshes the edge rule overestimates the stable step of the central-flux
    scheme (measured: unstable at the edge-rule dt on a 60 x 30 node mesh)."""
    i1, i2, i3 = tri[0] - 1, tri[1] - 1, tri[2] - 1
    L = [np.hypot(x[b] - x[a], y[b] - y[a]) for a, b in ((i1, i2), (i1, i3), (i2, i3))]
    area2 = np.abs((x[i2] - x[i1]) * (y[i3] - y[i1]) - (x[i3] - x[i1]) * (y[i2] - y[i1]))
    hmin = area2 / np.maximum(np.maximum(L[0], L[1]), L[2])
    cfl = float(np.min(hmin / float(nc) * ra / np.sqrt(g0 * hb1)))
    if cfl >= 100:
        return float(10 * int(cfl / 10))
    if cfl >= 10:
        return float(2 * int(cfl / 2))
    if cfl >= 1:
        return float(int(cfl))
    return float(int(cfl * 10)) / 10.0


def structured_mesh(nx, ny, xmin, xmax, ymin, ymax):
    """RMAL grid followed by the CM corner fix."""
    x, y, tri = M.rmal(nx, ny, xmin, xmax, ymin, ymax)
    tri = M.changemaille(len(x), tri)
    return x, y, tri


def _morton(ix, iy):
    """Interleave the bits of two non-negative integer arrays (Z order)."""
    z = np.zeros_like(ix, dtype=np.int64)
    for b in range(21):
        z |= ((ix >> b) & 1) << (2 * b)
        z |= ((iy >> b) & 1) << (2 * b + 1)
    return z


def jittered_mesh(nx, ny, lx, ly, jitter=0.3, seed=0):
    """Jittered-node Delaunay mesh over [-lx/2, lx/2] x [-ly/2, ly/2].

    Interior nodes of an nx x ny lattice move uniformly by up to
    +-jitter*spacing in each coordinate; boundary nodes slide along the
    boundary only.  The Delaunay triangulation of the nodes (2(nx-1)(ny-1)
    triangles) is oriented counter-clockwise, its triangles are numbered
    along a Z-order curve of their centroids (neighbours stay close in
    memory), then the CM corner fix is applied.  Nodes are numbered x
    fastest like src/rmal.f."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    hx, hy = lx / (nx - 1), ly / (ny - 1)
    xs = -lx / 2 + hx * np.arange(nx)
    ys = -ly / 2 + hy * np.arange(ny)
    X, Y = np.meshgrid(xs, ys)  # (ny, nx)
    jx = rng.uniform(-jitter, jitter, X.shape) * hx
    jy = rng.uniform(-jitter, jitter, X.shape) * hy
    interior_x = np.ones_like(X, bool)
    interior_x[:, 0] = interior_x[:, -1] = False
    interior_y = np.ones_like(X, bool)
    interior_y[0, :] = interior_y[-1, :] = False
    X = X + np.where(interior_x, jx, 0.0)
    Y = Y + np.where(interior_y, jy, 0.0)
    x, y = X.ravel(), Y.ravel()
    s = Delaunay(np.stack([x, y], 1)).simplices.astype(np.int64)
    area2 = (x[s[:, 1]] - x[s[:, 0]]) * (y[s[:, 2]] - y[s[:, 0]]) - \
            (x[s[:, 2]] - x[s[:, 0]]) * (y[s[:, 1]] - y[s[:, 0]])
    flip = area2 < 0
    s[flip, 1], s[flip, 2] = s[flip, 2].copy(), s[flip, 1].copy()
    cx = (x[s].mean(1) + lx / 2) / hx
    cy = (y[s].mean(1) + ly / 2) / hy
    order = np.argsort(_morton((cx * 2).astype(np.int64), (cy * 2).astype(np.int64)),
                       kind="stable")
    tri = np.asfortranarray((s[order] + 1).T.astype(np.int32))
    tri = M.changemaille(len(x), tri)
    return x, y, tri


def project(tables: dict, x, y, tri, f, ld: int = LDM) -> np.ndarray:
    """L2 projection of f(x, y) onto the modal basis, element by element
    (src/init.f:290-344: values at the quadrature points, test-function
    integrals, Cholesky solve with the mass matrix)."""
    import scipy.linalg as sl

    nm, ng = tables["nm"], tables["ng"]
    xi, yi, wi = tables["xi"], tables["yi"], tables["wi"]
    pi2 = tables["pi2"][:ng, :nm]
    i1, i2, i3 = tri[0] - 1, tri[1] - 1, tri[2] - 1
    agx = (x[i2] - x[i1]) * 0.5
    agy = (y[i2] - y[i1]) * 0.5
    bgx = (x[i3] - x[i1]) * 0.5
    bgy = (y[i3] - y[i1]) * 0.5
    X = agx[None] * (xi[:, None] + 1.0) + bgx[None] * (yi[:, None] + 1.0) + x[i1][None]
    Y = agy[None] * (xi[:, None] + 1.0) + bgy[None] * (yi[:, None] + 1.0) + y[i1][None]
    F = f(X, Y)  # (ng, ne)
    B = pi2.T @ (wi[:, None] * F)  # (nm, ne)
    U = np.triu(tables["amm"][:nm, :nm])
    Z = sl.solve_triangular(U, B, trans="T", lower=False)
    Z = sl.solve_triangular(U, Z, lower=False)
    out = np.zeros((ld, tri.shape[1]), order="F")
    out[:nm] = Z
    return out


def make_case(tables: dict, x, y, tri, *, f0=0.0, beta=0.0, g0=9.81, H=4000.0, mu0=0.0,
              mb0=0.0, wind=None, eta=None, u=None, v=None, ld: int = LDM) -> dict:
    """Mesh operands, physics fields and initial state of one run."""
    nc = tables["nc"]
    m = M.build(x, y, tri, nc, f0, beta)
    ne = m["ne"]
    zero = lambda: np.zeros((ld, ne), order="F")
    hb = zero()
    if callable(H):
        hb = project(tables, x, y, tri, H, ld)
    else:
        hb[0] = H
    w0 = project(tables, x, y, tri, wind, ld) if wind is not None else zero()
    z0 = zero()
    h0 = eta if isinstance(eta, np.ndarray) else (
        project(tables, x, y, tri, eta, ld) if eta is not None else zero())
    u0 = u if isinstance(u, np.ndarray) else (
        project(tables, x, y, tri, u, ld) if u is not None else zero())
    v0 = v if isinstance(v, np.ndarray) else (
        project(tables, x, y, tri, v, ld) if v is not None else zero())
    fc = np.zeros((ld, ne), order="F")
    fc[:] = m["fc"][:ld]
    return dict(mesh=m, phys=dict(mu0=mu0, mb0=mb0, g0=g0, fc=fc, w0=w0, z0=z0, hb=hb,
                                  beta=beta),
                state=(u0, v0, h0))


def config1(tables: dict) -> dict:
    """Linear basin: 1000 km square, 16x16 squares cut into 512 triangles,
    p=3, H=4000 m, 1 m Gaussian eta of radius 100 km, free-slip walls,
    dt from CFL 0.2 (= 20 s)."""
    x, y, tri = structured_mesh(17, 17, -500e3, 500e3, -500e3, 500e3)
    R = 100e3
    c = make_case(tables, x, y, tri, f0=1e-4, g0=9.81, H=4000.0,
                  eta=lambda X, Y: np.exp(-(X * X + Y * Y) / (R * R)))
    c["dt"] = cfl_dt(x, y, tri, tables["nc"], 4000.0, 9.81)
    c["variant"], c["nsteps"] = "lin", 1000
    return c


def config2(tables: dict, nx: int = 448, ny: int = 224, seed: int = 2001) -> dict:
    """Nonlinear free-slip double gyre: 4000 x 2000 km jittered Delaunay mesh
    (199,362 triangles at 448 x 224 nodes), p=6, beta plane f0=1e-4, beta=2e-11,
    H=1000 m, wind stress 0.1 N/m^2 (/rho0=1000) as sin(pi y / (Ly/2)),
    mu0=100 m^2/s, mb0=0, eta white noise sigma=1 cm in the element-mean
    mode (seed 2002)."""
    lx, ly = 4000e3, 2000e3
    x, y, tri = jittered_mesh(nx, ny, lx, ly, 0.3, seed)
    ne = tri.shape[1]
    rng = np.random.default_rng(seed + 1)
    eta = np.zeros((LDM, ne), order="F")
    eta[0] = rng.normal(0.0, 0.01, ne)
    tau = 0.1 / 1000.0
    c = make_case(tables, x, y, tri, f0=1e-4, beta=2e-11, g0=9.81, H=1000.0, mu0=100.0,
                  mb0=0.0, wind=lambda X, Y: tau * np.sin(math.pi * Y / (ly / 2)), eta=eta)
    c["dt"] = cfl_dt_altitude(x, y, tri, tables["nc"], 1000.0, 9.81)
    c["variant"], c["nsteps"] = "fs", 1000
    return c


def random_case(tables: dict, nx=9, ny=9, seed=0, jitter=0.25, scale=0.1) -> dict:
    """Small jittered mesh with random modal states and forcing, for parity
    tests of single right-hand-side evaluations."""
    nm = tables["nm"]
    L = 1000e3
    x, y, tri = jittered_mesh(nx, ny, L, L, jitter, seed)
    rng = np.random.default_rng(seed + 7)
    ne = tri.shape[1]

    def rnd(s, mean=0.0):
        a = np.zeros((LDM, ne), order="F")
        a[:nm] = rng.standard_normal((nm, ne)) * s
        a[0] += mean
        return a

    c = make_case(tables, x, y, tri, f0=1e-4, beta=2e-11, g0=9.81, H=1000.0, mu0=100.0,
                  mb0=1e-4)
    c["phys"]["hb"] = rnd(5.0, 1000.0)
    c["phys"]["w0"] = rnd(1e-4)
    c["phys"]["z0"] = rnd(1e-4)
    c["state"] = (rnd(scale), rnd(scale), rnd(scale))
    return c
