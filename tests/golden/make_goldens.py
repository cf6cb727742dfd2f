"""Write the golden vectors under tests/golden/ from the reference itself.

Runs only in a container that holds the reference sources: it builds the
reference Fortran from its own sources (oracle/Makefile -> oracle/_ref) and
calls it through the ctypes harness.  The outputs are data (inputs and the
reference's outputs), committed so that the CPU test suite can pin the
oracle without the reference tree.

    python tests/golden/make_goldens.py [/root/reference]

Fixtures:
  quadrature.npz       triangle rules (ng = 6, 12, 25, 27, 33 via the
                       reference reader src/readgauss.f; 42 parsed from the
                       same data file, ng > ng_max) and Gauss-Legendre rules
                       (ngf = 3..9)
  tables_p<p>.npz      operator tables of the driver set-up (src/shwwnl.f:127-156)
  mesh_c1.npz          config-1 mesh (RMAL 17x17 + CM) and its pointers
  mesh_jit.npz         jittered 11x9 mesh after the reference CM tool, pointers
  calvar_p<p>.npz      one cal_var per variant on seeded random states
  rk4_p3.npz           10 RK4 steps per variant from the config-1 initial state,
                       with energy/mass before and after
  calconv_p<p>.npz     checkconv indicators and cal_conv decisions
                       (src/adapt.f:1122-1332) on seeded random states, for
                       threshold sets that hit every branch
  interp_p<p>.npz      interpol onto a 4-split refinement (some elements kept)
                       and onto a larger jittered mesh; calh2 of a test function
                       (src/adapt.f:1662-1824, 303-373)
  calvar_p4.npz,       one cal_var per variant at the shapes of configurations
  calvar_p5.npz        5 (p=4, 25-point rule) and 4 (p=5, 25-point rule)
  amr_cycles.npz       five adaptations each of the config-1 and jittered meshes
                       through the refinement hierarchy (src/adapt.f:380-1258):
                       flags, new meshes, tabsol, tinter
  rk4_cfg.npz          10 RK4 steps with energy/mass before and after, from the
                       initial states of the configuration generators on reduced
                       meshes: c2 (fs, p=6, 1152 triangles), c4 (inv, p=5, 576),
                       c5 (fs, p=4, one member of about 1100)

    python tests/golden/make_goldens.py [/root/reference] [--only-calconv|--only-interp|
                                                           --only-configs|--only-amr]
"""
from __future__ import annotations

import os
import sys
import threading

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from spoc_amd import configs, mesh  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
LDM = 45


def run_big_stack(fn, *a, **k):
    """The reference keeps large work arrays on the stack (its nonlinear
    variants overflow the default 8 MB): run it on a thread with a big stack."""
    res = {}

    def tgt():
        res["v"] = fn(*a, **k)

    threading.stack_size(512 * 1024 * 1024)
    t = threading.Thread(target=tgt)
    t.start()
    t.join()
    threading.stack_size(0)
    return res["v"]


def quad_tables(refroot):
    R = oracle.RefLib("lin")
    q = {}
    for ng in (6, 12, 25, 27, 33):
        w, x, y, wf, tf = R.read_gauss(3, ng, 3, refroot)
        q[f"tri{ng}"] = np.stack([w, x, y])
    raw = np.loadtxt(os.path.join(refroot, "include", "tri2_42.dat"))
    q["tri42"] = np.ascontiguousarray(raw.T)
    # the text parser must agree with the reference reader on the other files
    for ng in (12, 33):
        alt = np.loadtxt(os.path.join(refroot, "include", f"tri2_{ng}.dat")).T
        assert np.array_equal(alt, q[f"tri{ng}"]), ng
    for ngf in range(3, 10):
        _, _, _, wf, tf = R.read_gauss(3, 6, ngf, refroot)
        q[f"legen{ngf}"] = np.stack([wf, tf])
    np.savez(os.path.join(OUT, "quadrature.npz"), **q)
    return q


def rule(q, nc, ng, ngf):
    t = q[f"tri{ng}"]
    g = q[f"legen{ngf}"]
    return t[0], t[1], t[2], g[0], g[1]


def calconv_goldens(q):
    """cal_conv on random states of random_case meshes.  Threshold sets:
    'mid' puts fac1 and fac2 exactly on indicator values (the tests are
    strict) and fac3 above the maximum; 'rollback' puts fac3 below it with
    no refinement or coarsening; 'flat' zeroes h, whose indicator then
    vanishes (max |h(1,e)| <= 1e-6)."""
    A = oracle.RefAdapt()
    for nc, ng, ngf, nx, ny in ((3, 12, 4, 17, 17), (6, 33, 7, 9, 9)):
        tb = _tables_for(q, nc, ng, ngf)
        nm = tb["nm"]
        c = configs.random_case(tb, nx, ny, seed=31 + nc, scale=0.01)
        m = c["mesh"]
        u0, v0, h0 = (a[:nm] for a in c["state"])
        pico = tb["pico"]
        amax = np.array([A.checkconv(nc, ng, ngf, m, pico, a) for a in (u0, v0, h0)])
        top = np.sort(amax.max(axis=0))
        sup = float(amax.max())
        sets = {"mid": (float(top[int(0.7 * len(top))]), float(top[int(0.3 * len(top))]),
                        2.0 * sup, (u0, v0, h0)),
                "rollback": (10.0 * sup, 0.0, 0.5 * sup, (u0, v0, h0)),
                "flat": (float(top[len(top) // 2]), float(np.sort(amax[:2].min(axis=0))[len(top) // 2]),
                         2.0 * sup, (u0, v0, np.zeros_like(h0)))}
        out = dict(nx=nx, ny=ny, seed=31 + nc, u0=u0, v0=v0, h0=h0, amax=amax)
        for name, (f1, f2, f3, st) in sets.items():
            r = A.cal_conv(nc, ng, ngf, m, pico, *st, f1, f2, f3)
            out[f"{name}_fac"] = np.array([f1, f2, f3])
            for k in ("fconv", "fderaff"):
                out[f"{name}_{k}"] = r[k]
            out[f"{name}_flags"] = np.array([r["conv"], r["deraff"], r["timedir"]], np.int32)
            if name == "flat":
                out["flat_amax_h"] = A.checkconv(nc, ng, ngf, m, pico, st[2])
        np.savez_compressed(os.path.join(OUT, f"calconv_p{nc}.npz"), **out)


def refine4(x, y, tri, keep):
    """Split every triangle not in keep into 4 through its edge midpoints
    (new nodes appended, shared midpoints once); kept triangles stay.
    Returns x, y, the new triangles and tinter (old index of a kept
    triangle, else 0)."""
    x, y = list(x), list(y)
    mid = {}

    def m(a, b):
        k = (min(a, b), max(a, b))
        if k not in mid:
            x.append(0.5 * (x[a - 1] + x[b - 1]))
            y.append(0.5 * (y[a - 1] + y[b - 1]))
            mid[k] = len(x)
        return mid[k]

    out, tinter = [], []
    for e in range(tri.shape[1]):
        a, b, c = (int(v) for v in tri[:, e])
        if e in keep:
            out.append((a, b, c))
            tinter.append(e + 1)
            continue
        ab, bc, ca = m(a, b), m(b, c), m(c, a)
        for t in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)):
            out.append(t)
            tinter.append(0)
    return (np.array(x), np.array(y), np.asfortranarray(np.array(out, np.int32).T),
            np.array(tinter, np.int32))


def interp_goldens(q):
    """interpol and calh2 (src/adapt.f:1662-1824, 303-373) of random fields of
    a random_case mesh onto (a) its 4-split refinement with every 5th
    element kept and (b) a different jittered mesh of a slightly larger
    square, whose quadrature points outside the old mesh take the wall
    projection of interlin."""
    for nc, ng, ngf, nx, ny in ((3, 12, 4, 9, 9), (6, 33, 7, 7, 7)):
        tb = _tables_for(q, nc, ng, ngf)
        nm = tb["nm"]
        c = configs.random_case(tb, nx, ny, seed=51 + nc, scale=0.01)
        m = c["mesh"]
        A = oracle.RefAdapt()
        A.setup(nc, *rule(q, nc, ng, ngf))
        fields = [c["state"][0][:nm], c["state"][2][:nm], c["phys"]["hb"][:nm]]
        keep = set(range(0, m["ne"], 5))
        xa, ya, ta, tia = refine4(m["x"], m["y"], m["in"], keep)
        xb, yb, tb_ = configs.jittered_mesh(nx + 2, ny + 1, 1.02e6, 1.03e6, 0.25, 90 + nc)
        xb, yb = xb - 0.01e6, yb - 0.015e6
        tib = np.zeros(tb_.shape[1], np.int32)
        out = dict(nx=nx, ny=ny, seed=51 + nc, xa=xa, ya=ya, ta=ta, tia=tia, xb=xb, yb=yb, tb=tb_)
        for f, u in enumerate(fields):
            out[f"u{f}"] = u
            out[f"a{f}"] = A.interpol(m, xa, ya, ta, tia, u)[:nm]
            out[f"b{f}"] = A.interpol(m, xb, yb, tb_, tib, u)[:nm]
        out["calh2"] = A.calh2(xa, ya, ta, tia, fields[2])[:nm]
        np.savez_compressed(os.path.join(OUT, f"interp_p{nc}.npz"), **out)


def _safe_families(amr, m, ne):
    """Families (four sibling leaves) that can be coarsened without leaving
    a triangle next to one two levels finer: hierarchy numbers of the four
    children, all <= ne (the reference reads the per-element coarsening
    flags by hierarchy number), with no current neighbour finer than them."""
    _, _, tri, tabsol = amr.mesh()
    lev = amr.levels()
    pos = {int(h): e for e, h in enumerate(tabsol)}
    nbr = [[] for _ in range(ne)]
    for f in range(m["nf"]):
        a, b = int(m["tseg"][0, f]) - 1, int(m["tseg"][1, f]) - 1
        if b >= 0:
            nbr[a].append(b)
            nbr[b].append(a)
    fams = []
    for h in sorted(pos):
        if h % 4 != 1 or h <= 4:
            continue
        kids = [h, h + 1, h + 2, h + 3]
        if any(k not in pos or k > ne for k in kids):
            continue
        es = [pos[k] for k in kids]
        if len({int(lev[e]) for e in es}) != 1:
            continue
        if all(lev[n] <= lev[es[0]] for e in es for n in nbr[e]):
            fams.append(kids)
    return fams


def amr_goldens():
    """The refinement hierarchy (init_adap, deraffiner, adapmaille,
    protridouble, elimpptri, inter_only_new; src/adapt.f:380-1258) over
    several adaptations of two meshes: seeded refinement flags (scattered
    and clustered, so that the closure and the green halves run at several
    levels), and coarsening flags on whole families of leaves.  Each cycle
    stores its flags and the reference's new mesh, tabsol and tinter."""
    import spoc_amd
    A = oracle.RefAdapt()
    out = {}
    d1 = np.load(os.path.join(OUT, "mesh_c1.npz"))
    dj = np.load(os.path.join(OUT, "mesh_jit.npz"))
    cases = {"c1": (d1["x"], d1["y"], d1["tri"], (0.03, 0.03, 0.02, 0.02, 0.02), (150e3, 0.0)),
             "jit": (dj["x"], dj["y"], dj["tri"], (0.05, 0.08, 0.06, 0.04, 0.05), (150e3, 0.0))}
    for name, (x, y, tri, rates, (rad, _)) in cases.items():
        rng = np.random.default_rng(61 if name == "c1" else 62)
        m = mesh.build(x, y, tri, 3)
        A.adap_init(x, y, tri, m["tseg"], m["itm"])
        ours = spoc_amd.Amr(m)  # (only to choose coarsening flags the reference can take)
        out[f"{name}_x"], out[f"{name}_y"], out[f"{name}_tri"] = x, y, tri
        out[f"{name}_tseg"], out[f"{name}_itm"] = m["tseg"], m["itm"]
        cx, cy, ct = x, y, tri
        cxm, cym = x.mean(), y.mean()
        for cyc, rate in enumerate(rates):
            ne = ct.shape[1]
            cen_x = cx[ct - 1].mean(axis=0)
            cen_y = cy[ct - 1].mean(axis=0)
            # (the cluster halves at each cycle: the reference's tables hold
            # 1,600 nodes and 1,600 triangles, include/maille.inc:6-7)
            near = (cen_x - cxm) ** 2 + (cen_y - cym) ** 2 < (rad / 2 ** cyc) ** 2
            fconv = ((rng.random(ne) < rate) | (near & (cyc % 2 == 0))).astype(np.int32)
            fder = np.zeros(ne, np.int32)
            if cyc == 3:
                fams = _safe_families(ours, mesh.build(cx, cy, ct, 3), ne)
                for kids in fams[:: max(1, len(fams) // 4)][:4]:
                    fder[np.array(kids) - 1] = 1
                fconv[np.array([k - 1 for kids in fams for k in kids], dtype=int)] = 0
            rx, ry, rt, rtab = A.adap_cycle(len(cx), fconv, fder)
            ours.adapt(fconv, fder)
            tin = A.inter_only_new(ct, rt)
            out[f"{name}_{cyc}_fconv"], out[f"{name}_{cyc}_fderaff"] = fconv, fder
            out[f"{name}_{cyc}_x"], out[f"{name}_{cyc}_y"] = rx, ry
            out[f"{name}_{cyc}_tri"], out[f"{name}_{cyc}_tabsol"] = rt, rtab
            out[f"{name}_{cyc}_tinter"] = tin
            print(f"amr {name} cycle {cyc}: refine {int(fconv.sum())} coarsen {int(fder.sum())} "
                  f"-> nn {len(rx)} ne {rt.shape[1]} (kept {int((tin > 0).sum())})")
            cx, cy, ct = rx, ry, rt
        out[f"{name}_cycles"] = len(rates)
    np.savez_compressed(os.path.join(OUT, "amr_cycles.npz"), **out)


def main(refroot="/root/reference", *flags):
    oracle.build(ref=True)
    q = quad_tables(refroot)
    if "--only-calconv" in flags:
        calconv_goldens(q)
        print("cal_conv goldens written to", OUT)
        return
    if "--only-interp" in flags:
        interp_goldens(q)
        print("interpolation goldens written to", OUT)
        return
    if "--only-amr" in flags:
        amr_goldens()
        print("refinement goldens written to", OUT)
        return
    if "--only-configs" in flags:
        config_goldens(q)
        print("configuration goldens written to", OUT)
        return
    for nc, ng, ngf in ((3, 12, 4), (6, 33, 7), (4, 25, 5), (5, 25, 6)):
        R = oracle.RefLib("fs")
        i1, i2 = R.setup(nc, *rule(q, nc, ng, ngf))
        t = R.tables()
        np.savez_compressed(os.path.join(OUT, f"tables_p{nc}_{ng}.npz"), info=np.array([i1, i2]),
                            **t)
    # meshes through the reference CM tool and pointer routines
    cm = oracle.RefCM()
    x, y, tri = mesh.rmal(17, 17, -500e3, 500e3, -500e3, 500e3)
    tri_c1 = cm.changemaille(len(x), tri)
    R = oracle.RefLib("lin")
    R.setup(3, *rule(q, 3, 12, 4))
    m1 = R.mesh(x, y, tri_c1, 3, 1e-4, 2e-11)
    np.savez_compressed(os.path.join(OUT, "mesh_c1.npz"), x=x, y=y, tri_in=tri, tri=tri_c1,
                        **{k: v for k, v in m1.items() if k != "nf"}, nf=m1["nf"])
    # jittered mesh: our generator's raw triangles through the reference CM
    xj, yj, trij_ours = configs.jittered_mesh(11, 9, 1000e3, 800e3, 0.3, 5)
    # rebuild the pre-CM triangles the generator produced (same seed), for CM parity
    rng_tri = _raw_jitter_tris(11, 9, 1000e3, 800e3, 0.3, 5)
    trij = cm.changemaille(len(xj), rng_tri)
    mj = R.mesh(xj, yj, trij, 3, 1e-4, 2e-11)
    np.savez_compressed(os.path.join(OUT, "mesh_jit.npz"), x=xj, y=yj, tri_in=rng_tri, tri=trij,
                        **{k: v for k, v in mj.items() if k != "nf"}, nf=mj["nf"])
    # single cal_var on seeded random states
    calvar_goldens(q, 3, 12, 4, 17, 17)
    calvar_goldens(q, 6, 33, 7, 9, 9)
    # 10 RK4 steps from the config-1 initial state
    tb = _tables_for(q, 3, 12, 4)
    c = configs.config1(tb)
    out = {}
    for var in ("lin", "fs", "ns", "inv"):
        Rv = oracle.RefLib(var)
        Rv.setup(3, *rule(q, 3, 12, 4))
        m = Rv.mesh(c["mesh"]["x"], c["mesh"]["y"], c["mesh"]["in"], 3, 1e-4, 0.0)
        p = dict(c["phys"])
        p["fc"] = m["fc"]
        p["mu0"] = 100.0
        m["in"], m["x"], m["y"] = c["mesh"]["in"], c["mesh"]["x"], c["mesh"]["y"]
        e0 = Rv.energy(m, p, *c["state"])
        r = run_big_stack(Rv.rk4, m, p, *c["state"], c["dt"], 10)
        e1 = Rv.energy(m, p, *r)
        for k, a in zip("uvh", r):
            out[f"{var}_{k}"] = a[:10]
        out[f"{var}_energy"] = np.array([e0, e1])
    np.savez_compressed(os.path.join(OUT, "rk4_p3.npz"), dt=c["dt"], mu0=100.0,
                        u0=c["state"][0][:10], v0=c["state"][1][:10], h0=c["state"][2][:10], **out)
    calconv_goldens(q)
    interp_goldens(q)
    config_goldens(q)
    amr_goldens()
    print("goldens written to", OUT)


def calvar_goldens(q, nc, ng, ngf, nx, ny):
    """One cal_var per variant on a seeded random_case state (the reference's
    own mesh pointers, its fc)."""
    out = {}
    nm = (nc + 1) * (nc + 2) // 2
    for var in ("lin", "fs", "ns", "inv"):
        Rv = oracle.RefLib(var)
        Rv.setup(nc, *rule(q, nc, ng, ngf))
        c = configs.random_case(_tables_for(q, nc, ng, ngf), nx, ny, seed=11 + nc)
        m = Rv.mesh(c["mesh"]["x"], c["mesh"]["y"], c["mesh"]["in"], nc, 1e-4, 2e-11)
        p = c["phys"]
        p["fc"] = m["fc"]
        u1 = run_big_stack(Rv.cal_var, m, p, *c["state"])
        for k, a in zip("uvh", u1):
            out[f"{var}_{k}1"] = a[:nm]
    st = c["state"]
    np.savez_compressed(os.path.join(OUT, f"calvar_p{nc}.npz"), nx=nx, ny=ny, seed=11 + nc,
                        u0=st[0][:nm], v0=st[1][:nm], h0=st[2][:nm], hb=p["hb"][:nm],
                        w0=p["w0"][:nm], z0=p["z0"][:nm], fc=p["fc"][:nm], **out)


# reduced meshes of the configuration generators (the reference holds at
# most 1600 triangles): name -> (p, ng, ngf, generator, variant)
CFG_CASES = {
    "c2": (6, 33, 7, lambda tb: configs.config2(tb, nx=25, ny=25), "fs"),
    "c4": (5, 25, 6, lambda tb: configs.config4(tb, nx=25, ny=13), "inv"),
    "c5": (4, 25, 5, lambda tb: configs.config5(tb, nmem=1, ntri=(1100, 1100))[0], "fs"),
}


def config_goldens(q):
    calvar_goldens(q, 4, 25, 5, 12, 10)
    calvar_goldens(q, 5, 25, 6, 12, 10)
    out = {}
    for name, (nc, ng, ngf, gen, var) in CFG_CASES.items():
        tb = _tables_for(q, nc, ng, ngf)
        c = gen(tb)
        nm = tb["nm"]
        Rv = oracle.RefLib(var)
        Rv.setup(nc, *rule(q, nc, ng, ngf))
        m = Rv.mesh(c["mesh"]["x"], c["mesh"]["y"], c["mesh"]["in"], nc, 1e-4, 2e-11)
        p = dict(c["phys"])
        p["fc"] = m["fc"]
        m["in"], m["x"], m["y"] = c["mesh"]["in"], c["mesh"]["x"], c["mesh"]["y"]
        e0 = run_big_stack(Rv.energy, m, p, *c["state"])
        r = run_big_stack(Rv.rk4, m, p, *c["state"], c["dt"], 10)
        e1 = run_big_stack(Rv.energy, m, p, *r)
        out[f"{name}_dt"] = c["dt"]
        out[f"{name}_ne"] = c["mesh"]["ne"]
        for k, a in zip("uvh", c["state"]):
            out[f"{name}_{k}0"] = a[:nm]
        for k, a in zip("uvh", r):
            out[f"{name}_{k}"] = a[:nm]
        out[f"{name}_energy"] = np.array([e0, e1])
    np.savez_compressed(os.path.join(OUT, "rk4_cfg.npz"), **out)


def _tables_for(q, nc, ng, ngf):
    from spoc_amd import nm_of
    O = oracle.Oracle()
    O.setup(nc, *rule(q, nc, ng, ngf))
    t = O.tables()
    w, x, y, wf, tf = rule(q, nc, ng, ngf)
    t.update(nc=nc, ng=ng, ngf=ngf, nm=nm_of(nc), wi=w, xi=x, yi=y, wif=wf, tif=tf)
    return t


def _raw_jitter_tris(nx, ny, lx, ly, jitter, seed):
    """The triangles of configs.jittered_mesh before its CM step."""
    saved = mesh.changemaille
    try:
        mesh.changemaille = lambda nn, tri: np.array(tri, dtype=np.int32, order="F")
        configs.M.changemaille = mesh.changemaille
        _, _, tri = configs.jittered_mesh(nx, ny, lx, ly, jitter, seed)
    finally:
        mesh.changemaille = saved
        configs.M.changemaille = saved
    return tri


if __name__ == "__main__":
    main(*sys.argv[1:])
