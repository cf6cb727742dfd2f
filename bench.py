#!/usr/bin/env python3
"""Benchmark: element-DOF RK4 stage updates/s of the SPOC right-hand side on
one MI355X, FP32-MFMA path against the plain FP64 GPU path.

Workload (default): nonlinear free-slip shallow water (cal_var of
src/calvarnl_fs.f) at p=6 (28 modes, 33-point triangle rule, 7 edge points)
on a jittered 448 x 224 node lattice over a 4000 x 2000 km basin (199,362
triangles), beta plane f0=1e-4 s^-1, beta=2e-11 m^-1 s^-1, H=1000 m,
double-gyre wind stress 0.1 N/m^2, mu0=100 m^2/s, eta white noise 1 cm.
Synthetic data, seeded.

Timed region: K RK4 steps of the device-resident state (inputs already in
HBM; the step replayed from a HIP graph), bracketed by device
synchronisation.  A step is 4 right-hand sides; element-DOF stage updates
per step = 4 * ne * nm * 3.  Per-kernel times for the roofline come from a
separate run of min(K, 20) steps with HIP events around each launch.

--workload selects one of the other configurations (c1 linear basin, c3
no-slip eddies at p=7, c4 inviscid at p=5, c5 ensemble of 64 members at
p=4); the default, c2, is the headline.

Prints ONE JSON line.  It runs on one GPU; --gpus above 1 is refused.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12        # B/s, MI355X HBM3E
FP32_PEAK = 157.3e12     # FLOP/s, dense FP32 MFMA
FP64_PEAK = 78.6e12      # FLOP/s, FP64 vector/MFMA


def quad_rule(ng, ngf):
    q = np.load(os.path.join(ROOT, "tests", "golden", "quadrature.npz"), allow_pickle=False)
    t, g = q[f"tri{ng}"], q[f"legen{ngf}"]
    return t[0], t[1], t[2], g[0], g[1]


METRIC = "element-DOF RK4 stage updates/s, FP32-MFMA vs FP64 GPU at \u03b7 rel-L2 \u22641e-6"

WORKLOADS = {
    "c2": dict(shape=(6, 33, 7), steps=100, gate=1000, name=lambda nc, ng, ngf, ne, dt, a: (
        f"calvarnl_fs p={nc} ng={ng} ngf={ngf} on {ne} triangles (jittered {a.nx}x{a.ny} nodes, "
        f"4000x2000 km, beta plane, double-gyre wind, mu0=100), dt={dt} s")),
    "c1": dict(shape=(3, 12, 4), steps=1000, gate=1000, name=lambda nc, ng, ngf, ne, dt, a: (
        f"calvar_lin p={nc} ng={ng} ngf={ngf} on {ne} triangles (1000 km square, 16x16 "
        f"squares), H=4000, 1 m Gaussian eta, dt={dt} s")),
    "c3": dict(shape=(7, 42, 8), steps=50, gate=2000, name=lambda nc, ng, ngf, ne, dt, a: (
        f"calvarnl_ns p={nc} ng={ng} ngf={ngf} on {ne} triangles (4000x2000 km), 20 seamounts, "
        f"k^-3 eddies RMS 0.3 m/s, nu=100, dt={dt} s")),
    "c4": dict(shape=(5, 25, 6), steps=100, gate=500, name=lambda nc, ng, ngf, ne, dt, a: (
        f"calvarnl_inv p={nc} ng={ng} ngf={ngf} on {ne} triangles (4000x2000 km), mu0=0, "
        f"1 m Gaussian eta, dt={dt} s")),
    "c5": dict(shape=(4, 25, 5), steps=50, gate=500, name=lambda nc, ng, ngf, ne, dt, a: (
        f"ensemble of {a.members} calvarnl_fs members, p={nc} ng={ng} ngf={ngf}, meshes refined "
        f"by the adaptive hierarchy around each member's eta bump to 5k-60k triangles, {ne} in "
        f"total, one union mesh per launch, dt={dt} s")),
}


def algorithmic_counts(nm, ng, ngf, variant, steps=1):
    """Useful (unpadded) FP32 and FP64 flops and compulsory HBM bytes per
    element per launch of each kernel, from its product shapes and the
    arrays it must read and write once (neighbour traces counted once, RK4
    traffic averaged over the four stages); a per-stage kernel's launch is
    one right-hand side, the persistent kernel's is `steps` RK4 steps."""
    e3 = 3 * ngf
    f4, f8 = 4, 8
    mesh = 16 * f8 + 6 * f4 + f4                       # geo64, nbr/nbe, bdl
    # RK4, averaged over the 4 stages: c0 read at every stage and the three
    # stored stage inputs Y2, Y3, Y4 at the last; one of them written at
    # stages 0-2, c0 at 3
    rk = 3 * nm * f8 * (1.0 + 0.75) + 3 * nm * f8 * (0.75 + 0.25)
    vol_fl = 2 * 4 * nm * nm                              # core: -g grad eta, div(hh0 u)
    face_fl = 2 * 6 * nm * e3                             # core: lifts, traces
    trl_by = 2 * e3 * f8 + rk + 2 * e3 * f8               # trL in, RK, trL out
    if variant == "lin":
        # (the stage input is c0 at stage 0 and Y4 at stage 3, read for the
        # RK4 update anyway)
        # persistent: c0 read and written once, the mesh read once, and per
        # stage only the traces (neighbours' in, own out); the rest of the
        # state stays in registers
        return {"lin_f64": (0, vol_fl + face_fl, 0.5 * 3 * nm * f8 + trl_by + mesh),
                "lin_persistent": (0, 4 * steps * (vol_fl + face_fl),
                                   2 * 3 * nm * f8 + mesh + 4 * steps * 4 * e3 * f8)}
    # B+C (FP32 only): q (Dx, Dy, LF), traces of F and of the mass flux (TR),
    # values at and projection from the quadrature points, the volume
    # divergence (6 products)
    bc_fl = 2 * (4 * nm * nm + 10 * nm * e3 + 11 * nm * ng + 6 * nm * nm)
    bc_by = (3 * nm * f8 + 2 * e3 * f4 + 16 * f4 + 6 * f4 + f4 + 4 * ng * f4 + 3 * f4
             + 3 * nm * f4 + 3 * e3 * f4)      # x, trA, geo, nbr/nbe, bdl, fields | gvol, trS
    # E+G: the FP64 core (volume and face terms; the flux traces' face
    # values are lifted by the same product) from the stage input, which is
    # c0 at stage 0 and Y4 at stage 3 (read for the RK4 update anyway: 2 of 4
    # stages count)
    eg_fl = 0
    eg_by = (0.5 * 3 * nm * f8 + 3 * nm * f4 + 3 * e3 * f4 + mesh + f8 + trl_by
             + 2 * e3 * f4)
    return {"bc_f32": (bc_fl, 0, bc_by), "eg_f32f64": (eg_fl, vol_fl + face_fl, eg_by)}


def measured_traffic(kernel, nc, ne):
    """HBM bytes per launch of `kernel` (averaged over the RK4 stages) from
    the committed counter profiles: FETCH_SIZE and WRITE_SIZE, each scaled
    by what it reads for a device copy of known size.  None when no profile
    of this degree and mesh size is committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None
    return d.get("workloads", {}).get(f"p{nc}_ne{ne}", {}).get(kernel)


def adjoint_check(ctx, t, c, nsteps, ndir, seed=4001):
    """Gradient of J = integral of eta^2 over the elements whose centroid is
    within 200 km of the centre, after nsteps, with respect to the initial
    eta and to hb: adjoint time (automatic checkpointing) and relative
    agreement with central differences of the FP64 path (step 1e-6 relative
    to the direction's scale) along ndir seeded random directions.  With
    ndir = 0 only the adjoint is timed, and J and its seed come from the
    mixed path's forward run (no FP64 forward runs)."""
    m, p = c["mesh"], c["phys"]
    nm, ne, variant, dt = t["nm"], m["ne"], c["variant"], c["dt"]
    M = t["ammo"][:nm, :nm]
    det = 1.0 / m["vect"][4]
    tri = np.asarray(m["in"]) - 1
    cx, cy = m["x"][tri].mean(axis=0), m["y"][tri].mean(axis=0)
    inside = (cx ** 2 + cy ** 2) < 200e3 ** 2
    x0 = [np.asfortranarray(a) for a in c["state"]]

    def fwd(x, hb, prec="fp64"):
        ctx.set_fields(hb, p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
        ctx.upload(*x)
        ctx.rk4(variant, dt, nsteps, precision=prec)
        h = ctx.download()[2][:nm]
        return h, float(np.sum(det[inside] * np.einsum("ie,ij,je->e", h[:, inside], M,
                                                        h[:, inside])))

    h, J0 = fwd(x0, p["hb"], "fp64" if ndir else "fp32")
    lam = np.zeros((x0[0].shape[0], ne), order="F")
    lam[:nm, inside] = 2.0 * det[inside] * (M @ h[:, inside])
    zero = np.zeros_like(lam)
    ctx.upload(*x0)
    ctx.adjoint_upload(zero, zero, lam)
    ctx.set_checkpoint(0)  # every stage input kept when it fits, else checkpoints every 50
    ctx.sync()
    t0 = time.perf_counter()
    ctx.rk4_adjoint(variant, dt, nsteps)
    ctx.sync()
    adj_s = time.perf_counter() - t0
    g_eta = ctx.adjoint_download()[2]
    g_hb = ctx.adjoint_hb_download()
    rng = np.random.default_rng(seed)
    errs = {"eta0": [], "hb": []}
    for _ in range(ndir):
        for key, scale in (("eta0", 0.01), ("hb", 1.0)):
            v = np.zeros_like(lam)
            v[:nm] = rng.standard_normal((nm, ne)) * scale
            eps = 1e-6 / scale * max(1.0, float(np.abs(x0[2][:nm]).max()) if key == "eta0"
                                     else float(np.abs(p["hb"][:nm]).max()))
            if key == "eta0":
                jp = fwd([x0[0], x0[1], x0[2] + eps * v], p["hb"])[1]
                jm = fwd([x0[0], x0[1], x0[2] - eps * v], p["hb"])[1]
                ad = float(np.sum(g_eta[:nm] * v[:nm]))
            else:
                jp = fwd(x0, p["hb"] + eps * v)[1]
                jm = fwd(x0, p["hb"] - eps * v)[1]
                ad = float(np.sum(g_hb[:nm] * v[:nm]))
            fd = (jp - jm) / (2 * eps)
            errs[key].append(abs(fd - ad) / max(abs(fd), 1e-300))
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
    return {"J": J0, "steps": nsteps, "adjoint_s": adj_s,
            "adjoint_ms_per_step": 1e3 * adj_s / max(nsteps, 1),
            "rel_err_vs_central_differences": {k: max(v) if v else None for k, v in errs.items()},
            "directions": ndir, "checkpoints": "automatic (all stage inputs kept when they fit "
                                                "in half the free device memory, else every 50 steps)"}


def calib_copy(nbytes):
    """One device-to-device copy of nbytes (reads and writes nbytes)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(nbytes)
    ok = hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    ok = ok and hip.hipMemset(a, 1, n) == 0 and hip.hipDeviceSynchronize() == 0
    ok = ok and hip.hipMemcpy(b, a, n, 3) == 0 and hip.hipDeviceSynchronize() == 0  # DtoD
    hip.hipFree(a)
    hip.hipFree(b)
    if not ok:
        raise RuntimeError("calibration copy failed")


def adapt_timing(ctx, state, ne, nc, ng, ngf, cpu, reps=20, fac=(0.1, 0.01, 1.0)):
    """The adaptive driver's refinement indicator and decisions (cal_conv,
    src/adapt.f:1122-1332) on the given state: wall time per call of
    spoc_cal_conv, flags copied to the host; with cpu, the reference
    Fortran's indicator (oracle/_ref) timed on the c2 generator's
    1152-triangle mesh, one core, scaled per element."""
    ctx.upload(*state)
    ctx.cal_conv(*fac)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = ctx.cal_conv(*fac, with_amax=False)
    dev = (time.perf_counter() - t0) / reps
    out = {"us_per_call": dev * 1e6, "ne": ne, "calls": reps,
           "decisions": {k: int(r[k]) for k in ("conv", "deraff", "timedir")} if ctx.nmem == 1
           else None, "fac": list(fac)}
    if not cpu:
        return out
    import oracle
    from spoc_amd import configs
    if not oracle.RefAdapt.available() or ng > 33:
        return out
    O = oracle.Oracle()
    O.setup(nc, *quad_rule(ng, ngf))
    tabs = O.tables()
    w, x, y, wf, tf = quad_rule(ng, ngf)
    tabs.update(nc=nc, ng=ng, ngf=ngf, nm=O.nm, wi=w, xi=x, yi=y, wif=wf, tif=tf)
    c = configs.config2(tabs, nx=25, ny=25)
    A = oracle.RefAdapt()
    m, st = c["mesh"], [a[:O.nm] for a in c["state"]]
    # (the three checkconv calls are cal_conv's work; cal_conv itself also
    # prints its decisions to stdout, which would mix with this JSON line)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        for a in st:
            A.checkconv(nc, ng, ngf, m, tabs["pico"], a)
        n += 1
    per_elem = (time.perf_counter() - t0) / n / m["ne"]
    out["cpu_reference"] = {"us_per_call_at_this_ne": per_elem * ne * 1e6, "cores": 1,
                            "sample": f"reference checkconv of u, v, h, p={nc} on {m['ne']} "
                                      f"triangles, {n} calls, scaled per element"}
    return out


def renumber(c, nc, how, seed=17):
    """The case with the caller's element numbering changed: "shuffled"
    (a seeded random permutation) or "rows" (centroids sorted by row, then
    by x, as a structured mesh generator numbers them); the mesh operands
    are rebuilt from the permuted triangles (new face numbering too) and the
    per-element fields and state permuted with them."""
    from spoc_amd import mesh as M
    m = c["mesh"]
    if how == "shuffled":
        perm = np.random.default_rng(seed).permutation(m["ne"])
    else:
        tri = np.asarray(m["in"]) - 1
        cx, cy = m["x"][tri].mean(0), m["y"][tri].mean(0)
        h = np.sqrt((np.ptp(cx) * np.ptp(cy)) / m["ne"] * 2.0)  # lattice spacing
        perm = np.lexsort((cx, np.floor((cy - cy.min()) / h)))
    m2 = M.build(m["x"], m["y"], np.asfortranarray(m["in"][:, perm]), nc, 0.0, 0.0)
    p = {k: (np.asfortranarray(v[:, perm]) if isinstance(v, np.ndarray) and v.ndim == 2 else v)
         for k, v in c["phys"].items()}
    out = dict(c)
    out.update(mesh=m2, phys=p, state=tuple(np.asfortranarray(a[:, perm]) for a in c["state"]))
    return out


def graph_prime_steps(K):
    """Steps of an untimed call that captures the HIP graph a K-step call
    replays (8 steps per graph from K >= 8, one step per graph from K >= 2;
    a single step runs eagerly)."""
    return 8 if K >= 8 else (2 if K >= 2 else 0)


def warm(ctx, variant, dt, W, K, prec):
    """W untimed warm-up steps, then (if those captured a graph of another
    shape) one untimed call that captures the graph the timed K steps
    replay, so that no capture falls inside the timed region."""
    if W:
        ctx.rk4(variant, dt, W, precision=prec)
    want = graph_prime_steps(K)
    if want and (W < 2 or (W >= 8) != (K >= 8)):
        ctx.rk4(variant, dt, want, precision=prec)
        return want
    return 0


def integrals(ctx):
    """(ke, pe, mass) of the resident state, summed over ensemble members."""
    if ctx.nmem > 1:
        ke, pe, mass = ctx.ensemble_energy()
        return float(ke.sum()), float(pe.sum()), float(mass.sum())
    return ctx.energy()


def mass_rel_l2(t, vect, a, b):
    """Relative L2 error of a against b in the mass-matrix norm:
    sqrt(sum_e det_e d^T M d / sum_e det_e b^T M b)."""
    nm = t["nm"]
    ammo = t["ammo"][:nm, :nm]
    det = 1.0 / vect[4]
    d = a[:nm] - b[:nm]
    n1 = np.sum(det * np.einsum("ie,ij,je->e", d, ammo, d))
    n2 = np.sum(det * np.einsum("ie,ij,je->e", b[:nm], ammo, b[:nm]))
    return float(np.sqrt(n1 / n2))


def gate_leg(ctx, t, variant, dt, state0, vect, nsteps):
    """The metric's accuracy gate: nsteps untimed RK4 steps of both paths
    from the same initial state; relative L2 (mass norm) of the mixed path
    against the FP64 path per field; kinetic and potential energy and mass
    (src/energy.f:3-96) at the start and the end on each path, the drift of
    the total energy and of the mass over the run on each path, and how far
    the mixed path's end values are from the FP64 path's."""
    out = {"steps": nsteps}
    fin, ends = {}, {}
    for prec in ("fp32", "fp64"):
        ctx.upload(*state0)
        e0 = integrals(ctx)
        ctx.rk4(variant, dt, nsteps, precision=prec)
        e1 = integrals(ctx)
        fin[prec], ends[prec] = ctx.download(), e1
        E0, E1 = e0[0] + e0[1], e1[0] + e1[1]
        out[f"integrals_{prec}"] = {"start": dict(zip(("ke", "pe", "mass"), e0)),
                                    "end": dict(zip(("ke", "pe", "mass"), e1))}
        out[f"drift_{prec}"] = {"energy": (E1 - E0) / abs(E0) if E0 else E1 - E0,
                                "mass": (e1[2] - e0[2]) / abs(e0[2]) if e0[2] else e1[2] - e0[2]}
    out["end_fp32_vs_fp64"] = {k: abs(a - b) / abs(b) if b else abs(a - b)
                               for k, a, b in zip(("ke", "pe", "mass"), ends["fp32"],
                                                  ends["fp64"])}
    out["rel_l2"] = {f: mass_rel_l2(t, vect, fin["fp32"][i], fin["fp64"][i])
                     for i, f in enumerate(("u", "v", "eta"))}
    out["eta_le_1e-6"] = out["rel_l2"]["eta"] <= 1e-6
    return out


def dropin_timing(ctx, t, variant, dt, state0, nsteps=2):
    """The unchanged driver's loop (src/shwwnl.f:238-284): per step four
    spoc_rhs calls on host (45, ne) arrays, each copying the state in and
    the tendencies out over PCIe, and the stage updates on the host (numpy,
    in the driver's expression order).  Returns ms per step and per call
    for both paths, and the PCIe bytes each call moves."""
    nm = t["nm"]
    u0 = [np.array(a, order="F") for a in state0]
    ne = u0[0].shape[1]
    outs = [[np.zeros_like(u0[0], order="F") for _ in range(3)] for _ in range(4)]
    res = {}
    for prec in ("fp64", "fp32"):
        ctx.cal_var(variant, *u0, precision=prec, out=outs[0])  # warm
        t_rhs = 0.0
        t0 = time.perf_counter()
        for _ in range(nsteps):
            x = u0
            for s in range(4):
                a = time.perf_counter()
                ctx.cal_var(variant, *x, precision=prec, out=outs[s])
                t_rhs += time.perf_counter() - a
                if s < 3:
                    f = 0.5 if s < 2 else 1.0
                    x = [u0[i] + dt * outs[s][i] * f if f != 1.0 else u0[i] + dt * outs[s][i]
                         for i in range(3)]
            for i in range(3):
                u0[i] = u0[i] + (dt / 6.0) * (((outs[0][i] + 2.0 * outs[1][i]) +
                                               2.0 * outs[2][i]) + outs[3][i])
        wall = time.perf_counter() - t0
        per_call = t_rhs / (4 * nsteps)
        res[prec] = {"ms_per_step": 1e3 * wall / nsteps, "ms_per_call": 1e3 * per_call}
    nbytes = 2 * 3 * nm * ne * 8
    res["pcie_bytes_per_call"] = nbytes
    res["pcie_GBps_fp32"] = nbytes / (res["fp32"]["ms_per_call"] * 1e-3) / 1e9
    res["steps"] = nsteps
    return res


def host_info():
    """CPU model, host core count and the reference's compile flags."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "compiler": "flang -O2 -fPIC (oracle/Makefile; no -march, so no FMA contraction)"}


def _cpu_case(variant, nc, ng, ngf, wl):
    """The bounded CPU sample of workload wl: the reference Fortran
    (oracle/_ref) when it can hold the rule, else the C restatement.
    Returns (run(nsteps), ne, nm, kind)."""
    import oracle
    from spoc_amd import configs
    O = oracle.Oracle()
    O.setup(nc, *quad_rule(ng, ngf))
    tabs = O.tables()
    w, x, y, wf, tf = quad_rule(ng, ngf)
    tabs.update(nc=nc, ng=ng, ngf=ngf, nm=(nc + 1) * (nc + 2) // 2, wi=w, xi=x, yi=y, wif=wf,
                tif=tf)
    small = {"c1": lambda: configs.config1(tabs),
             "c2": lambda: configs.config2(tabs, nx=25, ny=25),
             "c3": lambda: configs.config3(tabs, nx=25, ny=13),
             "c4": lambda: configs.config4(tabs, nx=25, ny=13),
             "c5": lambda: configs.config5(tabs, nmem=1, ntri=(1100, 1100))[0]}
    c = small[wl]()
    m, p = c["mesh"], c["phys"]
    ne, nm = m["ne"], tabs["nm"]
    # the reference's static arrays hold at most 33 quadrature points
    if oracle.ref_available() and ng <= 33:
        try:
            R = oracle.RefLib(variant)
            R.setup(nc, *quad_rule(ng, ngf))
            return (lambda n: R.rk4(m, p, *c["state"], c["dt"], n)), ne, nm, "reference"
        except OSError:
            pass
    return (lambda n: O.rk4(variant, m, p, *c["state"], c["dt"], n)), ne, nm, "port"


def _timed(runner, n=None, budget_s=15.0):
    """Run the CPU sample (on a thread with a large stack: the reference's
    nonlinear variants keep big local arrays) for n steps, or for about
    budget_s seconds; returns (steps, seconds)."""
    res = {}

    def work():
        k = n
        if k is None:
            t0 = time.perf_counter()
            runner(1)
            one = time.perf_counter() - t0
            k = max(1, int(budget_s / max(one, 1e-3)))
        t0 = time.perf_counter()
        runner(k)
        res["n"], res["s"] = k, time.perf_counter() - t0

    old = threading.stack_size(512 * 1024 * 1024)
    th = threading.Thread(target=work)
    th.start()
    th.join()
    threading.stack_size(old)
    return res["n"], res["s"]


def cpu_baseline(variant, nc, ng, ngf, budget_s=15.0, wl="c2", workers=16):
    """The reference Fortran (oracle/_ref) on a bounded sample of the same
    workload: same p, rule and physics on a reduced mesh (at most 1600
    triangles, what the reference's static arrays hold), one core; then the
    same sample in concurrent processes, the reference being serial, as an
    all-core aggregate: one per host core up to `workers` (16, the host's
    CPU share for one GPU on the measurement machine, whatever nproc says).
    Rules of more than 33 points exceed the reference's arrays; then the C
    restatement runs instead (kind "port")."""
    runner, ne, nm, kind = _cpu_case(variant, nc, ng, ngf, wl)
    n, sec = _timed(runner, budget_s=budget_s)
    val = 4 * ne * nm * 3 * n / sec
    out = {"value": val, "unit": "element-DOF stage updates/s", "cores": 1, "kind": kind,
           "sample": f"{variant} p={nc} ng={ng} on {ne} triangles (the workload's generator "
                     f"at a reduced node count), {n} RK4 steps in {sec:.1f} s, "
                     f"1 thread"}
    out.update(host_info())
    k = max(1, min(workers, os.cpu_count() or 1))
    if k > 1:
        import subprocess
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker",
               ",".join(map(str, (wl, variant, nc, ng, ngf, n)))]
        procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, text=True) for _ in range(k)]
        runs = []
        for pr in procs:
            o, _ = pr.communicate()
            if pr.returncode == 0:
                runs.append(json.loads(o.strip().splitlines()[-1]))
        if len(runs) == k:
            slowest = max(r["s"] for r in runs)
            out["all_cores"] = {
                "value": 4 * ne * nm * 3 * sum(r["n"] for r in runs) / slowest,
                "processes": k, "nproc": os.cpu_count(),
                "sample": f"{k} concurrent processes of the same sample (min(nproc, {workers})), "
                          f"{n} steps each, slowest {slowest:.1f} s"}
    return out


def cpu_worker(spec):
    """--cpu-worker wl,variant,nc,ng,ngf,nsteps: one process of the all-core
    CPU measurement (no GPU use); prints {"n": steps, "s": seconds}."""
    wl, variant, nc, ng, ngf, n = spec.split(",")
    runner, _, _, _ = _cpu_case(variant, int(nc), int(ng), int(ngf), wl)
    k, sec = _timed(runner, n=int(n))
    print(json.dumps({"n": k, "s": sec}), flush=True)


def phase(msg):
    """Progress marker on stderr (the JSON line alone goes to stdout), so
    that a run that stops names the step it stopped in."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--cpu-worker":
        cpu_worker(sys.argv[2])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS),
                    help="c2 (default): the headline configuration; c1, c3, c4, c5: the other "
                         "configurations (c5 is the ensemble)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--members", type=int, default=64, help="c5 ensemble size")
    ap.add_argument("--adjoint", type=int, default=0, metavar="NDIR",
                    help="c4: also time the checkpointed adjoint of the run (gradient of the "
                         "eta^2 integral over the central 200 km disc after --steps steps, "
                         "with respect to the initial eta and to hb) and check it against "
                         "central differences of the FP64 path along NDIR seeded directions")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="with --adjoint: time the adjoint alone, without the central-difference "
                         "check (no FP64 forward runs)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=448)
    ap.add_argument("--ny", type=int, default=224)
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="steps of the plain FP64 path to time and to compare eta against "
                         "(default: same as --steps)")
    ap.add_argument("--gate-steps", type=int, default=None,
                    help="steps of the untimed accuracy leg (both paths from the same state; "
                         "default: the configuration's run length, 1000 for c2); 0 skips it")
    ap.add_argument("--dropin-steps", type=int, default=2,
                    help="steps of the per-call drop-in loop (4 spoc_rhs calls per step on "
                         "host arrays); 0 skips it")
    ap.add_argument("--numbering", default="generator", choices=("generator", "shuffled", "rows"),
                    help="the caller's element numbering: the generator's (Z order of the "
                         "centroids), a random permutation, or row by row")
    ap.add_argument("--ordering", default="auto", choices=("auto", "hilbert", "zcurve", "input"),
                    help="the library's internal element numbering (spoc_set_ordering)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel of the RK4 loop eagerly instead of replaying "
                         "captured HIP graphs (spoc_set_graph 0)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--tiles", default=None, metavar="T0,T1",
                    help="force the element kernels' tile codes (waves per workgroup + 16 x "
                         "form; lin or bc, then eg) instead of autotuning")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="before the timed region, copy MIB MiB once device-to-device (the HIP "
                         "runtime's blit kernel) so that a counter profile holds a kernel of "
                         "known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import spoc_amd
    from spoc_amd import configs

    wl = args.workload
    nc, ng, ngf = WORKLOADS[wl]["shape"]
    t = spoc_amd.build_tables(nc, *quad_rule(ng, ngf))
    tb = time.time()
    ctx = spoc_amd.Context(t)
    ctx.set_ordering(args.ordering)
    if args.no_graph:
        ctx.set_graph(False)
    if wl == "c5":
        members = configs.config5(t, nmem=args.members)
        if args.numbering != "generator":
            members = [renumber(mm, nc, args.numbering, 17 + i) for i, mm in enumerate(members)]
        c = members[0]
        ctx.set_ensemble_mesh([mm["mesh"] for mm in members])
        P = [mm["phys"] for mm in members]
        cat = spoc_amd.Context.concat
        ctx.set_fields(cat([q["hb"] for q in P]), P[0]["g0"], cat([q["fc"] for q in P]),
                       cat([q["w0"] for q in P]), cat([q["z0"] for q in P]), P[0]["mu0"],
                       P[0]["mb0"])
        state0 = tuple(cat([mm["state"][f] for mm in members]) for f in range(3))
        vect = np.concatenate([mm["mesh"]["vect"] for mm in members], axis=1)
        ne = int(sum(mm["mesh"]["ne"] for mm in members))
    else:
        if wl == "c2":
            c = configs.config2(t, nx=args.nx, ny=args.ny)
        elif wl == "c1":
            c = configs.config1(t)
            c["variant"] = "lin"
        elif wl == "c3":
            c = configs.config3(t)
        else:
            c = configs.config4(t)
        if args.numbering != "generator":
            c = renumber(c, nc, args.numbering)
        m, p = c["mesh"], c["phys"]
        ctx.set_mesh(m)
        ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
        state0 = c["state"]
        vect = m["vect"]
        ne = m["ne"]
    setup_s = time.time() - tb
    phase(f"set up {wl}: {ne} elements in {setup_s:.1f} s")
    variant, dt, nm = c["variant"], c["dt"], t["nm"]
    K = args.steps if args.steps is not None else WORKLOADS[wl]["steps"]
    W = args.warmup
    if args.calib_copy:
        calib_copy(args.calib_copy << 20)
    # mixed-precision path: the autotuner's choice of kernel form and
    # workgroup size on this mesh (it would run by itself on the first call)
    if args.tiles:
        # forced tile codes (profiling passes: the same instances as a traced run)
        t0c, t1c = (int(v) for v in args.tiles.split(","))
        ctx.set_tiles(variant, t0c, t1c, autotune=False)
        tiles = (t0c, t1c)
    else:
        phase("mixed path: autotune")
        tiles = ctx.autotune(variant)
    # warm up, then time K steps from the same initial state
    phase("mixed path: warm-up")
    ctx.upload(*state0)
    prime = warm(ctx, variant, dt, W, K, "fp32")
    phase(f"mixed path: {K} timed steps")
    ctx.upload(*state0)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.rk4(variant, dt, K, precision="fp32")   # the RK4 step replayed from a HIP graph
    ctx.sync()
    wall = time.perf_counter() - t0
    s32 = ctx.download()
    # per-kernel times (HIP events around each launch; eager launches)
    phase("mixed path: per-kernel timing")
    ctx.upload(*state0)
    ctx.set_profiling(True)
    kprof = min(K, 20)
    ctx.rk4(variant, dt, kprof, precision="fp32")
    dev_ms, kern = ctx.timing()
    ctx.set_profiling(False)
    per_stage = None
    if "lin_persistent" in kern:
        # small linear meshes run every step in one persistent launch; the
        # same steps launched per stage (HIP graphs), for comparison
        ctx.set_persistent(False)
        ctx.upload(*state0)
        warm(ctx, variant, dt, W, K, "fp32")
        ctx.upload(*state0)
        ctx.sync()
        t1 = time.perf_counter()
        ctx.rk4(variant, dt, K, precision="fp32")
        ctx.sync()
        per_stage = 1e3 * (time.perf_counter() - t1) / K
        ctx.set_persistent(True)
    dof = 4 * ne * nm * 3
    value = dof * K / wall
    out = {
        "metric": METRIC,
        "value": value, "unit": "element-DOF stage updates/s", "n_gpus": 1, "steps": K,
        "warmup": W, "graph_prime_steps": prime, "ms_per_step": 1e3 * wall / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if variant != "lin" else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[wl]["name"](nc, ng, ngf, ne, dt, args),
                   "ne": ne, "nm": nm, "numbering": args.numbering, "ordering": args.ordering,
                   "graphs": not args.no_graph,
                   "persistent": per_stage is not None,
                   "tiles": {"waves_per_workgroup": [t & 15 for t in tiles],
                             "kernel_form": [t >> 4 for t in tiles]},
                   "precision": ("fp32-mfma quadrature, fp64-mfma linear gravity-wave core, "
                                 "fp64 state/RK4") if variant != "lin" else
                                "fp64-mfma (the linear variant is all linear core)"},
    }
    if per_stage is not None:
        out["per_stage_ms_per_step"] = per_stage
    # roofline of the dominant kernel (HIP events on the context stream)
    counts = algorithmic_counts(nm, ng, ngf, variant, steps=kprof)
    best = max(kern.items(), key=lambda kv: kv[1][0]) if kern else None
    if best:
        name, (ms_tot, nlaunch) = best
        fl32, fl64, by = counts[name]
        fl = fl32 + fl64
        avg = ms_tot / max(nlaunch, 1) / 1e3
        flops_s, bytes_s = fl * ne / avg, by * ne / avg
        ai = fl / by
        # MFMA peak of the kernel's mix: its FP32 and FP64 products at their
        # own peaks, one after the other
        fpeak = fl / (fl32 / FP32_PEAK + fl64 / FP64_PEAK)
        hbm_bound = ai < fpeak / HBM_PEAK
        if hbm_bound:
            rf = {"bound": "hbm", "achieved": bytes_s / 1e9, "peak": HBM_PEAK / 1e9,
                  "unit": "GB/s", "frac": bytes_s / HBM_PEAK}
        else:
            rf = {"bound": "mfma", "achieved": flops_s / 1e12, "peak": fpeak / 1e12,
                  "unit": "TFLOP/s", "frac": flops_s / fpeak}
        rf.update(kernel=name, avg_launch_us=avg * 1e6, traffic=measured_traffic(name, nc, ne),
                  algorithmic_flops_per_elem=fl, algorithmic_fp64_flops_per_elem=fl64,
                  algorithmic_bytes_per_elem=by,
                  other={k: {"ms": v[0], "launches": v[1]} for k, v in kern.items()})
        out["roofline"] = rf
    # plain FP64 path: same initial state, same warm-up and steps
    if not args.no_fp64:
        K64 = args.fp64_steps or K
        phase(f"FP64 path: warm-up and {K64} timed steps")
        ctx.upload(*state0)
        warm(ctx, variant, dt, W, K64, "fp64")
        ctx.upload(*state0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.rk4(variant, dt, K64, precision="fp64")
        ctx.sync()
        w64 = time.perf_counter() - t0
        s64 = ctx.download()
        out["fp64_value"] = dof * K64 / w64
        out["fp64_ms_per_step"] = 1e3 * w64 / K64
        out["fp32_over_fp64"] = value / out["fp64_value"]
        # the FP64 path's own kernels (HIP events, eager launches) and its
        # dominant kernel against HBM, with the counter-measured traffic
        phase("FP64 path: per-kernel timing")
        ctx.upload(*state0)
        ctx.set_profiling(True)
        ctx.rk4(variant, dt, min(K64, 4), precision="fp64")
        _, k64 = ctx.timing()
        ctx.set_profiling(False)
        if k64:
            name64, (ms64, n64) = max(k64.items(), key=lambda kv: kv[1][0])
            avg64 = ms64 / max(n64, 1) / 1e3
            tr64 = measured_traffic(name64, nc, ne)
            out["fp64_roofline"] = {
                "bound": "hbm", "kernel": name64, "avg_launch_us": avg64 * 1e6,
                "traffic": tr64,
                "achieved": tr64 / avg64 / 1e9 if tr64 else None, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": tr64 / avg64 / HBM_PEAK if tr64 else None,
                "per_kernel_us": {k: 1e3 * v[0] / max(v[1], 1) for k, v in k64.items()}}
        if K64 == K:
            out["rel_l2_vs_fp64"] = {f: mass_rel_l2(t, vect, s32[i], s64[i])
                                     for i, f in enumerate(("u", "v", "eta"))}
            out["rel_l2_vs_fp64"]["steps"] = K
    G = WORKLOADS[wl]["gate"] if args.gate_steps is None else args.gate_steps
    if G and not args.no_fp64:
        phase(f"accuracy leg: {G} steps on both paths")
        out["accuracy_gate"] = gate_leg(ctx, t, variant, dt, state0, vect, G)
    if args.dropin_steps and wl != "c5":
        phase("per-call drop-in loop")
        out["dropin"] = dropin_timing(ctx, t, variant, dt, state0, args.dropin_steps)
    if args.adjoint and wl == "c4":
        phase("adjoint and central differences")
        out["adjoint"] = adjoint_check(ctx, t, c, K, 0 if args.adjoint_only else args.adjoint)
    phase("refinement indicator")
    out["cal_conv"] = adapt_timing(ctx, s32, ne, nc, ng, ngf, cpu=not args.no_cpu)
    if not args.no_cpu:
        phase("CPU baseline")
        try:
            out["cpu_baseline"] = cpu_baseline(variant, nc, ng, ngf, wl=wl)
        except Exception as ex:  # reported, not fatal
            out["cpu_baseline"] = {"error": repr(ex)}
    out["setup_s"] = setup_s
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
