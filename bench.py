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
HBM), bracketed by device synchronisation.  A step is 4 right-hand sides;
element-DOF stage updates per step = 4 * ne * nm * 3.

Prints ONE JSON line.  Multi-GPU runs are not part of this build.
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


def algorithmic_counts(nm, ng, ngf, variant):
    """Useful (unpadded) flops and compulsory HBM bytes per element per
    right-hand side of each kernel, from its product shapes and the arrays it
    must read and write once (neighbour traces counted once, RK4 traffic
    averaged over the four stages)."""
    e3 = 3 * ngf
    f4, f8 = 4, 8
    mesh = 16 * f8 + 6 * f4 + f4                       # geo64, nbr/nbe, bdl
    rk = 3 * nm * f8 * (1 + 0.75) + 3 * nm * f8 * 2     # c0, acc reads; acc|c0, x64 writes
    core_fl = 2 * (4 * nm * nm + 6 * nm * e3)            # -g grad eta, div(hh0 u), traces
    core_by = 3 * nm * f8 + 2 * e3 * f8 + rk + 2 * e3 * f8   # x64, trL in, RK, trL out
    if variant == "lin":
        return {"lin_f64": (core_fl, core_by + mesh)}
    bc_fl = 2 * (4 * nm * nm + 10 * nm * e3 + 11 * nm * ng)
    bc_by = (3 * nm * f8 + 2 * e3 * f4 + 16 * f4 + 6 * f4 + f4 + 4 * ng * f4 + 3 * f4
             + (4 + 2 + 2) * nm * f4 + 3 * e3 * f4)
    eg_fl = 2 * (6 * nm * nm + 3 * nm * e3) + core_fl
    eg_by = (8 * nm * f4 + 3 * e3 * f4 + 16 * f4 + mesh + core_by + 2 * e3 * f4)
    return {"bc_f32": (bc_fl, bc_by), "eg_f32f64": (eg_fl, eg_by)}


def measured_traffic(kernel, ne):
    """HBM bytes per launch of `kernel` from the committed counter profile
    (FETCH_SIZE and WRITE_SIZE scaled by a calibration copy), when that
    profile is of this workload; else None."""
    path = os.path.join(ROOT, "profiles", "fs_p6_199k_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None
    if f"on {ne} triangles" not in d.get("workload", ""):
        return None
    return d["bytes_per_launch"].get(kernel)


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


def cpu_baseline(variant, nc, ng, ngf, budget_s=15.0):
    """The reference Fortran (oracle/_ref) on a bounded sample of the same
    workload: same p, rule and physics on a 25 x 25 node jittered mesh (1152
    triangles; the reference's static arrays hold at most 1600), one core."""
    import oracle
    from spoc_amd import configs
    t = {"nc": nc, "ng": ng, "ngf": ngf}
    O = oracle.Oracle()
    O.setup(nc, *quad_rule(ng, ngf))
    tabs = O.tables()
    w, x, y, wf, tf = quad_rule(ng, ngf)
    tabs.update(nc=nc, ng=ng, ngf=ngf, nm=(nc + 1) * (nc + 2) // 2, wi=w, xi=x, yi=y, wif=wf,
                tif=tf)
    c = configs.config2(tabs, nx=25, ny=25)
    m, p = c["mesh"], c["phys"]
    ne, nm = m["ne"], tabs["nm"]
    kind, runner = "port", None
    if oracle.ref_available():
        try:
            R = oracle.RefLib(variant)
            R.setup(nc, *quad_rule(ng, ngf))
            kind = "reference"
            runner = lambda n: R.rk4(m, p, *c["state"], c["dt"], n)
        except OSError:
            runner = None
    if runner is None:
        kind = "port"
        runner = lambda n: O.rk4(variant, m, p, *c["state"], c["dt"], n)
    res = {}

    def work():
        t0 = time.perf_counter()
        runner(1)
        one = time.perf_counter() - t0
        n = max(1, int(budget_s / max(one, 1e-3)))
        t0 = time.perf_counter()
        runner(n)
        res["n"], res["s"] = n, time.perf_counter() - t0

    old = threading.stack_size(512 * 1024 * 1024)
    th = threading.Thread(target=work)
    th.start()
    th.join()
    threading.stack_size(old)
    val = 4 * ne * nm * 3 * res["n"] / res["s"]
    return {"value": val, "unit": "element-DOF stage updates/s", "cores": 1, "kind": kind,
            "sample": f"{variant} p={nc} ng={ng} on {ne} triangles (25x25 jittered nodes), "
                      f"{res['n']} RK4 steps in {res['s']:.1f} s, 1 thread, "
                      f"{platform.processor() or platform.machine()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=448)
    ap.add_argument("--ny", type=int, default=224)
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="steps of the plain FP64 path to time and to compare eta against "
                         "(default: same as --steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="before the timed region, copy MIB MiB once device-to-device (the HIP "
                         "runtime's blit kernel) so that a counter profile holds a kernel of "
                         "known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import spoc_amd
    from spoc_amd import configs

    nc, ng, ngf, variant = 6, 33, 7, "fs"
    t = spoc_amd.build_tables(nc, *quad_rule(ng, ngf))
    tb = time.time()
    c = configs.config2(t, nx=args.nx, ny=args.ny)
    setup_s = time.time() - tb
    m, p = c["mesh"], c["phys"]
    ne, nm = m["ne"], t["nm"]
    ctx = spoc_amd.Context(t)
    ctx.set_mesh(m)
    ctx.set_fields(p["hb"], p["g0"], p["fc"], p["w0"], p["z0"], p["mu0"], p["mb0"])
    dt = c["dt"]
    K, W = args.steps, args.warmup
    if args.calib_copy:
        calib_copy(args.calib_copy << 20)
    # FP32-MFMA path: warm up, then time K steps from the same initial state
    ctx.upload(*c["state"])
    if W:
        ctx.rk4(variant, dt, W, precision="fp32")
    ctx.upload(*c["state"])
    ctx.sync()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    ctx.rk4(variant, dt, K, precision="fp32")
    ctx.sync()
    wall = time.perf_counter() - t0
    dev_ms, kern = ctx.timing()
    ctx.set_profiling(False)
    s32 = ctx.download()
    dof = 4 * ne * nm * 3
    value = dof * K / wall
    out = {
        "metric": "element-DOF RK4 stage updates/s, FP32-MFMA vs FP64 GPU at eta rel-L2 <=1e-6",
        "value": value, "unit": "element-DOF stage updates/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": 1e3 * wall / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"calvarnl_fs p={nc} ng={ng} ngf={ngf} on {ne} triangles "
                               f"(jittered {args.nx}x{args.ny} nodes, 4000x2000 km, beta plane, "
                               f"double-gyre wind, mu0=100), dt={dt} s",
                   "ne": ne, "nm": nm,
                   "precision": "fp32-mfma quadrature, fp64-mfma linear gravity-wave core, "
                                "fp64 state/RK4"},
    }
    # roofline of the dominant kernel (HIP events on the context stream)
    counts = algorithmic_counts(nm, ng, ngf, variant)
    best = max(kern.items(), key=lambda kv: kv[1][0]) if kern else None
    if best:
        name, (ms_tot, nlaunch) = best
        fl, by = counts[name]
        avg = ms_tot / max(nlaunch, 1) / 1e3
        flops_s, bytes_s = fl * ne / avg, by * ne / avg
        ai = fl / by
        fpeak = FP32_PEAK if name == "bc_f32" else FP64_PEAK
        hbm_bound = ai < fpeak / HBM_PEAK
        if hbm_bound:
            rf = {"bound": "hbm", "achieved": bytes_s / 1e9, "peak": HBM_PEAK / 1e9,
                  "unit": "GB/s", "frac": bytes_s / HBM_PEAK}
        else:
            rf = {"bound": "mfma", "achieved": flops_s / 1e12, "peak": fpeak / 1e12,
                  "unit": "TFLOP/s", "frac": flops_s / fpeak}
        rf.update(kernel=name, avg_launch_us=avg * 1e6, traffic=measured_traffic(name, ne),
                  algorithmic_flops_per_elem=fl, algorithmic_bytes_per_elem=by,
                  other={k: {"ms": v[0], "launches": v[1]} for k, v in kern.items()})
        out["roofline"] = rf
    # plain FP64 path: same initial state, same steps; eta error of FP32
    if not args.no_fp64:
        K64 = args.fp64_steps or K
        ctx.upload(*c["state"])
        ctx.sync()
        t0 = time.perf_counter()
        ctx.rk4(variant, dt, K64, precision="fp64")
        ctx.sync()
        w64 = time.perf_counter() - t0
        s64 = ctx.download()
        out["fp64_value"] = dof * K64 / w64
        out["fp32_over_fp64"] = value / out["fp64_value"]
        if K64 == K:
            ammo = t["ammo"][:nm, :nm]
            det = 1.0 / m["vect"][4]

            def mrel(a, b):
                d = a[:nm] - b[:nm]
                n1 = np.sum(det * np.einsum("ie,ij,je->e", d, ammo, d))
                n2 = np.sum(det * np.einsum("ie,ij,je->e", b[:nm], ammo, b[:nm]))
                return float(np.sqrt(n1 / n2))

            out["rel_l2_vs_fp64"] = {"u": mrel(s32[0], s64[0]), "v": mrel(s32[1], s64[1]),
                                      "eta": mrel(s32[2], s64[2]), "steps": K}
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(variant, nc, ng, ngf)
        except Exception as ex:  # reported, not fatal
            out["cpu_baseline"] = {"error": repr(ex)}
    out["setup_s"] = setup_s
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
