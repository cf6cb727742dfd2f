"""Build recipe for the gfx950 library (libspoc_amd.so) and the Fortran drop-in.

``python -m spoc_amd.build`` (or ``__graft_entry__.build()``) always
recompiles.  Every HIP translation unit is built for gfx950 only, with packed
FP32 instructions disabled; the FP64 kernels and the host table builder are
compiled without FMA contraction so that they reproduce the reference's
arithmetic exactly.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspoc_amd.so")
ROOT = os.path.dirname(HERE)

SHAPES = [(2, 6, 3), (3, 12, 4), (4, 25, 5), (5, 25, 6), (6, 33, 7), (7, 42, 8)]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
BASE = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument"]
EXACT = ["-ffp-contract=off"]


def _units():
    units = []
    for nc, ng, ngf in SHAPES:
        tag = f"{nc}_{ng}_{ngf}"
        d = [f"-DSPOC_NC={nc}", f"-DSPOC_NG={ng}", f"-DSPOC_NGF={ngf}"]
        units.append(("rhs_f64.hip", f"rhs_f64_{tag}.o", EXACT + d))
        units.append(("rhs_f32.hip", f"rhs_f32_{tag}.o", d))
        units.append(("adjoint.hip", f"adj_f64_{tag}.o", d))
    units.append(("f32_plan.cpp", "f32_plan.o", ["-x", "hip"]))
    units.append(("f64_plan.cpp", "f64_plan.o", EXACT + ["-x", "hip"]))
    units.append(("energy.hip", "energy.o", EXACT))
    units.append(("adapt.hip", "adapt.o", EXACT))
    units.append(("mesh.cpp", "mesh.o", EXACT))
    units.append(("files.cpp", "files.o", []))
    units.append(("interp.cpp", "interp.o", EXACT))
    units.append(("amr.cpp", "amr.o", EXACT))
    units.append(("order.cpp", "order.o", []))
    units.append(("abi.cpp", "abi.o", EXACT + ["-x", "hip"]))
    return units


def _compile(u):
    src, obj, flags = u
    base = [f for f in BASE if f != "-O3"] if "-O1" in flags else BASE
    cmd = [HIPCC] + base + flags + ["-c", os.path.join(CSRC, src), "-o", os.path.join(OBJ, obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    units = _units()
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, units))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + \
          [os.path.join(OBJ, o) for o in objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_fortran(verbose)
    return LIB


# Host sanitizer build (CPU only): the host translation units again with
# AddressSanitizer and UndefinedBehaviorSanitizer on their host side, linked
# with the device objects of build().  The result is loaded by setting
# SPOC_AMD_LIB to it and preloading the sanitizer runtime (SAN_RUNTIME).
SAN_DIR = os.path.join(OBJ, "san")
SAN_LIB = os.path.join(SAN_DIR, "libspoc_amd_san.so")
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
HOST_UNITS = ("f32_plan.cpp", "f64_plan.cpp", "energy.hip", "adapt.hip", "mesh.cpp", "files.cpp",
              "interp.cpp", "amr.cpp", "order.cpp", "abi.cpp")


def san_runtime() -> str:
    """Path of the compiler's shared AddressSanitizer runtime."""
    clang = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "llvm", "bin", "clang")
    if not os.path.exists(clang):
        clang = "/opt/rocm/llvm/bin/clang"
    r = subprocess.run([clang, "--print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True, check=True)
    return r.stdout.strip()


def build_sanitized(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(SAN_DIR, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    host, dev = [], []
    for src, obj, flags in _units():
        if src in HOST_UNITS:
            host.append((src, os.path.join("san", obj), flags + ["-O1", "-g"] + SAN_FLAGS))
        else:
            dev.append((src, obj, flags))
    missing = [u for u in dev if not os.path.exists(os.path.join(OBJ, u[1]))]
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, host + missing))
    objs += [u[1] for u in dev if u not in missing]
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", SAN_LIB] + SAN_FLAGS + \
          [os.path.join(OBJ, o) for o in objs]
    subprocess.run(cmd, check=True)
    if verbose:
        print(f"built {SAN_LIB}")
    return SAN_LIB


def build_fortran(verbose: bool = True) -> None:
    """Compile the ISO_C_BINDING drop-in objects: cal_var (one per variant)
    and the adaptive driver's cal_conv."""
    fc = shutil.which("flang") or shutil.which("amdflang")
    if not fc:
        if verbose:
            print("flang not found: Fortran drop-in not built")
        return
    fdir = os.path.join(HERE, "fortran")
    src = os.path.join(fdir, "spoc_calvar.F90")
    if not os.path.exists(src):
        return
    os.makedirs(OBJ, exist_ok=True)
    for v, code in (("lin", 0), ("fs", 1), ("ns", 2), ("inv", 3)):
        out = os.path.join(fdir, f"spoc_calvar_{v}.o")
        subprocess.run([fc, "-O2", "-fPIC", f"-DSPOC_VARIANT={code}", "-J", OBJ, "-c", src,
                        "-o", out], check=True)
    # the adaptive driver's cal_conv (uses the spoc_abi module written above)
    subprocess.run([fc, "-O2", "-fPIC", "-J", OBJ, "-c", os.path.join(fdir, "spoc_calconv.F90"),
                    "-o", os.path.join(fdir, "spoc_calconv.o")], check=True)
    if verbose:
        print("built Fortran drop-in objects in", fdir)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "sanitized":
        build_sanitized()
    else:
        build(int(sys.argv[1]) if len(sys.argv) > 1 else None)
