"""Run directories for the reference's fixed-mesh driver (program shwater,
src/shwwnl.f): oc.inp (src/shwwnl.f:84-98), cm.dat (src/cmal_s.f:20-33) and
init.bin (7 unformatted sequential records, src/shwwnl.f:180-187) through
the library's file writers, and the quadrature files include/tri2_<ng>.dat,
include/legen<ngf>.dat (src/readgauss.f:16-41) from the committed fixtures."""
from __future__ import annotations

import os
import resource
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.path.join(ROOT, "oracle", "_ref")


def driver(variant: str, kind: str) -> str:
    """Path of the reference driver ('ref') or the drop-in one ('spoc')."""
    return os.path.join(REF, f"shwwnl_{variant}_{kind}")


def _record(f, payload: bytes) -> None:
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def write_case(d: str, *, nc: int, ng: int, ngf: int, x, y, tri, state, hb, w0, z0, dt: float,
               nite: int, per2: int, mu0: float, mb0: float, f0: float, beta: float, g0: float,
               lx: float = 1.0, ly: float = 1.0) -> None:
    os.makedirs(os.path.join(d, "include"), exist_ok=True)
    q = np.load(os.path.join(HERE, "golden", "quadrature.npz"), allow_pickle=False)
    with open(os.path.join(d, "include", f"tri2_{ng}.dat"), "w") as f:
        for w, a, b in q[f"tri{ng}"].T:
            f.write(f"{float(w)!r} {float(a)!r} {float(b)!r}\n")
    with open(os.path.join(d, "include", f"legen{ngf}.dat"), "w") as f:
        for w, t in q[f"legen{ngf}"].T:
            f.write(f"{float(w)!r} {float(t)!r}\n")
    nm = (nc + 1) * (nc + 2) // 2
    with open(os.path.join(d, "oc.inp"), "w") as f:
        r = lambda v: repr(float(v))
        for line in (f"{nc} {ng} {ngf}", f"{r(lx)} {r(ly)}", f"0 0 {nite}", r(dt), r(mu0),
                     r(mb0), r(f0), r(beta), "1.0 1.0", r(g0), r(hb[0, 0]), "1", str(per2),
                     "0.0"):
            f.write(line + "\n")
    # cm.dat and init.bin through the library's own writers
    import spoc_amd
    spoc_amd.write_mesh_file(os.path.join(d, "cm.dat"), x, y, np.asarray(tri))
    spoc_amd.write_state_file(os.path.join(d, "init.bin"), nm, 0.0,
                              *(np.asarray(a) for a in (*state, w0, z0, hb)))


def write_init_python(path: str, nm: int, time: float, arrays) -> None:
    """init.bin written here in Python (record framing by hand), a second
    writer for the tests of the library's."""
    with open(path, "wb") as f:
        _record(f, struct.pack("<d", time))
        for a in arrays:
            _record(f, np.ascontiguousarray(np.asarray(a)[:nm].T, "<f8").tobytes())


def read_temp(path: str, nm: int, ne: int):
    """temp.bin -> (time, u, v, h, w, z, hb) with fields (nm, ne)."""
    with open(path, "rb") as f:
        raw = f.read()
    out, o = [], 0
    while o < len(raw):
        n = struct.unpack_from("<i", raw, o)[0]
        out.append(raw[o + 4:o + 4 + n])
        o += n + 8
    t = struct.unpack("<d", out[0])[0]
    return (t, *(np.frombuffer(b, "<f8").reshape(ne, nm).T for b in out[1:]))


def run(exe: str, d: str, timeout: float, env=None) -> subprocess.CompletedProcess:
    """Run a driver in d; the nonlinear variants need a large stack."""
    def stack():
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout,
                          preexec_fn=stack, env=e)
