"""ctypes access to the oracle -- TEST INFRASTRUCTURE ONLY.

Two checkers live here:

* ``Oracle``: our C restatement of the reference (``oracle/spoc_oracle.c``,
  built to ``oracle/liboracle.so``);
* ``RefLib``: the reference Fortran itself, compiled from its own sources
  into ``oracle/_ref/libspoc_ref_<variant>.so`` by ``oracle/Makefile``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, as the checker or the CPU baseline.  The product
package ``spoc_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
LDM, LDF, LDG = 45, 9, 64
VARIANTS = {"lin": 0, "fs": 1, "ns": 2, "inv": 3}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    assert a.dtype == np.float64 and a.flags.f_contiguous or a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, when ``ref``, the reference libraries and
    the reference driver built with its own cal_var and with the drop-in)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "drivers"], check=True)
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "adaptive"], check=True)


def ref_available() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, f"libspoc_ref_{v}.so")) for v in VARIANTS)


class Oracle:
    """The C restatement.  Arrays are (LDM, ne) Fortran-ordered float64."""

    def __init__(self):
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_tables_size.restype = C.c_size_t
        L.orc_po.restype = C.c_double
        L.orc_po.argtypes = [C.c_int, C.c_double]
        L.orc_dpo.restype = C.c_double
        L.orc_dpo.argtypes = [C.c_int, C.c_double]
        L.orc_setup.restype = C.c_int
        self._tab = None

    def setup(self, nc, wi, xi, yi, wif, tif):
        self._tab = C.create_string_buffer(self.lib.orc_tables_size())
        wi, xi, yi, wif, tif = (np.ascontiguousarray(a, np.float64) for a in (wi, xi, yi, wif, tif))
        self._keep = (wi, xi, yi, wif, tif)
        rc = self.lib.orc_setup(self._tab, nc, len(wi), len(wif), _d(wi), _d(xi), _d(yi), _d(wif),
                                _d(tif))
        self.nc, self.ng, self.ngf = nc, len(wi), len(wif)
        self.nm = (nc + 1) * (nc + 2) // 2
        return rc

    def tables(self):
        """Decode the table struct into numpy arrays (Fortran-ordered)."""
        raw = np.frombuffer(self._tab, dtype=np.uint8)
        off = 4 * 4
        def take(n, dt=np.float64):
            nonlocal off
            al = np.dtype(dt).itemsize
            off = (off + al - 1) // al * al
            a = raw[off:off + n * al].view(dt).copy()
            off += n * al
            return a
        out = {}
        out["wi"], out["xi"], out["yi"] = take(LDG), take(LDG), take(LDG)
        out["wif"], out["tif"] = take(LDF), take(LDF)
        out["pi2"] = take(LDG * LDM).reshape(LDM, LDG).T
        out["pico"] = take(LDF * LDM * 3).reshape(3, LDM, LDF).transpose(2, 1, 0)
        for k in ("amm", "amml", "agx", "agy", "ammo"):
            out[k] = take(LDM * LDM).reshape(LDM, LDM).T
        out["cip"] = take(2 * (LDM + 1), np.int32).reshape(LDM + 1, 2).T
        out["info1"], out["info2"] = take(2, np.int32)
        return out

    def cal_flu(self, m, a0):
        fb = np.zeros((LDF, m["nf"]), order="F")
        a0 = np.asfortranarray(a0)
        self.lib.orc_cal_flu(self._tab, m["nf"], _i(m["tseg"]), _d(fb), _d(a0), a0.shape[0],
                             _i(m["ssit"]))
        return fb

    def _args(self, m, p):
        return (_i(m["tseg"]), _i(m["ssit"]), _i(m["itm"]), _i(m["its"]), _d(m["norm"]),
                _i(m["bdl"]), _d(m["vect"]), _d(m["normt"]), C.c_double(p["mu0"]),
                C.c_double(p["mb0"]), _d(p["fc"]), C.c_double(p["g0"]), _d(p["w0"]), _d(p["z0"]),
                _d(p["hb"]))

    def cal_var(self, variant, m, p, u0, v0, h0):
        ld, ne = u0.shape
        u1, v1, h1 = (np.zeros((ld, ne), order="F") for _ in range(3))
        u0, v0, h0 = (np.asfortranarray(a) for a in (u0, v0, h0))
        self.lib.orc_cal_var(self._tab, VARIANTS[variant], m["nf"], ne, ld, _d(u0), _d(v0),
                             _d(h0), _d(u1), _d(v1), _d(h1), *self._args(m, p))
        return u1, v1, h1

    def rk4(self, variant, m, p, u0, v0, h0, dt, nsteps):
        ld, ne = u0.shape
        u, v, h = (np.array(a, order="F", copy=True) for a in (u0, v0, h0))
        self.lib.orc_rk4(self._tab, VARIANTS[variant], m["nf"], ne, ld, _d(u), _d(v), _d(h),
                         *self._args(m, p), C.c_double(dt), nsteps)
        return u, v, h

    def checkconv(self, m, a0):
        """Edge-jump indicator of one field (src/adapt.f:1260-1332)."""
        ld, ne = a0.shape
        amax = np.zeros(ne)
        a0 = np.asfortranarray(a0)
        self.lib.orc_checkconv(self._tab, m["nf"], ne, ld, _d(a0), _i(m["tseg"]), _i(m["ssit"]),
                               _d(amax))
        return amax

    def cal_conv(self, m, u0, v0, h0, fac1, fac2, fac3):
        """Refinement decisions (src/adapt.f:1122-1192): dict of fconv,
        fderaff (0/1 per element), conv, deraff, timedir (0/1), supmax and
        the three indicators amax (3, ne)."""
        ld, ne = u0.shape
        u0, v0, h0 = (np.asfortranarray(a) for a in (u0, v0, h0))
        fconv, fderaff = np.zeros(ne, np.int32), np.zeros(ne, np.int32)
        conv, deraff, timedir = C.c_int(), C.c_int(), C.c_int()
        sup = C.c_double()
        amax = np.zeros((3, ne))
        self.lib.orc_cal_conv(self._tab, m["nf"], ne, ld, _d(u0), _d(v0), _d(h0), _i(m["tseg"]),
                              _i(m["ssit"]), C.c_double(fac1), C.c_double(fac2),
                              C.c_double(fac3), _i(fconv), _i(fderaff), C.byref(conv),
                              C.byref(deraff), C.byref(timedir), C.byref(sup), _d(amax))
        return {"fconv": fconv, "fderaff": fderaff, "conv": conv.value, "deraff": deraff.value,
                "timedir": timedir.value, "supmax": sup.value, "amax": amax}

    def energy(self, m, p, u0, v0, h0):
        ld, ne = u0.shape
        ke, pe, mass = C.c_double(), C.c_double(), C.c_double()
        u0, v0, h0 = (np.asfortranarray(a) for a in (u0, v0, h0))
        self.lib.orc_energy(self._tab, ne, ld, _d(u0), _d(v0), _d(h0), C.c_double(p["g0"]),
                            _d(p["hb"]), _i(m["in"]), _d(m["x"]), _d(m["y"]), C.byref(ke),
                            C.byref(pe), C.byref(mass))
        return ke.value, pe.value, mass.value


class RefLib:
    """The reference Fortran built from its sources (one library per variant)."""

    def __init__(self, variant: str):
        path = os.path.join(REF_DIR, f"libspoc_ref_{variant}.so")
        self.lib = C.CDLL(path, mode=os.RTLD_LOCAL)
        self.lib.ref_po.restype = C.c_double
        self.lib.ref_po.argtypes = [C.c_int, C.c_double]
        self.lib.ref_dpo.restype = C.c_double
        self.lib.ref_dpo.argtypes = [C.c_int, C.c_double]
        self.variant = variant

    def read_gauss(self, nc, ng, ngf, refroot):
        """Read include/tri2_<ng>.dat, include/legen<ngf>.dat with the reference reader."""
        w, x, y = np.zeros(ng), np.zeros(ng), np.zeros(ng)
        wf, tf = np.zeros(ngf), np.zeros(ngf)
        cwd = os.getcwd()
        try:
            os.chdir(refroot)
            self.lib.ref_read_gauss(nc, ng, ngf, _d(w), _d(x), _d(y), _d(wf), _d(tf))
        finally:
            os.chdir(cwd)
        return w, x, y, wf, tf

    def setup(self, nc, wi, xi, yi, wif, tif):
        i1, i2 = C.c_int(), C.c_int()
        wi, xi, yi, wif, tif = (np.ascontiguousarray(a, np.float64) for a in (wi, xi, yi, wif, tif))
        self.lib.ref_setup(nc, len(wi), len(wif), _d(wi), _d(xi), _d(yi), _d(wif), _d(tif),
                           C.byref(i1), C.byref(i2))
        return i1.value, i2.value

    def tables(self):
        pi2 = np.zeros((33, LDM), order="F")
        pico = np.zeros((LDF, LDM, 3), order="F")
        mats = [np.zeros((LDM, LDM), order="F") for _ in range(5)]
        cip = np.zeros((2, LDM + 1), np.int32, order="F")
        self.lib.ref_get_tables(_d(pi2), _d(pico), *(_d(a) for a in mats), _i(cip))
        out = dict(pi2=pi2, pico=pico, cip=cip)
        for k, a in zip(("amm", "amml", "agx", "agy", "ammo"), mats):
            out[k] = a
        return out

    def mesh(self, x, y, tri, nc, f0, beta):
        """Reference pointer_mesh/check_mesh/cal_rotnf/cal_its/calpre/coriolis."""
        nn, ne = len(x), tri.shape[1]
        cap = max(2400, 2 * ne + 16)
        tri = np.asfortranarray(tri, dtype=np.int32)
        x, y = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64)
        nf, nbdl = C.c_int(), C.c_int()
        itm = np.zeros((3, ne), np.int32, order="F")
        its = np.zeros((3, ne), np.int32, order="F")
        tseg = np.zeros((2, cap), np.int32, order="F")
        pseg = np.zeros((2, cap), np.int32, order="F")
        ssit = np.zeros((2, cap), np.int32, order="F")
        norm = np.zeros((2, cap), order="F")
        vect = np.zeros((5, ne), order="F")
        normt = np.zeros((2, ne), order="F")
        bdl = np.zeros(ne, np.int32)
        fc = np.zeros((LDM, ne), order="F")
        self.lib.ref_mesh(nn, ne, _i(tri), _d(x), _d(y), nc, C.c_double(f0), C.c_double(beta),
                          C.byref(nf), _i(itm), _i(its), _i(tseg), _i(pseg), _i(ssit), _d(norm),
                          _d(vect), _d(normt), _i(bdl), _d(fc), C.byref(nbdl))
        n = nf.value
        return dict(nf=n, itm=itm, its=its, tseg=np.asfortranarray(tseg[:, :n]),
                    pseg=np.asfortranarray(pseg[:, :n]), ssit=np.asfortranarray(ssit[:, :n]),
                    norm=np.asfortranarray(norm[:, :n]), vect=vect, normt=normt, bdl=bdl, fc=fc)

    def cal_var(self, m, p, u0, v0, h0, time=0.0):
        ld, ne = u0.shape
        assert ld == LDM
        u1, v1, h1 = (np.zeros((LDM, ne), order="F") for _ in range(3))
        u0, v0, h0 = (np.asfortranarray(a) for a in (u0, v0, h0))
        self.lib.ref_cal_var(m["nf"], ne, _d(u0), _d(v0), _d(h0), _d(u1), _d(v1), _d(h1),
                             _i(m["tseg"]), _i(m["ssit"]), _i(m["itm"]), _i(m["its"]),
                             _d(m["norm"]), _i(m["bdl"]), _d(m["vect"]), _d(m["normt"]),
                             C.c_double(p["mu0"]), C.c_double(p["mb0"]), _d(p["fc"]),
                             C.c_double(p.get("beta", 0.0)), _d(p["w0"]), _d(p["z0"]),
                             C.c_double(p["g0"]), _d(p["hb"]), C.c_double(time))
        return u1, v1, h1

    def rk4(self, m, p, u0, v0, h0, dt, nsteps):
        ld, ne = u0.shape
        assert ld == LDM
        u, v, h = (np.array(a, order="F", copy=True) for a in (u0, v0, h0))
        self.lib.ref_rk4(m["nf"], ne, _d(u), _d(v), _d(h), _i(m["tseg"]), _i(m["ssit"]),
                         _i(m["itm"]), _i(m["its"]), _d(m["norm"]), _i(m["bdl"]), _d(m["vect"]),
                         _d(m["normt"]), C.c_double(p["mu0"]), C.c_double(p["mb0"]), _d(p["fc"]),
                         C.c_double(p.get("beta", 0.0)), _d(p["w0"]), _d(p["z0"]),
                         C.c_double(p["g0"]), _d(p["hb"]), C.c_double(dt), nsteps)
        return u, v, h

    def energy(self, m, p, u0, v0, h0):
        ne = u0.shape[1]
        ke, pe, mass = C.c_double(), C.c_double(), C.c_double()
        u0, v0, h0 = (np.asfortranarray(a) for a in (u0, v0, h0))
        self.lib.ref_energy(ne, _d(u0), _d(v0), _d(h0), C.c_double(p["g0"]), _d(p["hb"]),
                            C.byref(ke), C.byref(pe), C.byref(mass), _i(m["in"]), _d(m["x"]),
                            _d(m["y"]))
        return ke.value, pe.value, mass.value


class RefCM:
    """The reference CM tool (corner fix + edge-1-on-boundary rotation)."""

    def __init__(self):
        self.lib = C.CDLL(os.path.join(REF_DIR, "libspoc_ref_cm.so"), mode=os.RTLD_LOCAL)

    def changemaille(self, nn, tri):
        tri = np.array(tri, dtype=np.int32, order="F", copy=True)
        self.lib.ref_changemaille(nn, tri.shape[1], _i(tri))
        return tri


class RefAdapt:
    """The reference's AMR routines (src/adapt.f), compiled from its
    sources: the decisions checkconv and cal_conv, the interpolations, and
    the refinement hierarchy.  Modal arrays are (45, ne) Fortran-ordered;
    pico is the (9, 45, 3) table."""

    def __init__(self):
        self.lib = C.CDLL(os.path.join(REF_DIR, "libspoc_ref_adapt.so"), mode=os.RTLD_LOCAL)

    @staticmethod
    def available() -> bool:
        return os.path.exists(os.path.join(REF_DIR, "libspoc_ref_adapt.so"))

    @staticmethod
    def _pad(a):
        out = np.zeros((LDM, a.shape[1]), order="F")
        out[:a.shape[0]] = a
        return out

    def setup(self, nc, wi, xi, yi, wif, tif):
        """Install the driver's tables in this library's COMMONs (ref_setup)."""
        i1, i2 = C.c_int(), C.c_int()
        wi, xi, yi, wif, tif = (np.ascontiguousarray(a, np.float64) for a in (wi, xi, yi, wif, tif))
        self.lib.ref_setup(nc, len(wi), len(wif), _d(wi), _d(xi), _d(yi), _d(wif), _d(tif),
                           C.byref(i1), C.byref(i2))
        return i1.value, i2.value

    def interpol(self, old, xn, yn, trin, tinter, u):
        """interpol (src/adapt.f:1662-1739) of one field; setup first."""
        ne = old["in"].shape[1]
        u = self._pad(u)
        xn, yn = np.ascontiguousarray(xn, np.float64), np.ascontiguousarray(yn, np.float64)
        trin = np.asfortranarray(trin, np.int32)
        tinter = np.ascontiguousarray(tinter, np.int32)
        unew = np.zeros((LDM, trin.shape[1]), order="F")
        x, y = np.ascontiguousarray(old["x"]), np.ascontiguousarray(old["y"])
        self.lib.ref_interpol(ne, len(x), _d(x), _d(y), _i(np.asfortranarray(old["in"], np.int32)),
                              _i(old["tseg"]), _i(old["pseg"]), _i(old["itm"]), _i(tinter),
                              _d(u), _d(unew), trin.shape[1], len(xn), _d(xn), _d(yn), _i(trin))
        return unew

    def calh2(self, x, y, tri, tinter, h0):
        """calh2 (src/adapt.f:303-373) of the harness's test function."""
        tri = np.asfortranarray(tri, np.int32)
        h0 = self._pad(h0)
        h1 = np.zeros((LDM, tri.shape[1]), order="F")
        self.lib.ref_calh2(tri.shape[1], _d(np.ascontiguousarray(x)), _d(np.ascontiguousarray(y)),
                           _i(tri), _i(np.ascontiguousarray(tinter, np.int32)), _d(h0), _d(h1))
        return h1

    def checkconv(self, nc, ng, ngf, m, pico, a0):
        ne = a0.shape[1]
        a0 = self._pad(a0)
        pico = np.asfortranarray(pico)
        amax = np.zeros(ne)
        self.lib.ref_checkconv(nc, ng, ngf, ne, m["nf"], _d(a0), _d(pico), _i(m["tseg"]),
                               _i(m["ssit"]), _d(amax))
        return amax

    def cal_conv(self, nc, ng, ngf, m, pico, u0, v0, h0, fac1, fac2, fac3):
        ne = u0.shape[1]
        u0, v0, h0 = (self._pad(a) for a in (u0, v0, h0))
        pico = np.asfortranarray(pico)
        fconv, fderaff = np.zeros(ne, np.int32), np.zeros(ne, np.int32)
        conv, deraff, timedir = C.c_int(), C.c_int(), C.c_int()
        self.lib.ref_cal_conv(nc, ng, ngf, ne, m["nf"], _d(pico), _i(m["tseg"]), _i(m["ssit"]),
                              _d(u0), _d(v0), _d(h0), C.c_double(fac1), C.c_double(fac2),
                              C.c_double(fac3), _i(fconv), _i(fderaff), C.byref(conv),
                              C.byref(deraff), C.byref(timedir))
        return {"fconv": fconv, "fderaff": fderaff, "conv": conv.value, "deraff": deraff.value,
                "timedir": timedir.value}

    # the reference's hierarchy sizes (include/maille.inc:6-7)
    NNDIM, NEDIM = 800, 1600

    def adap_init(self, x, y, tri, tseg, itm):
        """init_adap (src/adapt.f:380-449): the hierarchy of this mesh
        (tri, itm (3, ne) and tseg (2, nf), 1-based)."""
        tri = np.asfortranarray(tri, np.int32)
        self.lib.ref_adap_init(len(x), tri.shape[1], _d(np.ascontiguousarray(x, np.float64)),
                               _d(np.ascontiguousarray(y, np.float64)), _i(tri),
                               _i(np.asfortranarray(tseg, np.int32)),
                               _i(np.asfortranarray(itm, np.int32)))

    def adap_cycle(self, nn, fconv, fderaff):
        """One adaptation of the current hierarchy (global_refinment's
        sequence, src/adapt.f:89-119) from per-element flags; returns the new
        mesh x, y, tri (3, ne) and tabsol."""
        ne = len(fconv)
        fconv = np.ascontiguousarray(fconv, np.int32)
        fder = np.zeros(2 * self.NEDIM, np.int32)
        fder[:ne] = fderaff
        xn, yn = np.zeros(2 * self.NNDIM), np.zeros(2 * self.NNDIM)
        trin = np.zeros((3, self.NEDIM), np.int32, order="F")
        tabs = np.zeros(self.NEDIM, np.int32)
        nnn, nen = C.c_int(), C.c_int()
        self.lib.ref_adap_cycle(nn, ne, _i(fconv), _i(fder), C.byref(nnn), C.byref(nen), _d(xn),
                                _d(yn), _i(trin), _i(tabs))
        n, e = nnn.value, nen.value
        return xn[:n].copy(), yn[:n].copy(), np.asfortranarray(trin[:, :e]), tabs[:e].copy()

    def inter_only_new(self, tri, trin):
        """inter_only_new (src/adapt.f:1233-1258): tinter (nenew)."""
        tri = np.asfortranarray(tri, np.int32)
        ext = np.zeros((3, trin.shape[1] + 1), np.int32, order="F")  # a zero column past the end
        ext[:, :trin.shape[1]] = trin
        tinter = np.zeros(trin.shape[1], np.int32)
        self.lib.ref_inter_only_new(tri.shape[1], trin.shape[1], _i(tri), _i(ext), _i(tinter))
        return tinter
