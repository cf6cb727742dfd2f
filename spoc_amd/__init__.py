"""MI355X (gfx950) right-hand side of SPOC's discontinuous-Galerkin shallow-water
model, behind a C-ABI (include/spoc_abi.h).

This module is the host-side mirror of the reference's operator interface:
``Context.cal_var`` takes the same operands as the reference subroutine
``cal_var`` (src/calvar_lin.f:7-10 and the three nonlinear variants) and
returns the tendencies ``(u1, v1, h1)``; ``Context.rk4`` runs the driver's RK4
loop (src/shwwnl.f:233-284) on device-resident state.  All arithmetic runs
in ``libspoc_amd.so``; if the library is missing this module raises instead
of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPOC_AMD_LIB names another build of the library (the host sanitizer build)
LIB_PATH = os.environ.get("SPOC_AMD_LIB") or os.path.join(HERE, "libspoc_amd.so")

LIN, FS, NS, INV = 0, 1, 2, 3
VARIANTS = {"lin": LIN, "fs": FS, "ns": NS, "inv": INV}
FP64, FP32_MFMA = 0, 1
PRECISIONS = {"fp64": FP64, "fp32": FP32_MFMA}
LDM = 45  # nnmod, the reference's leading dimension of modal arrays

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


class SpocError(RuntimeError):
    """A nonzero SPOC_E* status; .code holds it."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


ENAN = -7  # SPOC_ENAN: the state is no longer finite
ENOMEM = -8  # SPOC_ENOMEM: host memory or a host thread could not be had


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libspoc_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SpocError(f"{path} not found: build it with `python -m spoc_amd.build`")
    lib = C.CDLL(path)
    lib.spoc_last_error.restype = C.c_char_p
    for name in ("spoc_ctx_create", "spoc_ctx_destroy", "spoc_mesh_set", "spoc_fields_set",
                 "spoc_rhs", "spoc_state_upload", "spoc_state_download", "spoc_rk4_steps",
                 "spoc_energy", "spoc_sync", "spoc_tables_build", "spoc_last_timing",
                 "spoc_set_profiling", "spoc_set_graph", "spoc_set_persistent", "spoc_version",
                 "spoc_ensemble_mesh_set", "spoc_ensemble_energy", "spoc_rhs_adjoint",
                 "spoc_adjoint_upload", "spoc_adjoint_download", "spoc_rk4_adjoint_steps",
                 "spoc_autotune", "spoc_set_tiles", "spoc_adjoint_hb_download",
                 "spoc_set_checkpoint", "spoc_checkconv", "spoc_cal_conv", "spoc_mesh_build",
                 "spoc_mesh_changemaille", "spoc_state_file_read", "spoc_state_file_write",
                 "spoc_mesh_file_size", "spoc_mesh_file_read", "spoc_mesh_file_write",
                 "spoc_interpol", "spoc_calh2", "spoc_set_flux", "spoc_set_ordering",
                 "spoc_mesh_order", "spoc_energy_noalloc", "spoc_amr_create", "spoc_amr_adapt",
                 "spoc_amr_mesh", "spoc_amr_levels", "spoc_amr_inter_only_new"):
        getattr(lib, name).restype = C.c_int
    lib.spoc_ctx_destroy.argtypes = [C.c_void_p]
    lib.spoc_amr_destroy.argtypes = [C.c_void_p]
    lib.spoc_amr_destroy.restype = None
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise SpocError(f"spoc error {rc}: {load_library().spoc_last_error().decode()}", rc)


def _d(a: np.ndarray):
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def nm_of(nc: int) -> int:
    return (nc + 1) * (nc + 2) // 2


# -- the driver's files (host side, no device) --------------------------------
def read_state_file(path: str, nm: int, ne: int, ld: int = LDM):
    """init.bin / temp.bin (src/shwwnl.f:180-187): (time, u, v, h, w, z, hb)
    with (ld, ne) Fortran-ordered arrays."""
    lib = load_library()
    out = [np.zeros((ld, ne), order="F") for _ in range(6)]
    t = C.c_double()
    _check(lib.spoc_state_file_read(os.fsencode(path), nm, ne, ld, C.byref(t),
                                    *(_d(a) for a in out)))
    return (t.value, *out)


def write_state_file(path: str, nm: int, time: float, u, v, h, w, z, hb) -> None:
    """Write a state file the reference driver reads as init.bin
    (save_temp, src/shwwnl_adap.f:703-724)."""
    lib = load_library()
    arrs = [np.asfortranarray(a, np.float64) for a in (u, v, h, w, z, hb)]
    ld, ne = arrs[0].shape
    _check(lib.spoc_state_file_write(os.fsencode(path), nm, ne, ld, C.c_double(time),
                                     *(_d(a) for a in arrs)))


def read_mesh_file(path: str, scx: float = 1.0, scy: float = 1.0):
    """cm.dat (src/cmal_s.f:7-48): x, y (nn) scaled by scx, scy and the
    triangles (3, ne), 1-based int32."""
    lib = load_library()
    nn, ne = C.c_int(), C.c_int()
    _check(lib.spoc_mesh_file_size(os.fsencode(path), C.byref(nn), C.byref(ne)))
    x, y = np.zeros(nn.value), np.zeros(nn.value)
    tri = np.zeros((3, ne.value), np.int32, order="F")
    _check(lib.spoc_mesh_file_read(os.fsencode(path), nn.value, ne.value, C.c_double(scx),
                                   C.c_double(scy), _d(x), _d(y), _i(tri)))
    return x, y, tri


def write_mesh_file(path: str, x, y, tri, scx: float = 1.0, scy: float = 1.0) -> None:
    """cm.dat the reference reads (save_mesh's layout, src/cmal_s.f:56-78)."""
    lib = load_library()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    tri = np.asfortranarray(tri, np.int32)
    _check(lib.spoc_mesh_file_write(os.fsencode(path), len(x), tri.shape[1], _d(x), _d(y),
                                    _i(tri), C.c_double(scx), C.c_double(scy)))


def build_tables(nc: int, wi, xi, yi, wif, tif) -> dict:
    """Operator tables as the reference driver builds them (src/shwwnl.f:127-156),
    returned in the COMMON layouts (pi2(33|ng,45), pico(9,45,3), (45,45))."""
    lib = load_library()
    ng, ngf = len(wi), len(wif)
    ldg, ldf, ldm = max(33, ng), 9, LDM
    arr = lambda *s: np.zeros(s, order="F")
    t = dict(pi2=arr(ldg, ldm), pico=arr(ldf, ldm, 3), amm=arr(ldm, ldm), amml=arr(ldm, ldm),
             agx=arr(ldm, ldm), agy=arr(ldm, ldm), ammo=arr(ldm, ldm))
    q = [np.ascontiguousarray(a, np.float64) for a in (wi, xi, yi, wif, tif)]
    _check(lib.spoc_tables_build(nc, ng, ngf, *(_d(a) for a in q), ldg, ldf, ldm,
                                 _d(t["pi2"]), _d(t["pico"]), _d(t["amm"]), _d(t["amml"]),
                                 _d(t["agx"]), _d(t["agy"]), _d(t["ammo"])))
    t.update(nc=nc, ng=ng, ngf=ngf, nm=nm_of(nc), wi=q[0], xi=q[1], yi=q[2], wif=q[3], tif=q[4],
             ldg=ldg)
    return t


# -- the adaptive driver's re-projection onto a new mesh (host side) ----------
def interpol(tables: dict, old: dict, new_x, new_y, new_tri, tinter, fields):
    """Carry modal fields (each (ld, ne_old)) onto a new mesh (interpol +
    interlin, src/adapt.f:1662-1824): old is a mesh dict of mesh.build
    (x, y, in, itm, tseg, pseg); tinter[i] > 0 copies old element tinter[i].
    Returns the new fields, (ld, ne_new) each."""
    lib = load_library()
    t = tables
    fields = [np.asfortranarray(f, np.float64) for f in fields]
    ld = fields[0].shape[0]
    new_x = np.ascontiguousarray(new_x, np.float64)
    new_y = np.ascontiguousarray(new_y, np.float64)
    new_tri = np.asfortranarray(new_tri, np.int32)
    tinter = np.ascontiguousarray(tinter, np.int32)
    nenew = new_tri.shape[1]
    out = [np.zeros((ld, nenew), order="F") for _ in fields]
    P = C.POINTER(C.c_double)
    u = (P * len(fields))(*(_d(f) for f in fields))
    un = (P * len(out))(*(_d(f) for f in out))
    tri, itm = np.asfortranarray(old["in"], np.int32), np.asfortranarray(old["itm"], np.int32)
    tseg, pseg = (np.asfortranarray(old[k], np.int32) for k in ("tseg", "pseg"))
    x, y = np.ascontiguousarray(old["x"], np.float64), np.ascontiguousarray(old["y"], np.float64)
    _check(lib.spoc_interpol(t["nc"], t["ng"], _d(t["wi"]), _d(t["xi"]), _d(t["yi"]),
                             _d(t["pi2"]), t["ldg"], _d(t["amm"]), LDM, len(x), tri.shape[1],
                             tseg.shape[1], _d(x), _d(y), _i(tri), _i(itm), _i(tseg), _i(pseg),
                             len(new_x), nenew, _d(new_x), _d(new_y), _i(new_tri), _i(tinter),
                             len(fields), ld, u, un))
    return out


_FF = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_void_p)


def calh2(tables: dict, ff, x, y, tri, tinter=None, h0=None, ld: int = LDM):
    """L2 projection of ff(x, y) onto the mesh (calh2, src/adapt.f:303-373);
    elements with tinter[i] > 0 copy h0's element tinter[i]."""
    lib = load_library()
    t = tables
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    tri = np.asfortranarray(tri, np.int32)
    ne = tri.shape[1]
    tinter = np.zeros(ne, np.int32) if tinter is None else np.ascontiguousarray(tinter, np.int32)
    h1 = np.zeros((ld, ne), order="F")
    cb = _FF(lambda a, b, _u: float(ff(a, b)))
    h0a = np.asfortranarray(h0, np.float64) if h0 is not None else None
    h0p = _d(h0a) if h0a is not None else None
    ne0 = h0a.shape[1] if h0a is not None else 0
    _check(lib.spoc_calh2(t["nc"], t["ng"], _d(t["wi"]), _d(t["xi"]), _d(t["yi"]), _d(t["pi2"]),
                          t["ldg"], _d(t["amm"]), LDM, cb, None, len(x), ne, _d(x), _d(y),
                          _i(tri), _i(tinter), ne0, ld, h0p, _d(h1)))
    return h1


class Amr:
    """The adaptive driver's refinement hierarchy (spoc_amr_*; src/adapt.f
    init_adap, deraffiner, adapmaille, protridouble, elimpptri), host side.
    Built from a mesh dict of mesh.build (x, y, in, itm, tseg, nf);
    adapt(fconv, fderaff) runs one adaptation from per-element 0/1 flags
    (as cal_conv returns them) and returns the new mesh (x, y, tri (3, ne)
    1-based)."""

    def __init__(self, m: dict):
        self.lib = load_library()
        self.h = C.c_void_p()
        x, y = (np.ascontiguousarray(m[k], np.float64) for k in ("x", "y"))
        tri, itm, tseg = (np.asfortranarray(m[k], np.int32) for k in ("in", "itm", "tseg"))
        _check(self.lib.spoc_amr_create(C.byref(self.h), len(x), tri.shape[1], int(m["nf"]), _d(x),
                                        _d(y), _i(tri), _i(tseg), _i(itm)))
        self.nn, self.ne = len(x), tri.shape[1]

    def adapt(self, fconv, fderaff=None):
        fconv = np.ascontiguousarray(fconv, np.int32)
        fder = np.zeros(self.ne, np.int32) if fderaff is None else \
            np.ascontiguousarray(fderaff, np.int32)
        if len(fconv) != self.ne or len(fder) != self.ne:
            raise SpocError(f"flags must have the mesh's {self.ne} entries", -1)
        nn, ne = C.c_int(), C.c_int()
        _check(self.lib.spoc_amr_adapt(self.h, self.ne, _i(fconv), _i(fder), C.byref(nn),
                                       C.byref(ne)))
        self.nn, self.ne = nn.value, ne.value
        return self.mesh()[:3]

    def mesh(self):
        """Current mesh: x, y (nn), tri (3, ne) 1-based, tabsol (ne)."""
        x, y = np.zeros(self.nn), np.zeros(self.nn)
        tri = np.zeros((3, self.ne), np.int32, order="F")
        tabsol = np.zeros(self.ne, np.int32)
        _check(self.lib.spoc_amr_mesh(self.h, _d(x), _d(y), _i(tri), _i(tabsol)))
        return x, y, tri, tabsol

    def levels(self) -> np.ndarray:
        lev = np.zeros(self.ne, np.int32)
        _check(self.lib.spoc_amr_levels(self.h, _i(lev)))
        return lev

    def close(self):
        if self.h:
            self.lib.spoc_amr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def inter_only_new(tri, new_tri) -> np.ndarray:
    """tinter for interpol: the old triangle that reappears unchanged as each
    new triangle, or 0 (inter_only_new, src/adapt.f:1233-1258)."""
    lib = load_library()
    tri, new_tri = (np.asfortranarray(a, np.int32) for a in (tri, new_tri))
    tinter = np.zeros(new_tri.shape[1], np.int32)
    _check(lib.spoc_amr_inter_only_new(tri.shape[1], _i(tri), new_tri.shape[1], _i(new_tri),
                                       _i(tinter)))
    return tinter


ORDERINGS = {"input": 0, "hilbert": 1, "zcurve": 2, "auto": 3}


def mesh_order(meshes, mode: str = "hilbert") -> np.ndarray:
    """The library's internal element order of one mesh dict or a list of
    member meshes (spoc_mesh_order): perm[i] = the caller's element (0-based,
    union numbering) at internal position i."""
    lib = load_library()
    meshes = meshes if isinstance(meshes, (list, tuple)) else [meshes]
    ne = np.array([int(m["itm"].shape[1]) for m in meshes], np.int32)
    nf = np.array([int(m["nf"]) for m in meshes], np.int32)
    cat = lambda k, dt: np.ascontiguousarray(
        np.concatenate([np.asfortranarray(m[k], dt).ravel(order="F") for m in meshes]))
    perm = np.zeros(int(ne.sum()), np.int32)
    _check(lib.spoc_mesh_order(ORDERINGS[mode], len(meshes), _i(ne), _i(nf), _i(cat("itm", np.int32)),
                               _i(cat("tseg", np.int32)), _i(cat("ssit", np.int32)),
                               _d(cat("vect", np.float64)), _i(perm)))
    return perm - 1


class Context:
    """One device context: tables, a mesh, physics fields and RK4 state."""

    def __init__(self, tables: dict, device: int = 0):
        self.lib = load_library()
        self.t = tables
        self.nc, self.ng, self.ngf, self.nm = tables["nc"], tables["ng"], tables["ngf"], tables["nm"]
        self._h = C.c_void_p()
        t = tables
        _check(self.lib.spoc_ctx_create(C.byref(self._h), device, self.nc, self.ng, self.ngf,
                                        _d(t["wi"]), _d(t["wif"]), _d(t["pi2"]), t["ldg"],
                                        _d(t["pico"]), 9, LDM, _d(t["amm"]), _d(t["amml"]),
                                        _d(t["agx"]), _d(t["agy"]), _d(t["ammo"])))
        self.ne = self.nf = 0
        self.nmem = 1

    def close(self):
        if self._h:
            self.lib.spoc_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- operands -----------------------------------------------------------
    def set_mesh(self, m: dict) -> None:
        """m: the arrays of spoc_amd.mesh.build (or the reference's own)."""
        self._mesh = m  # keep the host arrays alive while in use
        ne, nf = int(m["itm"].shape[1]), int(m["nf"])
        have_nodes = "in" in m and "x" in m
        tri = np.asfortranarray(m["in"], np.int32) if have_nodes else None
        x = np.ascontiguousarray(m["x"], np.float64) if have_nodes else None
        y = np.ascontiguousarray(m["y"], np.float64) if have_nodes else None
        self._nodes = (tri, x, y)
        _check(self.lib.spoc_mesh_set(
            self._h, ne, nf, _i(np.asfortranarray(m["itm"], np.int32)),
            _i(np.asfortranarray(m["its"], np.int32)), _i(np.asfortranarray(m["tseg"], np.int32)),
            _i(np.asfortranarray(m["ssit"], np.int32)), _d(np.asfortranarray(m["norm"])),
            _d(np.asfortranarray(m["vect"])), _d(np.asfortranarray(m["normt"])),
            _i(np.ascontiguousarray(m["bdl"], np.int32)),
            len(x) if have_nodes else 0, _i(tri) if have_nodes else None,
            _d(x) if have_nodes else None, _d(y) if have_nodes else None))
        self.ne, self.nf = ne, nf
        self.nmem = 1

    def set_ensemble_mesh(self, meshes) -> None:
        """Several member meshes (dicts as for set_mesh) run as one union mesh;
        fields and state are then the members' arrays concatenated along
        elements (see concat)."""
        self._mesh = meshes
        ne = np.array([int(m["itm"].shape[1]) for m in meshes], np.int32)
        nf = np.array([int(m["nf"]) for m in meshes], np.int32)
        have_nodes = all("in" in m and "x" in m for m in meshes)
        nn = np.array([len(m["x"]) for m in meshes], np.int32) if have_nodes else None
        cat = lambda k, dt: np.ascontiguousarray(
            np.concatenate([np.asfortranarray(m[k], dt).ravel(order="F") for m in meshes]))
        arrs = dict(itm=cat("itm", np.int32), its=cat("its", np.int32), tseg=cat("tseg", np.int32),
                    ssit=cat("ssit", np.int32), norm=cat("norm", np.float64),
                    vect=cat("vect", np.float64), normt=cat("normt", np.float64),
                    bdl=cat("bdl", np.int32))
        if have_nodes:
            arrs.update(tri=cat("in", np.int32), x=cat("x", np.float64), y=cat("y", np.float64))
        self._ens = (ne, nf, nn, arrs)
        _check(self.lib.spoc_ensemble_mesh_set(
            self._h, len(meshes), _i(ne), _i(nf), _i(nn) if have_nodes else None,
            _i(arrs["itm"]), _i(arrs["its"]), _i(arrs["tseg"]), _i(arrs["ssit"]),
            _d(arrs["norm"]), _d(arrs["vect"]), _d(arrs["normt"]), _i(arrs["bdl"]),
            _i(arrs["tri"]) if have_nodes else None, _d(arrs["x"]) if have_nodes else None,
            _d(arrs["y"]) if have_nodes else None))
        self.ne, self.nf = int(ne.sum()), int(nf.sum())
        self.nmem = len(meshes)

    @staticmethod
    def concat(arrays):
        """Members' (ld, ne_m) arrays -> one (ld, sum ne_m) array."""
        return np.asfortranarray(np.concatenate([np.asarray(a) for a in arrays], axis=1))

    def ensemble_energy(self):
        n = self.nmem
        ke, pe, mass = (np.zeros(n) for _ in range(3))
        _check(self.lib.spoc_ensemble_energy(self._h, n, _d(ke), _d(pe), _d(mass)))
        return ke, pe, mass

    def set_fields(self, hb, g0, fc=None, w0=None, z0=None, mu0=0.0, mb0=0.0) -> None:
        ld = hb.shape[0]
        f = [np.asfortranarray(a, np.float64) if a is not None else None for a in (fc, w0, z0, hb)]
        _check(self.lib.spoc_fields_set(self._h, ld, *(None if a is None else _d(a) for a in f),
                                        C.c_double(mu0), C.c_double(mb0), C.c_double(g0)))

    # -- the reference call -------------------------------------------------
    def cal_var(self, variant, u0, v0, h0, precision="fp64", out=None):
        """Tendencies (u1, v1, h1) of the given state, shaped like u0 (into
        `out`, three Fortran-ordered float64 arrays of that shape, if given)."""
        var = VARIANTS[variant] if isinstance(variant, str) else variant
        prec = PRECISIONS[precision] if isinstance(precision, str) else precision
        u0, v0, h0 = (np.asfortranarray(a, np.float64) for a in (u0, v0, h0))
        ld = u0.shape[0]
        if out is None:
            out = [np.zeros_like(u0, order="F") for _ in range(3)]
        else:
            for a in out:
                if a.shape != u0.shape or a.dtype != np.float64 or not a.flags.f_contiguous:
                    raise ValueError("cal_var: out arrays must be Fortran-ordered float64 "
                                     "of the state's shape")
        _check(self.lib.spoc_rhs(self._h, var, prec, ld, _d(u0), _d(v0), _d(h0),
                                 *(_d(a) for a in out)))
        return tuple(out)

    def cal_var_adjoint(self, variant, au, av, ah, state=None, with_hb=False):
        """Transpose of the right-hand side's Jacobian at `state` applied to
        (au, av, ah); returns (ru, rv, rh) shaped like au, and the
        bathymetry adjoint as a 4th entry when with_hb."""
        var = VARIANTS[variant] if isinstance(variant, str) else variant
        a = [np.asfortranarray(x, np.float64) for x in (au, av, ah)]
        ld = a[0].shape[0]
        st_arrays = None if state is None else [np.asfortranarray(x, np.float64) for x in state]
        st = [None] * 3 if st_arrays is None else [_d(x) for x in st_arrays]
        out = [np.zeros_like(a[0], order="F") for _ in range(4)]
        _check(self.lib.spoc_rhs_adjoint(self._h, var, ld, *st, *(_d(x) for x in a),
                                         *(_d(x) for x in out[:3]),
                                         _d(out[3]) if with_hb else None))
        return tuple(out) if with_hb else tuple(out[:3])

    def adjoint_hb_download(self, ld: int = LDM):
        out = np.zeros((ld, self.ne), order="F")
        _check(self.lib.spoc_adjoint_hb_download(self._h, ld, _d(out)))
        return out

    def set_checkpoint(self, every: int) -> None:
        """Nonlinear adjoint: checkpoints every `every` steps (recomputing
        the steps between), or 0 (the default) to keep every step's stage
        inputs when they fit in device memory."""
        _check(self.lib.spoc_set_checkpoint(self._h, every))

    def adjoint_upload(self, au, av, ah) -> None:
        a = [np.asfortranarray(x, np.float64) for x in (au, av, ah)]
        _check(self.lib.spoc_adjoint_upload(self._h, a[0].shape[0], *(_d(x) for x in a)))

    def adjoint_download(self, ld: int = LDM):
        out = [np.zeros((ld, self.ne), order="F") for _ in range(3)]
        _check(self.lib.spoc_adjoint_download(self._h, ld, *(_d(x) for x in out)))
        return tuple(out)

    def rk4_adjoint(self, variant, dt: float, nsteps: int) -> float:
        """nsteps transposed RK4 steps, backwards; returns device ms."""
        var = VARIANTS[variant] if isinstance(variant, str) else variant
        _check(self.lib.spoc_rk4_adjoint_steps(self._h, var, C.c_double(dt), nsteps))
        return self.timing()[0]

    # -- device-resident RK4 ------------------------------------------------
    def upload(self, u0, v0, h0) -> None:
        u0, v0, h0 = (np.asfortranarray(a, np.float64) for a in (u0, v0, h0))
        _check(self.lib.spoc_state_upload(self._h, u0.shape[0], _d(u0), _d(v0), _d(h0)))

    def download(self, ld: int = LDM):
        out = [np.zeros((ld, self.ne), order="F") for _ in range(3)]
        _check(self.lib.spoc_state_download(self._h, ld, *(_d(a) for a in out)))
        return tuple(out)

    def rk4(self, variant, dt: float, nsteps: int, precision="fp64") -> float:
        """Advance nsteps; returns the device time of the call in ms."""
        var = VARIANTS[variant] if isinstance(variant, str) else variant
        prec = PRECISIONS[precision] if isinstance(precision, str) else precision
        _check(self.lib.spoc_rk4_steps(self._h, var, prec, C.c_double(dt), nsteps))
        return self.timing()[0]

    def sync(self) -> None:
        _check(self.lib.spoc_sync(self._h))

    def energy(self):
        ke, pe, mass = C.c_double(), C.c_double(), C.c_double()
        _check(self.lib.spoc_energy(self._h, C.byref(ke), C.byref(pe), C.byref(mass)))
        return ke.value, pe.value, mass.value

    def energy_noalloc(self, out_dev_ptr: int) -> None:
        """Queue (ke, pe, mass) per member into the device buffer at
        out_dev_ptr (3 nmem doubles) without any allocation; asynchronous on
        the context's stream (sync() before reading)."""
        _check(self.lib.spoc_energy_noalloc(self._h, C.c_void_p(out_dev_ptr)))

    def set_profiling(self, on: bool) -> None:
        _check(self.lib.spoc_set_profiling(self._h, int(on)))

    # -- the adaptive driver's refinement indicator ------------------------
    def checkconv(self, field):
        """Jump indicator of field 0 (u), 1 (v) or 2 (h) of the resident
        state, per element (checkconv, src/adapt.f:1260-1332)."""
        amax = np.zeros(self.ne)
        _check(self.lib.spoc_checkconv(self._h, int(field), _d(amax)))
        return amax

    def cal_conv(self, fac1, fac2, fac3, with_amax=True):
        """Refinement decisions of the resident state (cal_conv,
        src/adapt.f:1122-1192): dict with per-element fconv, fderaff (0/1)
        and amax (3, ne; None without with_amax), and conv, deraff, timedir
        (0/1) and supmax, one per ensemble member (scalars for a single
        mesh)."""
        ne, n = self.ne, self.nmem
        fconv, fderaff = np.zeros(ne, np.int32), np.zeros(ne, np.int32)
        conv, deraff, timedir = (np.zeros(n, np.int32) for _ in range(3))
        sup = np.zeros(n)
        amax = np.zeros((3, ne)) if with_amax else None
        _check(self.lib.spoc_cal_conv(self._h, C.c_double(fac1), C.c_double(fac2),
                                      C.c_double(fac3), _i(fconv), _i(fderaff), _i(conv),
                                      _i(deraff), _i(timedir), _d(sup),
                                      _d(amax) if with_amax else None))
        r = {"fconv": fconv, "fderaff": fderaff, "amax": amax}
        for k, v in (("conv", conv), ("deraff", deraff), ("timedir", timedir), ("supmax", sup)):
            r[k] = v if n > 1 else v[0].item()
        return r

    def autotune(self, variant):
        """Time the element kernels in each compiled form at 2, 4 and 8 waves
        per workgroup on this mesh; returns the (lin|bc, eg) tile codes kept
        (waves per workgroup + 16 x form)."""
        var = VARIANTS[variant] if isinstance(variant, str) else variant
        out = (C.c_int * 2)()
        _check(self.lib.spoc_autotune(self._h, var, out))
        return out[0], out[1]

    def set_flux(self, alpha: float = 0.0) -> None:
        """Face flux of the FP32-MFMA path: 0 central (the reference's),
        alpha > 0 local Lax-Friedrichs with speed alpha sqrt(g hh0)."""
        _check(self.lib.spoc_set_flux(self._h, C.c_double(alpha)))

    def set_tiles(self, variant, tile0=0, tile1=0, autotune=True) -> None:
        var = VARIANTS[variant] if isinstance(variant, str) else variant
        _check(self.lib.spoc_set_tiles(self._h, var, tile0, tile1, int(autotune)))

    def set_ordering(self, mode: str = "auto") -> None:
        """Internal element numbering from the next set_mesh: along the
        Hilbert or Z-order curve of the centroids, the caller's own
        ("input"), or ("auto", the default) Hilbert unless the caller's keeps
        at least as many neighbours in a block of 16.  Results and every
        array crossing the C-ABI are the same either way."""
        _check(self.lib.spoc_set_ordering(self._h, ORDERINGS[mode]))

    def set_graph(self, on: bool) -> None:
        """Replay the RK4 step from a HIP graph (default on)."""
        _check(self.lib.spoc_set_graph(self._h, int(on)))

    def set_persistent(self, on) -> None:
        """Run the linear variant's RK4 steps on the mixed path in one
        persistent launch on small meshes (at most 1,536 elements at p <= 3,
        1,024 at p = 4, 768 at p = 5; default on); the results are the same
        bit for bit."""
        _check(self.lib.spoc_set_persistent(self._h, int(on)))  # (2: test of the failure path)

    def timing(self):
        tot = C.c_double()
        n = C.c_int(16)
        names = C.create_string_buffer(1024)
        ms = (C.c_double * 16)()
        cnt = (C.c_int * 16)()
        _check(self.lib.spoc_last_timing(self._h, C.byref(tot), C.byref(n), names, 1024, ms, cnt))
        parts = names.raw.split(b"\0")
        kern = {parts[i].decode(): (ms[i], cnt[i]) for i in range(n.value)}
        return tot.value, kern
