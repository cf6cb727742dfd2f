"""Host-side mesh preparation, reproducing the reference's integer conventions.

The reference builds these arrays in Fortran once per (re)mesh; they are the
integer/geometry operands of the right-hand side.  Face numbering, edge
orientation and the boundary flags must match the reference bit for bit, so
each routine states the rule it follows:

* ``rmal``            structured mesh, src/rmal.f:1-48
* ``changemaille``    corner fix + edge-1-on-boundary rotation, src/chmal.f:52-154
* ``build``           face numbering (src/cmal_s.f:88-216), check_mesh
                      (src/cmal_s.f:222-277), cal_its (src/mat1.f:28-64),
                      cal_rotnf (src/mat1.f:7-23), calpre (src/shwwnl.f:350-422),
                      coriolis (src/mat1.f:69-83)

``changemaille`` and ``build`` run in the native library (spoc_amd/csrc/mesh.cpp,
C-ABI ``spoc_mesh_changemaille``, ``spoc_mesh_build``).

All index arrays are int32 and 1-based, as in the reference, laid out
column-major like the Fortran arrays: ``tri[:, e]`` are the 3 nodes of
element e, ``itm[:, e]`` its 3 faces, ``tseg[:, f]`` the 2 elements of face f.
"""
from __future__ import annotations

import numpy as np

LDM = 45


def rmal(nx: int, ny: int, xmin: float, xmax: float, ymin: float, ymax: float):
    """Structured rectangle mesh: nx*ny nodes, 2(nx-1)(ny-1) CCW triangles."""
    i = np.arange(1, nx + 1, dtype=np.float64)
    j = np.arange(1, ny + 1, dtype=np.float64)
    # node (i, j) numbered i + nx*(j-1), x varies fastest (src/rmal.f:17-25)
    xs = xmin + (i - 1.0) * (xmax - xmin) / float(nx - 1)
    ys = ymin + (j - 1.0) * (ymax - ymin) / float(ny - 1)
    x = np.tile(xs, ny)
    y = np.repeat(ys, nx)
    tris = []
    for jj in range(1, ny):
        for ii in range(1, nx):
            n1 = ii + nx * (jj - 1)
            tris.append((n1, n1 + 1, n1 + nx))
            tris.append((n1 + 1, n1 + 1 + nx, n1 + nx))
    tri = np.asfortranarray(np.array(tris, dtype=np.int32).T)
    return x, y, tri


def _lib():
    from . import load_library
    return load_library()


def _i32p(a):
    import ctypes as C
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64p(a):
    import ctypes as C
    return a.ctypes.data_as(C.POINTER(C.c_double))


def changemaille(nn: int, tri: np.ndarray) -> np.ndarray:
    """Flip the diagonal of triangles with two boundary faces, then rotate each
    boundary triangle's local numbering so that its edge 1 is the boundary
    edge (src/chmal.f:52-142; native, spoc_mesh_changemaille).  Returns the new
    connectivity."""
    tri = np.array(tri, dtype=np.int32, order="F", copy=True)
    rc = _lib().spoc_mesh_changemaille(int(nn), int(tri.shape[1]), _i32p(tri))
    if rc != 0:
        raise ValueError(f"spoc_mesh_changemaille failed ({rc})")
    return tri


def build(x: np.ndarray, y: np.ndarray, tri: np.ndarray, nc: int, f0: float = 0.0,
          beta: float = 0.0, check: bool = True) -> dict:
    """All mesh operands of cal_var, computed as the reference driver does
    (src/shwwnl.f:162-174; native, spoc_mesh_build)."""
    import ctypes as C
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    tri = np.asfortranarray(tri, dtype=np.int32)
    nn, ne = len(x), tri.shape[1]
    itm = np.zeros((3, ne), np.int32, order="F")
    its = np.zeros((3, ne), np.int32, order="F")
    tseg = np.zeros((2, 3 * ne), np.int32, order="F")
    pseg = np.zeros((2, 3 * ne), np.int32, order="F")
    ssit = np.zeros((2, 3 * ne), np.int32, order="F")
    norm = np.zeros((2, 3 * ne), order="F")
    vect = np.zeros((5, ne), order="F")
    normt = np.zeros((2, ne), order="F")
    bdl = np.zeros(ne, np.int32)
    fc = np.zeros((LDM, ne), order="F")
    deta = np.zeros(ne)
    nf = C.c_int()
    rc = _lib().spoc_mesh_build(
        nn, ne, _i32p(tri), _f64p(x), _f64p(y), int(nc), C.c_double(f0), C.c_double(beta), LDM,
        int(bool(check)), C.byref(nf), _i32p(itm), _i32p(its), _i32p(tseg), _i32p(pseg),
        _i32p(ssit), _f64p(norm), _f64p(vect), _f64p(normt), _i32p(bdl), _f64p(fc), _f64p(deta))
    if rc != 0:
        raise ValueError(f"spoc_mesh_build rejected the mesh ({rc})")
    n = nf.value
    cut = lambda a: np.asfortranarray(a[:, :n])
    return dict(nn=nn, ne=ne, nf=n, x=x, y=y, **{"in": tri}, itm=itm, its=its, tseg=cut(tseg),
                pseg=cut(pseg), ssit=cut(ssit), norm=cut(norm), vect=vect, normt=normt,
                bdl=bdl, fc=fc, deta=deta)
