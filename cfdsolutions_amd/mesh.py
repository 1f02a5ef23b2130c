"""Mesh generation and median-dual preprocessing (libcfd_mesh.so).

Produces the same arrays as the reference preprocessing
(Euler3D/src/MeshIO.cpp, Connectivity_Maps.cpp, Area_Volume.cpp), bit for bit,
in O(N+E).  Boundary tags: 1 x-min, 2 x-max, 3 y-min (bump wall), 4 y-max,
5 z-min, 6 z-max.
"""
import ctypes
import os

import numpy as np

from . import _build

BC_EULER_WALL = 0
BC_FAR_FIELD = 1
TAG_BUMP_WALL = 3


class _Info(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("n_node", "n_tri", "n_tet", "n_edge", "n_bedge", "n_n2n")]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = _build.lib_path("libcfd_mesh.so")
        if not os.path.exists(path):
            raise RuntimeError("libcfd_mesh.so is not built; run __graft_entry__.build()")
        _LIB = ctypes.CDLL(path)
        _LIB.cfdm_last_error.restype = ctypes.c_char_p
    return _LIB


def _check(rc):
    if rc != 0:
        raise RuntimeError("cfd_mesh error %d: %s" % (rc, _lib().cfdm_last_error().decode()))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Mesh:
    """A preprocessed tetrahedral mesh: the arrays the residual path consumes."""

    def __init__(self, handle):
        lib = _lib()
        _check(lib.cfdm_preprocess(handle))
        info = _Info()
        _check(lib.cfdm_get_info(handle, ctypes.byref(info)))
        n, e, b = info.n_node, info.n_edge, info.n_bedge
        self.n_node, self.n_edge, self.n_bedge, self.n_tet, self.n_tri = n, e, b, info.n_tet, info.n_tri
        self.xyz = np.zeros((n, 3))
        self.tet = np.zeros((info.n_tet, 4), np.int32)
        self.tri = np.zeros((info.n_tri, 3), np.int32)
        self.tri_tag = np.zeros(info.n_tri, np.int32)
        self.edge = np.zeros((e, 2), np.int32)
        self.edge_area = np.zeros((e, 3))
        self.bedge = np.zeros((b, 2), np.int32)
        self.bedge_area = np.zeros((b, 3))
        self.bedge_tag = np.zeros(b, np.int32)
        self.vol = np.zeros(n)
        self.n2n_ia = np.zeros(n + 1, np.int32)
        self.n2n_ja = np.zeros(info.n_n2n, np.int32)
        _check(lib.cfdm_get_arrays(handle, _p(self.xyz), _p(self.tet), _p(self.tri), _p(self.tri_tag), _p(self.edge),
                                   _p(self.edge_area), _p(self.bedge), _p(self.bedge_area), _p(self.bedge_tag),
                                   _p(self.vol), _p(self.n2n_ia), _p(self.n2n_ja)))
        # BC deck of the bump channel: x-min/x-max far field, all other walls.
        self.bedge_type = np.where(self.bedge_tag <= 2, BC_FAR_FIELD, BC_EULER_WALL).astype(np.int32)

    @property
    def n_total(self):
        return self.n_node + self.n_bedge


def box(nx, ny, nz, lx=3.0, ly=1.0, lz=1.0, bump_h=0.1, jitter=0.0, seed=20261202, reorder=True,
        ugrid_path=None):
    """Generate (and optionally write as UGRID) a Kuhn-split hex box; returns Mesh."""
    lib = _lib()
    h = ctypes.c_void_p()
    _check(lib.cfdm_box_generate(int(nx), int(ny), int(nz), ctypes.c_double(lx), ctypes.c_double(ly),
                                 ctypes.c_double(lz), ctypes.c_double(bump_h), ctypes.c_double(jitter),
                                 ctypes.c_uint64(seed), int(bool(reorder)), ctypes.byref(h)))
    try:
        if ugrid_path:
            _check(lib.cfdm_write_ugrid(h, str(ugrid_path).encode()))
            # round trip through the file so the mesh is exactly what a reader sees
            h2 = ctypes.c_void_p()
            _check(lib.cfdm_read_ugrid(str(ugrid_path).encode(), ctypes.byref(h2)))
            lib.cfdm_free(h)
            h = h2
        return Mesh(h)
    finally:
        lib.cfdm_free(h)


class Content:
    """UGRID content of a mesh before preprocessing: xyz [N,3], tri [NT,3] in UGRID winding, tri_tag [NT],
    tet [NC,4] (the input of cfd_preprocess)."""

    def __init__(self, handle):
        lib = _lib()
        info = _Info()
        _check(lib.cfdm_get_info(handle, ctypes.byref(info)))
        self.n_node, self.n_tri, self.n_tet = info.n_node, info.n_tri, info.n_tet
        self.xyz = np.zeros((info.n_node, 3))
        self.tet = np.zeros((info.n_tet, 4), np.int32)
        self.tri = np.zeros((info.n_tri, 3), np.int32)
        self.tri_tag = np.zeros(info.n_tri, np.int32)
        _check(lib.cfdm_get_arrays(handle, _p(self.xyz), _p(self.tet), _p(self.tri), _p(self.tri_tag), None, None,
                                   None, None, None, None, None, None))


def box_content(nx, ny, nz, lx=3.0, ly=1.0, lz=1.0, bump_h=0.1, jitter=0.0, seed=20261202, reorder=True):
    """The UGRID content box() would preprocess, unpreprocessed."""
    lib = _lib()
    h = ctypes.c_void_p()
    _check(lib.cfdm_box_generate(int(nx), int(ny), int(nz), ctypes.c_double(lx), ctypes.c_double(ly),
                                 ctypes.c_double(lz), ctypes.c_double(bump_h), ctypes.c_double(jitter),
                                 ctypes.c_uint64(seed), int(bool(reorder)), ctypes.byref(h)))
    try:
        return Content(h)
    finally:
        lib.cfdm_free(h)


def read_ugrid(path):
    lib = _lib()
    h = ctypes.c_void_p()
    _check(lib.cfdm_read_ugrid(str(path).encode(), ctypes.byref(h)))
    try:
        return Mesh(h)
    finally:
        lib.cfdm_free(h)


def bump_channel(nx, ny, nz, bump_h=0.1, **kw):
    """The bump channel of the configurations: [0,3]x[0,1]x[0,1], arc of height h over x in [1,2]."""
    return box(nx, ny, nz, 3.0, 1.0, 1.0, bump_h=bump_h, **kw)


def jittered_cube(n, jitter=0.2, seed=20261202, **kw):
    """Unit cube n^3 hexes, interior nodes jittered by U(-jitter, jitter)*h (residual microbenchmark)."""
    return box(n, n, n, 1.0, 1.0, 1.0, bump_h=0.0, jitter=jitter, seed=seed, **kw)


def write_restart(path, q, n_node, n_ghost, iteration=0):
    """Restart file of a conservative state q [5][n_node + n_ghost] (Restart_Writer layout)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    if q.size != 5 * (n_node + n_ghost):
        raise ValueError("state has %d values, expected %d" % (q.size, 5 * (n_node + n_ghost)))
    _check(_lib().cfdm_restart_write(str(path).encode(), int(iteration), int(n_node), int(n_ghost), _p(q)))


def read_restart(path, n_node, n_ghost):
    """Reads a restart file; returns (iteration, q [5][n_node + n_ghost])."""
    q = np.zeros((5, n_node + n_ghost))
    it = ctypes.c_int32(0)
    _check(_lib().cfdm_restart_read(str(path).encode(), int(n_node), int(n_ghost), ctypes.byref(it), _p(q)))
    return int(it.value), q


def write_vtk(ugrid_path, vtk_path, q, gamma=1.4, r_gas=None):
    """VTK solution file (VTK_Writer layout) of the mesh in a UGRID file and a conservative state q [5][>= N]."""
    lib = _lib()
    h = ctypes.c_void_p()
    _check(lib.cfdm_read_ugrid(str(ugrid_path).encode(), ctypes.byref(h)))
    try:
        _check(lib.cfdm_preprocess(h))
        q = np.ascontiguousarray(q, dtype=np.float64)
        r = 1.0 / gamma if r_gas is None else r_gas
        _check(lib.cfdm_vtk_write(h, str(vtk_path).encode(), _p(q), int(q.shape[1]), ctypes.c_double(gamma),
                                  ctypes.c_double(r)))
    finally:
        lib.cfdm_free(h)

