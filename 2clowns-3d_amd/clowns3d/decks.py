"""The reference's case files through the C-ABI (clowns_deck_*).

``read_case`` reads a 2CLOWNS-3D case -- grid file (".xyz" or ".xyzn"),
data file, optional open-ratio file -- (src2LOWNS/data_io.F90:87-364) into a
``configs.Grid`` in the box layout the assembly and the solver take, plus the
data file's pressure.  ``write_data`` writes a data file in output_data's
format (data_io.F90:708-783), e.g. with the solved pressure, and
``write_grid`` a sparse grid file in output_xyzn's (data_io.F90:1505-1560).
"""
import ctypes

import numpy as np

from . import _lib as L
from . import configs


class DeckDesc(ctypes.Structure):
    _fields_ = [("is_", ctypes.c_int32), ("js", ctypes.c_int32), ("ks", ctypes.c_int32),
                ("ie", ctypes.c_int32), ("je", ctypes.c_int32), ("ke", ctypes.c_int32),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("sparse", ctypes.c_int32), ("big_endian", ctypes.c_int32),
                ("inne", ctypes.c_int32), ("inne2", ctypes.c_int32), ("has_dt0", ctypes.c_int32),
                ("t0", ctypes.c_double), ("g", ctypes.c_double), ("nu", ctypes.c_double),
                ("dt0", ctypes.c_double)]


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _b(path):
    return None if path is None else str(path).encode()


def read_case(grid_path, data_path, fdata_path=None):
    """Returns (grid, p, desc): grid a configs.Grid with solid, F, un, fb, a,
    xf (low-face coordinates) and dx (set_dx of the grid file's
    coordinates), p the data file's pressure per box cell, desc the
    clowns_deck_desc (extents, t0, g, nu, dt0, byte order)."""
    d = DeckDesc()
    L.check(L.lib().clowns_deck_open(_b(grid_path), _b(data_path), ctypes.byref(d)), "clowns_deck_open")
    g = configs.Grid(d.nx, d.ny, d.nz)
    X, Y, Z = g.X, g.Y, g.Z
    nc = g.ncell
    solid = np.zeros(nc, np.int8)
    F = np.zeros(nc)
    un = np.zeros((3, nc))
    p = np.zeros(nc)
    a = np.zeros((3, nc))
    fb = np.zeros(nc)
    x = np.zeros((3, d.stride))
    dx = np.zeros((3, d.stride))
    L.check(L.lib().clowns_deck_read(_b(grid_path), _b(data_path), _b(fdata_path), ctypes.byref(d), _p(solid),
                                     _p(F), _p(un), _p(p), _p(a), _p(fb), _p(x), _p(dx)), "clowns_deck_read")
    g.solid[:] = solid.reshape(Z, Y, X)
    g.F[:] = F.reshape(Z, Y, X)
    g.un[:] = un.reshape(3, Z, Y, X)
    g.a[:] = a.reshape(3, Z, Y, X)
    g.fb[:] = fb.reshape(Z, Y, X)
    g.xf[:] = x
    g.dx[:] = dx
    g.h = tuple(float(dx[l, 1]) for l in range(3))
    if d.has_dt0 and d.dt0 > 0:
        g.dt = d.dt0
    g.name = str(data_path)
    return g, p.reshape(Z, Y, X), d


def faces(grid_path, desc):
    """The grid file's faces X(l, ls-2 .. le+1) as stored, shaped
    (3, stride+2), zero past each axis' n+4 values."""
    out = np.zeros((3, desc.stride + 2))
    L.check(L.lib().clowns_deck_faces(_b(grid_path), ctypes.byref(desc), _p(out)), "clowns_deck_faces")
    return out


def write_grid(path, desc, solid, face_arrays, big_endian=True):
    """A sparse (.xyzn) grid file: the extents of desc, the faces (as
    ``faces`` returns them) and every solid box cell."""
    solid = np.ascontiguousarray(solid, dtype=np.int8).ravel()
    fa = np.ascontiguousarray(face_arrays, dtype=np.float64)
    if fa.shape != (3, desc.stride + 2):
        raise ValueError(f"faces must be shaped (3, {desc.stride + 2})")
    L.check(L.lib().clowns_deck_write_grid(_b(path), ctypes.byref(desc), _p(solid), _p(fa),
                                           1 if big_endian else 0), "clowns_deck_write_grid")


def write_data(path, desc, nf, nb, un, F, p, t, dt, big_endian=True):
    """A data file in output_data's format: nf/nb the cell flags per box cell
    (clowns3d.classify), un (3, ncell), F and p per box cell."""
    nf = np.ascontiguousarray(nf, dtype=np.int8).ravel()
    nb = np.ascontiguousarray(nb, dtype=np.int8).ravel()
    un = np.ascontiguousarray(un, dtype=np.float64).reshape(3, -1)
    F = np.ascontiguousarray(F, dtype=np.float64).ravel()
    p = np.ascontiguousarray(p, dtype=np.float64).ravel()
    L.check(L.lib().clowns_deck_write_data(_b(path), ctypes.byref(desc), _p(nf), _p(nb), _p(un), _p(F), _p(p),
                                           float(t), float(dt), 1 if big_endian else 0), "clowns_deck_write_data")
