"""Geometry: the lattice (core/python/spirit/geometry.py)."""
import ctypes

import numpy as np

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b, _f, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_float, ctypes.c_double
_F3 = ctypes.c_float * 3

BRAVAIS_LATTICE_IRREGULAR = 0
BRAVAIS_LATTICE_RECTILINEAR = 1
BRAVAIS_LATTICE_SC = 2
BRAVAIS_LATTICE_HEX2D = 3
BRAVAIS_LATTICE_HEX2D_60 = 4
BRAVAIS_LATTICE_HEX2D_120 = 5
BRAVAIS_LATTICE_HCP = 6
BRAVAIS_LATTICE_BCC = 7
BRAVAIS_LATTICE_FCC = 8


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Set_Type = _fn("Geometry_Set_Bravais_Lattice_Type", [_vp, _i])
_Set_N_Cells = _fn("Geometry_Set_N_Cells", [_vp, ctypes.POINTER(_i)])
_Set_mu_s = _fn("Geometry_Set_mu_s", [_vp, _f, _i, _i])
_Set_Lattice_Constant = _fn("Geometry_Set_Lattice_Constant", [_vp, _f])
_Set_Bravais_Vectors = _fn("Geometry_Set_Bravais_Vectors", [_vp, _F3, _F3, _F3])
_Get_NOS = _fn("Geometry_Get_NOS", [_vp], _i)
_Get_Positions = _fn("Geometry_Get_Positions", [_vp, _i, _i], ctypes.POINTER(_d))
_Get_Bounds = _fn("Geometry_Get_Bounds", [_vp, _F3, _F3, _i, _i])
_Get_Center = _fn("Geometry_Get_Center", [_vp, _F3, _i, _i])
_Get_Type = _fn("Geometry_Get_Bravais_Lattice_Type", [_vp, _i, _i], _i)
_Get_Bravais_Vectors = _fn("Geometry_Get_Bravais_Vectors", [_vp, _F3, _F3, _F3, _i, _i])
_Get_N_Cells = _fn("Geometry_Get_N_Cells", [_vp, ctypes.POINTER(_i), _i, _i])
_Get_Dimensionality = _fn("Geometry_Get_Dimensionality", [_vp, _i, _i], _i)
_Get_N_Cell_Atoms = _fn("Geometry_Get_N_Cell_Atoms", [_vp, _i, _i], _i)
_Get_mu_s = _fn("Geometry_Get_mu_s", [_vp, ctypes.POINTER(_f), _i, _i])


def set_bravais_lattice_type(p_state, lattice_type, idx_image=-1, idx_chain=-1):
    _Set_Type(ctypes.c_void_p(p_state), _i(lattice_type))


def set_n_cells(p_state, n_cells=[1, 1, 1], idx_image=-1, idx_chain=-1):
    _Set_N_Cells(ctypes.c_void_p(p_state), (_i * 3)(*n_cells))


def set_mu_s(p_state, mu_s, idx_image=-1, idx_chain=-1):
    _Set_mu_s(ctypes.c_void_p(p_state), _f(mu_s), _i(idx_image), _i(idx_chain))


def set_lattice_constant(p_state, lattice_constant, idx_image=-1, idx_chain=-1):
    _Set_Lattice_Constant(ctypes.c_void_p(p_state), _f(lattice_constant))


def set_bravais_vectors(p_state, ta=[1.0, 0.0, 0.0], tb=[0.0, 1.0, 0.0], tc=[0.0, 0.0, 1.0], idx_image=-1,
                        idx_chain=-1):
    _Set_Bravais_Vectors(ctypes.c_void_p(p_state), _F3(*ta), _F3(*tb), _F3(*tc))


def get_nos(p_state):
    return int(_Get_NOS(ctypes.c_void_p(p_state)))


def get_positions(p_state, idx_image=-1, idx_chain=-1):
    nos = get_nos(p_state)
    ptr = _Get_Positions(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))
    return np.ctypeslib.as_array(ptr, shape=(nos, 3))


def get_bounds(p_state, idx_image=-1, idx_chain=-1):
    lo, hi = _F3(), _F3()
    _Get_Bounds(ctypes.c_void_p(p_state), lo, hi, _i(idx_image), _i(idx_chain))
    return [float(x) for x in lo], [float(x) for x in hi]


def get_center(p_state, idx_image=-1, idx_chain=-1):
    c = _F3()
    _Get_Center(ctypes.c_void_p(p_state), c, _i(idx_image), _i(idx_chain))
    return [float(x) for x in c]


def get_bravais_lattice_type(p_state, idx_image=-1, idx_chain=-1):
    return int(_Get_Type(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_bravais_vectors(p_state, idx_image=-1, idx_chain=-1):
    a, b, c = _F3(), _F3(), _F3()
    _Get_Bravais_Vectors(ctypes.c_void_p(p_state), a, b, c, _i(idx_image), _i(idx_chain))
    return [float(x) for x in a], [float(x) for x in b], [float(x) for x in c]


def get_n_cells(p_state, idx_image=-1, idx_chain=-1):
    n = (_i * 3)()
    _Get_N_Cells(ctypes.c_void_p(p_state), n, _i(idx_image), _i(idx_chain))
    return [int(x) for x in n]


def get_dimensionality(p_state, idx_image=-1, idx_chain=-1):
    return int(_Get_Dimensionality(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_n_cell_atoms(p_state, idx_image=-1, idx_chain=-1):
    return int(_Get_N_Cell_Atoms(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_mu_s(p_state, idx_image=-1, idx_chain=-1):
    n = get_n_cell_atoms(p_state, idx_image, idx_chain)
    arr = (_f * n)()
    _Get_mu_s(ctypes.c_void_p(p_state), arr, _i(idx_image), _i(idx_chain))
    return [float(x) for x in arr]
