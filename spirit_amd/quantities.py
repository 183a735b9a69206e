"""Quantities (core/python/spirit/quantities.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
_F3 = ctypes.c_float * 3

_Get_Magnetization = _spirit.Quantity_Get_Magnetization
_Get_Magnetization.argtypes = [_vp, _F3, _i, _i]
_Get_Magnetization.restype = None
_Get_Average_Spin = _spirit.Quantity_Get_Average_Spin
_Get_Average_Spin.argtypes = [_vp, _F3, _i, _i]
_Get_Average_Spin.restype = None
_Get_Topological_Charge = _spirit.Quantity_Get_Topological_Charge
_Get_Topological_Charge.argtypes = [_vp, _i, _i]
_Get_Topological_Charge.restype = _f


def get_magnetization(p_state, idx_image=-1, idx_chain=-1):
    m = _F3()
    _Get_Magnetization(ctypes.c_void_p(p_state), m, _i(idx_image), _i(idx_chain))
    return [float(x) for x in m]


def get_average_spin(p_state, idx_image=-1, idx_chain=-1):
    s = _F3()
    _Get_Average_Spin(ctypes.c_void_p(p_state), s, _i(idx_image), _i(idx_chain))
    return [float(x) for x in s]


def get_topological_charge(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Topological_Charge(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


_Get_Topological_Charge_Density = _spirit.Quantity_Get_Topological_Charge_Density
_Get_Topological_Charge_Density.argtypes = [_vp, ctypes.POINTER(_f), ctypes.POINTER(_i), _i, _i]
_Get_Topological_Charge_Density.restype = _i


def get_topological_charge_density(p_state, idx_image=-1, idx_chain=-1):
    """(charge of each triangle, the triangles' three spin indices) of a 2D lattice with one basis
    atom; empty lists otherwise."""
    n = _Get_Topological_Charge_Density(ctypes.c_void_p(p_state), None, None, _i(idx_image), _i(idx_chain))
    dens = (_f * max(n, 1))()
    tri = (_i * max(3 * n, 1))()
    _Get_Topological_Charge_Density(ctypes.c_void_p(p_state), dens, tri, _i(idx_image), _i(idx_chain))
    return [float(dens[k]) for k in range(n)], [[int(tri[3 * k]), int(tri[3 * k + 1]), int(tri[3 * k + 2])]
                                                for k in range(n)]
