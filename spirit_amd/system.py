"""System: one image's spins, effective field and energy (core/python/spirit/system.py)."""
import ctypes

import numpy as np

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b, _f, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_float, ctypes.c_double


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Get_Index = _fn("System_Get_Index", [_vp], _i)
_Get_NOS = _fn("System_Get_NOS", [_vp, _i, _i], _i)
_Get_Spin_Directions = _fn("System_Get_Spin_Directions", [_vp, _i, _i], ctypes.POINTER(_d))
_Get_Effective_Field = _fn("System_Get_Effective_Field", [_vp, _i, _i], ctypes.POINTER(_d))
_Get_Energy = _fn("System_Get_Energy", [_vp, _i, _i], _f)
_Get_Rx = _fn("System_Get_Rx", [_vp, _i, _i], _f)
_Update_Data = _fn("System_Update_Data", [_vp, _i, _i])
_Print_Energy_Array = _fn("System_Print_Energy_Array", [_vp, _i, _i])
_Get_Energy_Array_Names = _fn("System_Get_Energy_Array_Names", [_vp, ctypes.c_char_p, _i, _i], _i)
_Get_Energy_Array = _fn("System_Get_Energy_Array", [_vp, ctypes.POINTER(_f), _b, _i, _i], _i)


def get_index(p_state):
    return int(_Get_Index(ctypes.c_void_p(p_state)))


def get_nos(p_state, idx_image=-1, idx_chain=-1):
    return int(_Get_NOS(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_spin_directions(p_state, idx_image=-1, idx_chain=-1):
    """A zero-copy (nos, 3) view of the image's spins; writes are seen by the next method start."""
    nos = get_nos(p_state, idx_image, idx_chain)
    ptr = _Get_Spin_Directions(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))
    return np.ctypeslib.as_array(ptr, shape=(nos, 3))


def get_effective_field(p_state, idx_image=-1, idx_chain=-1):
    nos = get_nos(p_state, idx_image, idx_chain)
    ptr = _Get_Effective_Field(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))
    return np.ctypeslib.as_array(ptr, shape=(nos, 3))


def get_energy(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Energy(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_rx(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Rx(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def update_data(p_state, idx_image=-1, idx_chain=-1):
    _Update_Data(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def print_energy_array(p_state, idx_image=-1, idx_chain=-1):
    _Print_Energy_Array(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def get_energy_array_names(p_state, idx_image=-1, idx_chain=-1):
    """Names of the energy contributions (System_Get_Energy_Array_Names)."""
    n = _Get_Energy_Array_Names(ctypes.c_void_p(p_state), None, _i(idx_image), _i(idx_chain))
    buf = ctypes.create_string_buffer(n + 1)
    _Get_Energy_Array_Names(ctypes.c_void_p(p_state), buf, _i(idx_image), _i(idx_chain))
    names = buf.value.decode()
    return names.split("|") if names else []


def get_energy_contributions(p_state, divide_by_nspins=True, idx_image=-1, idx_chain=-1):
    """{name: energy} of the contributions of the last update_data, as the reference's
    system.get_energy_contributions."""
    names = get_energy_array_names(p_state, idx_image, idx_chain)
    values = get_energy_array(p_state, divide_by_nspins, idx_image, idx_chain)
    return dict(zip(names, values))


def get_energy_array(p_state, divide_by_nspins=True, idx_image=-1, idx_chain=-1):
    """Energy contributions of the last update_data (System_Get_Energy_Array)."""
    n = _Get_Energy_Array(ctypes.c_void_p(p_state), None, _b(divide_by_nspins), _i(idx_image), _i(idx_chain))
    arr = (_f * max(n, 1))()
    _Get_Energy_Array(ctypes.c_void_p(p_state), arr, _b(divide_by_nspins), _i(idx_image), _i(idx_chain))
    return [float(arr[k]) for k in range(n)]
