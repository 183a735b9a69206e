"""Transitions: initial guesses for a chain (core/python/spirit/transition.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float

_Homogeneous = _spirit.Transition_Homogeneous
_Homogeneous.argtypes = [_vp, _i, _i, _i]
_Homogeneous.restype = None
_Insert = _spirit.Transition_Homogeneous_Insert_Interpolated
_Insert.argtypes = [_vp, _i, _i]
_Insert.restype = None
_Noise = _spirit.Transition_Add_Noise_Temperature
_Noise.argtypes = [_vp, _f, _i, _i, _i]
_Noise.restype = None


def homogeneous(p_state, idx_1, idx_2, idx_chain=-1):
    _Homogeneous(ctypes.c_void_p(p_state), _i(idx_1), _i(idx_2), _i(idx_chain))


def homogeneous_insert_interpolated(p_state, n_interpolate, idx_chain=-1):
    _Insert(ctypes.c_void_p(p_state), _i(n_interpolate), _i(idx_chain))


def add_noise(p_state, temperature, idx_1, idx_2, idx_chain=-1):
    _Noise(ctypes.c_void_p(p_state), _f(temperature), _i(idx_1), _i(idx_2), _i(idx_chain))
