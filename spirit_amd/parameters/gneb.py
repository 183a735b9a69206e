"""Parameters of the GNEB method (core/python/spirit/parameters/gneb.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_float

IMAGE_NORMAL = 0
IMAGE_CLIMBING = 1
IMAGE_FALLING = 2
IMAGE_STATIONARY = 3


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Set_N_Iterations = _fn("Parameters_GNEB_Set_N_Iterations", [_vp, _i, _i, _i])
_Set_Convergence = _fn("Parameters_GNEB_Set_Convergence", [_vp, _f, _i, _i])
_Set_Spring_Constant = _fn("Parameters_GNEB_Set_Spring_Constant", [_vp, _f, _i, _i])
_Set_Climbing_Falling = _fn("Parameters_GNEB_Set_Climbing_Falling", [_vp, _i, _i, _i])
_Set_Image_Type_Automatically = _fn("Parameters_GNEB_Set_Image_Type_Automatically", [_vp, _i])
_Set_N_Energy_Interpolations = _fn("Parameters_GNEB_Set_N_Energy_Interpolations", [_vp, _i, _i])
_Get_Convergence = _fn("Parameters_GNEB_Get_Convergence", [_vp, _i, _i], _f)
_Get_Spring_Constant = _fn("Parameters_GNEB_Get_Spring_Constant", [_vp, _i, _i], _f)
_Get_Climbing_Falling = _fn("Parameters_GNEB_Get_Climbing_Falling", [_vp, _i, _i], _i)
_Get_N_Energy_Interpolations = _fn("Parameters_GNEB_Get_N_Energy_Interpolations", [_vp, _i], _i)


def set_iterations(p_state, n_iterations, n_iterations_log, idx_chain=-1):
    _Set_N_Iterations(ctypes.c_void_p(p_state), _i(n_iterations), _i(n_iterations_log), _i(idx_chain))


def set_convergence(p_state, convergence, idx_image=-1, idx_chain=-1):
    _Set_Convergence(ctypes.c_void_p(p_state), _f(convergence), _i(idx_image), _i(idx_chain))


def set_spring_constant(p_state, spring_constant, idx_image=-1, idx_chain=-1):
    _Set_Spring_Constant(ctypes.c_void_p(p_state), _f(spring_constant), _i(idx_image), _i(idx_chain))


def set_climbing_falling(p_state, image_type, idx_image=-1, idx_chain=-1):
    _Set_Climbing_Falling(ctypes.c_void_p(p_state), _i(image_type), _i(idx_image), _i(idx_chain))


def set_image_type_automatically(p_state, idx_chain=-1):
    _Set_Image_Type_Automatically(ctypes.c_void_p(p_state), _i(idx_chain))


def set_n_energy_interpolations(p_state, n, idx_chain=-1):
    _Set_N_Energy_Interpolations(ctypes.c_void_p(p_state), _i(n), _i(idx_chain))


def get_convergence(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Convergence(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_spring_constant(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Spring_Constant(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_climbing_falling(p_state, idx_image=-1, idx_chain=-1):
    return int(_Get_Climbing_Falling(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_n_energy_interpolations(p_state, idx_chain=-1):
    return int(_Get_N_Energy_Interpolations(ctypes.c_void_p(p_state), _i(idx_chain)))
