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


_Set_Spring_Force_Ratio = _fn("Parameters_GNEB_Set_Spring_Force_Ratio", [_vp, _f, _i])
_Set_Path_Shortening_Constant = _fn("Parameters_GNEB_Set_Path_Shortening_Constant", [_vp, _f, _i])
_Set_Moving_Endpoints = _fn("Parameters_GNEB_Set_Moving_Endpoints", [_vp, _b, _i])
_Set_Translating_Endpoints = _fn("Parameters_GNEB_Set_Translating_Endpoints", [_vp, _b, _i])
_Set_Equilibrium_Delta_Rx = _fn("Parameters_GNEB_Set_Equilibrium_Delta_Rx", [_vp, _f, _f, _i])
_Get_N_Iterations = _fn("Parameters_GNEB_Get_N_Iterations", [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _i])
_Get_Spring_Force_Ratio = _fn("Parameters_GNEB_Get_Spring_Force_Ratio", [_vp, _i], _f)
_Get_Path_Shortening_Constant = _fn("Parameters_GNEB_Get_Path_Shortening_Constant", [_vp, _i], _f)
_Get_Moving_Endpoints = _fn("Parameters_GNEB_Get_Moving_Endpoints", [_vp, _i], _b)
_Get_Translating_Endpoints = _fn("Parameters_GNEB_Get_Translating_Endpoints", [_vp, _i], _b)
_Get_Equilibrium_Delta_Rx = _fn("Parameters_GNEB_Get_Equilibrium_Delta_Rx",
                                [_vp, ctypes.POINTER(_f), ctypes.POINTER(_f), _i])


def set_spring_force(p_state, spring_constant=1, ratio=0, idx_image=-1, idx_chain=-1):
    """Spring constant and the energy/Rx ratio of the spring force."""
    _Set_Spring_Constant(ctypes.c_void_p(p_state), _f(spring_constant), _i(idx_image), _i(idx_chain))
    _Set_Spring_Force_Ratio(ctypes.c_void_p(p_state), _f(ratio), _i(idx_chain))


def set_path_shortening_constant(p_state, shortening_constant, idx_image=-1, idx_chain=-1):
    _Set_Path_Shortening_Constant(ctypes.c_void_p(p_state), _f(shortening_constant), _i(idx_chain))


def set_moving_endpoints(p_state, moving_endpoints, idx_chain=-1):
    _Set_Moving_Endpoints(ctypes.c_void_p(p_state), _b(moving_endpoints), _i(idx_chain))


def set_translating_endpoints(p_state, translating_endpoints, idx_chain=-1):
    _Set_Translating_Endpoints(ctypes.c_void_p(p_state), _b(translating_endpoints), _i(idx_chain))


def set_equilibrium_delta_Rx(p_state, delta_Rx_left, delta_Rx_right, idx_chain=-1):
    _Set_Equilibrium_Delta_Rx(ctypes.c_void_p(p_state), _f(delta_Rx_left), _f(delta_Rx_right), _i(idx_chain))


def get_iterations(p_state, idx_image=-1, idx_chain=-1):
    n, n_log = _i(), _i()
    _Get_N_Iterations(ctypes.c_void_p(p_state), ctypes.byref(n), ctypes.byref(n_log), _i(idx_chain))
    return int(n.value), int(n_log.value)


def get_spring_force(p_state, idx_image=-1, idx_chain=-1):
    return (float(_Get_Spring_Constant(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))),
            float(_Get_Spring_Force_Ratio(ctypes.c_void_p(p_state), _i(idx_chain))))


def get_path_shortening_constant(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Path_Shortening_Constant(ctypes.c_void_p(p_state), _i(idx_chain)))


def get_moving_endpoints(p_state, idx_chain=-1):
    return bool(_Get_Moving_Endpoints(ctypes.c_void_p(p_state), _i(idx_chain)))


def get_translating_endpoints(p_state, idx_chain=-1):
    return bool(_Get_Translating_Endpoints(ctypes.c_void_p(p_state), _i(idx_chain)))


def get_equilibrium_delta_Rx(p_state, idx_chain=-1):
    a, b = _f(), _f()
    _Get_Equilibrium_Delta_Rx(ctypes.c_void_p(p_state), ctypes.byref(a), ctypes.byref(b), _i(idx_chain))
    return float(a.value), float(b.value)
