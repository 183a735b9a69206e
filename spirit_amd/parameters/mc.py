"""Monte Carlo parameters; mirrors the reference's ``spirit.parameters.mc``
(core/python/spirit/parameters/mc.py) for the entry points of ``Parameters_MC.h``."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_float


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Set_N_Iterations = _fn("Parameters_MC_Set_N_Iterations", [_vp, _i, _i, _i, _i])
_Set_Temperature = _fn("Parameters_MC_Set_Temperature", [_vp, _f, _i, _i])
_Set_Metropolis_Cone = _fn("Parameters_MC_Set_Metropolis_Cone", [_vp, _b, _f, _b, _f, _i, _i])
_Set_Random_Sample = _fn("Parameters_MC_Set_Random_Sample", [_vp, _b, _i, _i])
_Get_N_Iterations = _fn("Parameters_MC_Get_N_Iterations", [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _i, _i])
_Get_Temperature = _fn("Parameters_MC_Get_Temperature", [_vp, _i, _i], _f)
_Get_Metropolis_Cone = _fn("Parameters_MC_Get_Metropolis_Cone",
                           [_vp, ctypes.POINTER(_b), ctypes.POINTER(_f), ctypes.POINTER(_b), ctypes.POINTER(_f), _i, _i])
_Get_Random_Sample = _fn("Parameters_MC_Get_Random_Sample", [_vp, _i, _i], _b)


def set_iterations(p_state, n_iterations, n_iterations_log, idx_image=-1, idx_chain=-1):
    _Set_N_Iterations(ctypes.c_void_p(p_state), n_iterations, n_iterations_log, idx_image, idx_chain)


def set_temperature(p_state, temperature, idx_image=-1, idx_chain=-1):
    _Set_Temperature(ctypes.c_void_p(p_state), temperature, idx_image, idx_chain)


def set_metropolis_cone(p_state, use_cone=True, cone_angle=30, use_adaptive_cone=True, target_acceptance_ratio=0.5,
                        idx_image=-1, idx_chain=-1):
    _Set_Metropolis_Cone(ctypes.c_void_p(p_state), use_cone, cone_angle, use_adaptive_cone, target_acceptance_ratio,
                         idx_image, idx_chain)


def set_random_sample(p_state, random_sample=True, idx_image=-1, idx_chain=-1):
    _Set_Random_Sample(ctypes.c_void_p(p_state), random_sample, idx_image, idx_chain)


def get_iterations(p_state, idx_image=-1, idx_chain=-1):
    n, n_log = _i(), _i()
    _Get_N_Iterations(ctypes.c_void_p(p_state), ctypes.byref(n), ctypes.byref(n_log), idx_image, idx_chain)
    return int(n.value), int(n_log.value)


def get_temperature(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Temperature(ctypes.c_void_p(p_state), idx_image, idx_chain))


def get_metropolis_cone(p_state, idx_image=-1, idx_chain=-1):
    cone, angle, adaptive, target = _b(), _f(), _b(), _f()
    _Get_Metropolis_Cone(ctypes.c_void_p(p_state), ctypes.byref(cone), ctypes.byref(angle), ctypes.byref(adaptive),
                         ctypes.byref(target), idx_image, idx_chain)
    return bool(cone.value), float(angle.value), bool(adaptive.value), float(target.value)


def get_random_sample(p_state, idx_image=-1, idx_chain=-1):
    return bool(_Get_Random_Sample(ctypes.c_void_p(p_state), idx_image, idx_chain))
