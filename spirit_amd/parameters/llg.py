"""Parameters of the LLG method (core/python/spirit/parameters/llg.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_float
_F3 = ctypes.c_float * 3


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Set_N_Iterations = _fn("Parameters_LLG_Set_N_Iterations", [_vp, _i, _i, _i, _i])
_Set_Direct_Minimization = _fn("Parameters_LLG_Set_Direct_Minimization", [_vp, _b, _i, _i])
_Set_Convergence = _fn("Parameters_LLG_Set_Convergence", [_vp, _f, _i, _i])
_Set_Time_Step = _fn("Parameters_LLG_Set_Time_Step", [_vp, _f, _i, _i])
_Set_Damping = _fn("Parameters_LLG_Set_Damping", [_vp, _f, _i, _i])
_Set_Temperature = _fn("Parameters_LLG_Set_Temperature", [_vp, _f, _i, _i])
_Set_Output_General = _fn("Parameters_LLG_Set_Output_General", [_vp, _b, _b, _b, _i, _i])
_Get_N_Iterations = _fn("Parameters_LLG_Get_N_Iterations", [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _i, _i])
_Get_Direct_Minimization = _fn("Parameters_LLG_Get_Direct_Minimization", [_vp, _i, _i], _b)
_Get_Convergence = _fn("Parameters_LLG_Get_Convergence", [_vp, _i, _i], _f)
_Get_Time_Step = _fn("Parameters_LLG_Get_Time_Step", [_vp, _i, _i], _f)
_Get_Damping = _fn("Parameters_LLG_Get_Damping", [_vp, _i, _i], _f)
_Get_Temperature = _fn("Parameters_LLG_Get_Temperature", [_vp, _i, _i], _f)


def set_iterations(p_state, n_iterations, n_iterations_log, idx_image=-1, idx_chain=-1):
    _Set_N_Iterations(ctypes.c_void_p(p_state), _i(n_iterations), _i(n_iterations_log), _i(idx_image), _i(idx_chain))


def set_direct_minimization(p_state, use_minimization, idx_image=-1, idx_chain=-1):
    _Set_Direct_Minimization(ctypes.c_void_p(p_state), _b(use_minimization), _i(idx_image), _i(idx_chain))


def set_convergence(p_state, convergence, idx_image=-1, idx_chain=-1):
    _Set_Convergence(ctypes.c_void_p(p_state), _f(convergence), _i(idx_image), _i(idx_chain))


def set_timestep(p_state, dt, idx_image=-1, idx_chain=-1):
    _Set_Time_Step(ctypes.c_void_p(p_state), _f(dt), _i(idx_image), _i(idx_chain))


def set_damping(p_state, damping, idx_image=-1, idx_chain=-1):
    _Set_Damping(ctypes.c_void_p(p_state), _f(damping), _i(idx_image), _i(idx_chain))


_Set_Temperature_Gradient = _fn("Parameters_LLG_Set_Temperature_Gradient", [_vp, _f, ctypes.POINTER(_F3), _i, _i])
_Get_Temperature_Gradient = _fn("Parameters_LLG_Get_Temperature_Gradient",
                                [_vp, ctypes.POINTER(_f), ctypes.POINTER(_F3), _i, _i])
_Set_STT = _fn("Parameters_LLG_Set_STT", [_vp, _b, _f, ctypes.POINTER(_F3), _i, _i])
_Get_STT = _fn("Parameters_LLG_Get_STT", [_vp, ctypes.POINTER(_b), ctypes.POINTER(_f), ctypes.POINTER(_F3), _i, _i])
_Set_Non_Adiabatic_Damping = _fn("Parameters_LLG_Set_Non_Adiabatic_Damping", [_vp, _f, _i, _i])
_Get_Non_Adiabatic_Damping = _fn("Parameters_LLG_Get_Non_Adiabatic_Damping", [_vp, _i, _i], _f)


def set_temperature(p_state, temperature, gradient_inclination=0, gradient_direction=(1.0, 0.0, 0.0), idx_image=-1,
                    idx_chain=-1):
    """Homogeneous base temperature [K] and a linear gradient of `gradient_inclination` [K per
    length unit] along `gradient_direction` (core/python/spirit/parameters/llg.py set_temperature)."""
    _Set_Temperature(ctypes.c_void_p(p_state), _f(temperature), _i(idx_image), _i(idx_chain))
    d = _F3(*gradient_direction)
    _Set_Temperature_Gradient(ctypes.c_void_p(p_state), _f(gradient_inclination), ctypes.byref(d), _i(idx_image),
                              _i(idx_chain))


def set_stt(p_state, use_gradient, magnitude, direction, idx_image=-1, idx_chain=-1):
    """Spin-transfer torque: the spin derivative along `direction` (`use_gradient`) or the
    monolayer approximation with polarisation `direction`, of strength `magnitude`."""
    d = _F3(*direction)
    _Set_STT(ctypes.c_void_p(p_state), _b(use_gradient), _f(magnitude), ctypes.byref(d), _i(idx_image), _i(idx_chain))


def set_non_adiabatic_damping(p_state, beta, idx_image=-1, idx_chain=-1):
    """The non-adiabatic parameter beta of the spin-transfer torque."""
    _Set_Non_Adiabatic_Damping(ctypes.c_void_p(p_state), _f(beta), _i(idx_image), _i(idx_chain))


def set_output_general(p_state, any=True, initial=False, final=False, idx_image=-1, idx_chain=-1):
    _Set_Output_General(ctypes.c_void_p(p_state), _b(any), _b(initial), _b(final), _i(idx_image), _i(idx_chain))


def get_iterations(p_state, idx_image=-1, idx_chain=-1):
    n, nl = _i(), _i()
    _Get_N_Iterations(ctypes.c_void_p(p_state), ctypes.byref(n), ctypes.byref(nl), _i(idx_image), _i(idx_chain))
    return int(n.value), int(nl.value)


def get_direct_minimization(p_state, idx_image=-1, idx_chain=-1):
    return bool(_Get_Direct_Minimization(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_convergence(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Convergence(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_timestep(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Time_Step(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_damping(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Damping(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def get_temperature(p_state, idx_image=-1, idx_chain=-1):
    """(base temperature, gradient inclination, gradient direction)."""
    T = float(_Get_Temperature(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))
    incl, d = _f(), _F3()
    _Get_Temperature_Gradient(ctypes.c_void_p(p_state), ctypes.byref(incl), ctypes.byref(d), _i(idx_image),
                              _i(idx_chain))
    return T, float(incl.value), [float(x) for x in d]


def get_stt(p_state, idx_image=-1, idx_chain=-1):
    """(magnitude, direction, use_gradient) of the spin-transfer torque."""
    g, m, d = _b(), _f(), _F3()
    _Get_STT(ctypes.c_void_p(p_state), ctypes.byref(g), ctypes.byref(m), ctypes.byref(d), _i(idx_image),
             _i(idx_chain))
    return float(m.value), [float(x) for x in d], bool(g.value)


def get_non_adiabatic_damping(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Non_Adiabatic_Damping(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


_Set_Output_Configuration = _fn("Parameters_LLG_Set_Output_Configuration", [_vp, _b, _b, _i, _i, _i])
_Get_Output_Configuration = _fn("Parameters_LLG_Get_Output_Configuration",
                                [_vp, ctypes.POINTER(_b), ctypes.POINTER(_b), ctypes.POINTER(_i), _i, _i])


def set_output_configuration(p_state, step=False, archive=True, filetype=3, idx_image=-1, idx_chain=-1):
    """Spin configuration files per log step: one file each (`step`) and/or one archive file
    with a segment appended per log step (`archive`), in OVF format `filetype`
    (core/python/spirit/parameters/llg.py set_output_configuration)."""
    _Set_Output_Configuration(_vp(p_state), _b(step), _b(archive), _i(filetype), _i(idx_image), _i(idx_chain))


def get_output_configuration(p_state, idx_image=-1, idx_chain=-1):
    step, archive, ft = _b(), _b(), _i()
    _Get_Output_Configuration(_vp(p_state), ctypes.byref(step), ctypes.byref(archive), ctypes.byref(ft),
                              _i(idx_image), _i(idx_chain))
    return step.value, archive.value, ft.value
