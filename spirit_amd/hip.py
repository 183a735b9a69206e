"""Engine-only extensions of the C API (include/Spirit/Spirit_HIP.h)."""
import ctypes

import numpy as np

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_PD = ctypes.POINTER(_d)
_F3 = ctypes.c_float * 3


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Set_Host_Sync = _fn("Spirit_HIP_Set_Host_Sync", [_vp, _i])
_Set_Thermal_Noise = _fn("Spirit_HIP_Set_Thermal_Noise", [_vp, _PD, _i, _i, _i])
_Set_Site_Anisotropy = _fn("Spirit_HIP_Hamiltonian_Set_Site_Anisotropy", [_vp, _PD, _F3, _i, _i])
_Gradient_and_Energy = _fn("Spirit_HIP_Gradient_and_Energy", [_vp, _PD, _PD, _i, _i], _i)
_Get_Last_Step_Time = _fn("Spirit_HIP_Get_Last_Step_Time", [_vp, _i, _i], _d)
_Device_Name = _fn("Spirit_HIP_Device_Name", [], ctypes.c_char_p)
_Profile_Enable = _fn("Spirit_HIP_Profile_Enable", [_vp, ctypes.c_bool, _i, _i])
_Profile_Get = _fn("Spirit_HIP_Profile_Get", [_vp, ctypes.c_char_p, _PD, ctypes.POINTER(_i), _i, _i])
_Profile_Reset = _fn("Spirit_HIP_Profile_Reset", [_vp, _i, _i])
_DDI_Info = _fn("Spirit_HIP_DDI_Info", [_vp, ctypes.c_char_p, _i, _i, _i], _i)
_LLG_Start_Batch = _fn("Spirit_HIP_Simulation_LLG_Start_Batch",
                       [ctypes.POINTER(_vp), _i, _i, _i, _i, ctypes.c_void_p])


def llg_start_batch(states, solver_type, n_iterations=-1, n_iterations_log=-1):
    """Simulation_LLG_Start on the active image of every state in `states`, all replicas advanced
    together on the GPU.  Returns one simulation_run_info per state."""
    from spirit_amd import simulation
    n = len(states)
    arr = (_vp * n)(*[ctypes.c_void_p(p).value for p in states])
    infos = (simulation.simulation_run_info * n)()
    spiritlib.wrap_function(_LLG_Start_Batch, [arr, _i(n), _i(solver_type), _i(n_iterations), _i(n_iterations_log),
                                               ctypes.cast(infos, ctypes.c_void_p)])
    return list(infos)


def profile_enable(p_state, enable=True, idx_image=-1, idx_chain=-1):
    _Profile_Enable(ctypes.c_void_p(p_state), ctypes.c_bool(enable), _i(idx_image), _i(idx_chain))


def profile_get(p_state, kernel, idx_image=-1, idx_chain=-1):
    """(total milliseconds, launch count) of a kernel class since the last reset."""
    ms, n = _d(0), _i(0)
    _Profile_Get(ctypes.c_void_p(p_state), ctypes.c_char_p(kernel.encode()), ctypes.byref(ms), ctypes.byref(n),
                 _i(idx_image), _i(idx_chain))
    return ms.value, n.value


def profile_reset(p_state, idx_image=-1, idx_chain=-1):
    _Profile_Reset(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def set_host_sync(p_state, mode):
    _Set_Host_Sync(ctypes.c_void_p(p_state), _i(1 if mode else 0))


def set_thermal_noise(p_state, noise, idx_image=-1, idx_chain=-1):
    """Inject N(0,1) noise of shape (n_steps, nos, 3) for the next n_steps LLG steps (None: off)."""
    if noise is None:
        _Set_Thermal_Noise(ctypes.c_void_p(p_state), None, 0, _i(idx_image), _i(idx_chain))
        return
    arr = np.ascontiguousarray(noise, dtype=np.float64)
    _Set_Thermal_Noise(ctypes.c_void_p(p_state), arr.ctypes.data_as(_PD), _i(arr.shape[0]), _i(idx_image),
                       _i(idx_chain))


def set_site_anisotropy(p_state, K, normal=(0.0, 0.0, 1.0), idx_image=-1, idx_chain=-1):
    if K is None:
        _Set_Site_Anisotropy(ctypes.c_void_p(p_state), None, _F3(*normal), _i(idx_image), _i(idx_chain))
        return
    arr = np.ascontiguousarray(K, dtype=np.float64)
    _Set_Site_Anisotropy(ctypes.c_void_p(p_state), arr.ctypes.data_as(_PD), _F3(*normal), _i(idx_image),
                         _i(idx_chain))


def gradient_and_energy(p_state, nos, idx_image=-1, idx_chain=-1):
    g = np.zeros((nos, 3))
    e = _d(0)
    rc = _Gradient_and_Energy(ctypes.c_void_p(p_state), g.ctypes.data_as(_PD), ctypes.byref(e), _i(idx_image),
                              _i(idx_chain))
    if rc != 0:
        raise RuntimeError("Spirit_HIP_Gradient_and_Energy failed")
    return g, e.value


_Gradient_and_Energy_NoAlloc = _fn("Spirit_HIP_Gradient_and_Energy_NoAlloc", [_vp, _vp, _vp, _vp, _vp, _i, _i], _i)
NOALLOC_MAX_SPINS = 1048576  # SPIRIT_HIP_NOALLOC_MAX_SPINS


def gradient_and_energy_noalloc(p_state, spins_ptr, gradient_ptr, energy_ptr, stream=None, idx_image=-1,
                                idx_chain=-1):
    """Spirit_HIP_Gradient_and_Energy_NoAlloc on caller-owned device arrays (raw addresses: spins and
    gradient 3*nos doubles AoS, energy one double) on `stream` (a hipStream_t address or None).
    Returns the entry point's status (0 on success)."""
    return _Gradient_and_Energy_NoAlloc(ctypes.c_void_p(p_state), _vp(spins_ptr), _vp(gradient_ptr), _vp(energy_ptr),
                                        _vp(stream), _i(idx_image), _i(idx_chain))


_Get_Hessian = _fn("Spirit_HIP_Hamiltonian_Get_Hessian", [_vp, _PD, _i, _i], _i)
_Hessian_VP = _fn("Spirit_HIP_Hamiltonian_Hessian_Vector_Product", [_vp, _PD, _PD, _i, _i], _i)


def hessian(p_state, nos, idx_image=-1, idx_chain=-1):
    """Dense embedding Hessian (3 nos x 3 nos) of the image's Hamiltonian, computed on the GPU
    (Hamiltonian_Heisenberg::Hessian)."""
    H = np.zeros((3 * nos, 3 * nos))
    if _Get_Hessian(ctypes.c_void_p(p_state), H.ctypes.data_as(_PD), _i(idx_image), _i(idx_chain)) != 0:
        raise RuntimeError("Spirit_HIP_Hamiltonian_Get_Hessian failed")
    return H


def hessian_vector_product(p_state, v, idx_image=-1, idx_chain=-1):
    """H v for the embedding Hessian, without forming it; v of shape (nos, 3)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(v)
    if _Hessian_VP(ctypes.c_void_p(p_state), v.ctypes.data_as(_PD), out.ctypes.data_as(_PD), _i(idx_image),
                   _i(idx_chain)) != 0:
        raise RuntimeError("Spirit_HIP_Hamiltonian_Hessian_Vector_Product failed")
    return out


_Energy_per_Spin = _fn("Spirit_HIP_Get_Energy_Contributions_per_Spin", [_vp, _PD, _i, _i], _i)


def energy_contributions_per_spin(p_state, nos, idx_image=-1, idx_chain=-1):
    """Per-spin energy contributions {name: (nos,) array} of the image's current spins."""
    from spirit_amd import system
    names = system.get_energy_array_names(p_state, idx_image, idx_chain)
    out = np.zeros((max(len(names), 1), nos))
    _Energy_per_Spin(ctypes.c_void_p(p_state), out.ctypes.data_as(_PD), _i(idx_image), _i(idx_chain))
    return {n: out[k] for k, n in enumerate(names)}


def last_step_ms(p_state, idx_image=-1, idx_chain=-1):
    return float(_Get_Last_Step_Time(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def device_name():
    return bytes(_Device_Name()).decode()


def ddi_info(p_state, idx_image=-1, idx_chain=-1):
    """Description of the DDI plan of the image's running LLG method, or None without DDI."""
    buf = ctypes.create_string_buffer(256)
    if not _DDI_Info(ctypes.c_void_p(p_state), buf, _i(256), _i(idx_image), _i(idx_chain)):
        return None
    return buf.value.decode()
