"""Hamiltonian: parameters of the Heisenberg Hamiltonian (core/python/spirit/hamiltonian.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_float
_F3 = ctypes.c_float * 3

CHIRALITY_BLOCH = 1
CHIRALITY_NEEL = 2
CHIRALITY_BLOCH_INVERSE = -1
CHIRALITY_NEEL_INVERSE = -2

DDI_METHOD_NONE = 0
DDI_METHOD_FFT = 1
DDI_METHOD_FMM = 2
DDI_METHOD_CUTOFF = 3


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Set_BC = _fn("Hamiltonian_Set_Boundary_Conditions", [_vp, ctypes.POINTER(_b), _i, _i])
_Set_Field = _fn("Hamiltonian_Set_Field", [_vp, _f, _F3, _i, _i])
_Set_Anisotropy = _fn("Hamiltonian_Set_Anisotropy", [_vp, _f, _F3, _i, _i])
_Set_Cubic = _fn("Hamiltonian_Set_Cubic_Anisotropy", [_vp, _f, _i, _i])
_Set_Exchange = _fn("Hamiltonian_Set_Exchange", [_vp, _i, ctypes.POINTER(_f), _i, _i])
_Set_DMI = _fn("Hamiltonian_Set_DMI", [_vp, _i, ctypes.POINTER(_f), _i, _i, _i])
_Set_DDI = _fn("Hamiltonian_Set_DDI", [_vp, _i, ctypes.POINTER(_i), _f, _b, _i, _i])
_Get_Name = _fn("Hamiltonian_Get_Name", [_vp, _i, _i], ctypes.c_char_p)
_Get_BC = _fn("Hamiltonian_Get_Boundary_Conditions", [_vp, ctypes.POINTER(_b), _i, _i])
_Get_Field = _fn("Hamiltonian_Get_Field", [_vp, ctypes.POINTER(_f), _F3, _i, _i])
_Get_Anisotropy = _fn("Hamiltonian_Get_Anisotropy", [_vp, ctypes.POINTER(_f), _F3, _i, _i])
_Get_Cubic = _fn("Hamiltonian_Get_Cubic_Anisotropy", [_vp, ctypes.POINTER(_f), _i, _i])
_Get_DDI = _fn("Hamiltonian_Get_DDI", [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_f),
                                       ctypes.POINTER(_b), _i, _i])
_Get_Exchange_N_Pairs = _fn("Hamiltonian_Get_Exchange_N_Pairs", [_vp, _i, _i], _i)
_Get_DMI_N_Pairs = _fn("Hamiltonian_Get_DMI_N_Pairs", [_vp, _i, _i], _i)
_Get_Exchange_Pairs = _fn("Hamiltonian_Get_Exchange_Pairs", [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                                             ctypes.POINTER(_f), _i, _i])


def set_boundary_conditions(p_state, boundaries, idx_image=-1, idx_chain=-1):
    arr = (_b * 3)(*boundaries)
    _Set_BC(ctypes.c_void_p(p_state), arr, _i(idx_image), _i(idx_chain))


def set_field(p_state, magnitude, direction, idx_image=-1, idx_chain=-1):
    _Set_Field(ctypes.c_void_p(p_state), _f(magnitude), _F3(*direction), _i(idx_image), _i(idx_chain))


def set_anisotropy(p_state, magnitude, direction, idx_image=-1, idx_chain=-1):
    _Set_Anisotropy(ctypes.c_void_p(p_state), _f(magnitude), _F3(*direction), _i(idx_image), _i(idx_chain))


def set_cubic_anisotropy(p_state, magnitude, idx_image=-1, idx_chain=-1):
    _Set_Cubic(ctypes.c_void_p(p_state), _f(magnitude), _i(idx_image), _i(idx_chain))


def set_exchange(p_state, n_shells, J_ij, idx_image=-1, idx_chain=-1):
    arr = (_f * n_shells)(*J_ij)
    _Set_Exchange(ctypes.c_void_p(p_state), _i(n_shells), arr, _i(idx_image), _i(idx_chain))


def set_dmi(p_state, n_shells, D_ij, chirality=CHIRALITY_BLOCH, idx_image=-1, idx_chain=-1):
    arr = (_f * n_shells)(*D_ij)
    _Set_DMI(ctypes.c_void_p(p_state), _i(n_shells), arr, _i(chirality), _i(idx_image), _i(idx_chain))


def set_ddi(p_state, ddi_method, n_periodic_images=[4, 4, 4], radius=0.0, pb_zero_padding=True, idx_image=-1,
            idx_chain=-1):
    arr = (_i * 3)(*n_periodic_images)
    _Set_DDI(ctypes.c_void_p(p_state), _i(ddi_method), arr, _f(radius), _b(pb_zero_padding), _i(idx_image),
             _i(idx_chain))


def get_name(p_state, idx_image=-1, idx_chain=-1):
    return bytes(_Get_Name(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))).decode()


def get_boundary_conditions(p_state, idx_image=-1, idx_chain=-1):
    arr = (_b * 3)()
    _Get_BC(ctypes.c_void_p(p_state), arr, _i(idx_image), _i(idx_chain))
    return [bool(x) for x in arr]


def get_field(p_state, idx_image=-1, idx_chain=-1):
    m, n = _f(), _F3()
    _Get_Field(ctypes.c_void_p(p_state), ctypes.byref(m), n, _i(idx_image), _i(idx_chain))
    return float(m.value), [float(x) for x in n]


def get_anisotropy(p_state, idx_image=-1, idx_chain=-1):
    m, n = _f(), _F3()
    _Get_Anisotropy(ctypes.c_void_p(p_state), ctypes.byref(m), n, _i(idx_image), _i(idx_chain))
    return float(m.value), [float(x) for x in n]


def get_cubic_anisotropy(p_state, idx_image=-1, idx_chain=-1):
    m = _f()
    _Get_Cubic(ctypes.c_void_p(p_state), ctypes.byref(m), _i(idx_image), _i(idx_chain))
    return float(m.value)


def get_ddi(p_state, idx_image=-1, idx_chain=-1):
    method, n, r, zp = _i(), (_i * 3)(), _f(), _b()
    _Get_DDI(ctypes.c_void_p(p_state), ctypes.byref(method), n, ctypes.byref(r), ctypes.byref(zp), _i(idx_image),
             _i(idx_chain))
    return {"ddi_method": method.value, "n_periodic_images": [int(x) for x in n], "radius": r.value,
            "pb_zero_padding": bool(zp.value)}


def get_exchange_pairs(p_state, idx_image=-1, idx_chain=-1):
    """(pairs (n,2), translations (n,3), J (n,)) of the redundant exchange table."""
    import numpy as np
    n = _Get_Exchange_N_Pairs(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))
    idx = np.zeros((max(n, 1), 2), dtype=np.int32)
    tr = np.zeros((max(n, 1), 3), dtype=np.int32)
    J = np.zeros(max(n, 1), dtype=np.float32)
    _Get_Exchange_Pairs(ctypes.c_void_p(p_state), idx.ctypes.data_as(ctypes.POINTER(_i)),
                        tr.ctypes.data_as(ctypes.POINTER(_i)), J.ctypes.data_as(ctypes.POINTER(_f)), _i(idx_image),
                        _i(idx_chain))
    return idx[:n], tr[:n], J[:n]


_Write_Hessian = _fn("Hamiltonian_Write_Hessian", [_vp, ctypes.c_char_p, _b, _i, _i])


def write_hessian(p_state, filename, triplet_format=True, idx_image=-1, idx_chain=-1):
    """Writes the 3N x 3N embedding Hessian (anisotropy, exchange, DMI) to a file, as
    "row col value" triplets (recommended for large, sparse Hessians) or as a dense matrix
    (core/python/spirit/hamiltonian.py:338-358)."""
    _Write_Hessian(ctypes.c_void_p(p_state), filename.encode("utf-8"), _b(triplet_format), _i(idx_image),
                   _i(idx_chain))
