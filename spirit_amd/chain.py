"""Chain of images (core/python/spirit/chain.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool, ctypes.c_float


def _fn(name, argtypes, restype=None):
    f = getattr(_spirit, name)
    f.argtypes = argtypes
    f.restype = restype
    return f


_Get_NOI = _fn("Chain_Get_NOI", [_vp, _i], _i)
_next = _fn("Chain_next_Image", [_vp, _i], _b)
_prev = _fn("Chain_prev_Image", [_vp, _i], _b)
_Jump = _fn("Chain_Jump_To_Image", [_vp, _i, _i], _b)
_Set_Length = _fn("Chain_Set_Length", [_vp, _i, _i])
_Image_to_Clipboard = _fn("Chain_Image_to_Clipboard", [_vp, _i, _i])
_Replace = _fn("Chain_Replace_Image", [_vp, _i, _i])
_Insert_Before = _fn("Chain_Insert_Image_Before", [_vp, _i, _i])
_Insert_After = _fn("Chain_Insert_Image_After", [_vp, _i, _i])
_Push_Back = _fn("Chain_Push_Back", [_vp, _i])
_Delete = _fn("Chain_Delete_Image", [_vp, _i, _i], _b)
_Pop_Back = _fn("Chain_Pop_Back", [_vp, _i], _b)
_Get_Rx = _fn("Chain_Get_Rx", [_vp, ctypes.POINTER(_f), _i])
_Get_Energy = _fn("Chain_Get_Energy", [_vp, ctypes.POINTER(_f), _i])
_Update_Data = _fn("Chain_Update_Data", [_vp, _i])
_Setup_Data = _fn("Chain_Setup_Data", [_vp, _i])


def get_noi(p_state, idx_chain=-1):
    return int(_Get_NOI(ctypes.c_void_p(p_state), _i(idx_chain)))


def next_image(p_state, idx_chain=-1):
    return bool(_next(ctypes.c_void_p(p_state), _i(idx_chain)))


def prev_image(p_state, idx_chain=-1):
    return bool(_prev(ctypes.c_void_p(p_state), _i(idx_chain)))


def jump_to_image(p_state, idx_image=-1, idx_chain=-1):
    return bool(_Jump(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def set_length(p_state, n_images, idx_chain=-1):
    _Set_Length(ctypes.c_void_p(p_state), _i(n_images), _i(idx_chain))


def image_to_clipboard(p_state, idx_image=-1, idx_chain=-1):
    _Image_to_Clipboard(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def replace_image(p_state, idx_image=-1, idx_chain=-1):
    _Replace(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def insert_image_before(p_state, idx_image=-1, idx_chain=-1):
    _Insert_Before(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def insert_image_after(p_state, idx_image=-1, idx_chain=-1):
    _Insert_After(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def push_back(p_state, idx_chain=-1):
    _Push_Back(ctypes.c_void_p(p_state), _i(idx_chain))


def delete_image(p_state, idx_image=-1, idx_chain=-1):
    return bool(_Delete(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain)))


def pop_back(p_state, idx_chain=-1):
    return bool(_Pop_Back(ctypes.c_void_p(p_state), _i(idx_chain)))


def get_rx(p_state, idx_chain=-1):
    noi = get_noi(p_state, idx_chain)
    arr = (_f * noi)()
    _Get_Rx(ctypes.c_void_p(p_state), arr, _i(idx_chain))
    return [float(x) for x in arr]


def get_energy(p_state, idx_chain=-1):
    noi = get_noi(p_state, idx_chain)
    arr = (_f * noi)()
    _Get_Energy(ctypes.c_void_p(p_state), arr, _i(idx_chain))
    return [float(x) for x in arr]


def update_data(p_state, idx_chain=-1):
    _Update_Data(ctypes.c_void_p(p_state), _i(idx_chain))


def setup_data(p_state, idx_chain=-1):
    _Setup_Data(ctypes.c_void_p(p_state), _i(idx_chain))


_Get_Rx_Interpolated = _fn("Chain_Get_Rx_Interpolated", [_vp, ctypes.POINTER(_f), _i])
_Get_Energy_Interpolated = _fn("Chain_Get_Energy_Interpolated", [_vp, ctypes.POINTER(_f), _i])


def _n_interpolated(p_state, idx_chain):
    from spirit_amd.parameters import gneb
    noi = get_noi(p_state, idx_chain)
    return noi + (noi - 1) * gneb.get_n_energy_interpolations(p_state, idx_chain)


def get_reaction_coordinate(p_state, idx_chain=-1):
    """Reaction coordinate Rx of every image (geodesic path length)."""
    return get_rx(p_state, idx_chain)


def get_reaction_coordinate_interpolated(p_state, idx_chain=-1):
    n = _n_interpolated(p_state, idx_chain)
    arr = (_f * n)()
    _Get_Rx_Interpolated(ctypes.c_void_p(p_state), arr, _i(idx_chain))
    return [float(x) for x in arr]


def get_energy_interpolated(p_state, idx_chain=-1):
    n = _n_interpolated(p_state, idx_chain)
    arr = (_f * n)()
    _Get_Energy_Interpolated(ctypes.c_void_p(p_state), arr, _i(idx_chain))
    return [float(x) for x in arr]
