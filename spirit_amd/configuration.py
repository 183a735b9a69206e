"""Configurations: spin configurations inside a spatial filter (core/python/spirit/configuration.py)."""
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


_FILT = [_F3, _F3, _f, _f, _b]
_Domain = _fn("Configuration_Domain", [_vp, _F3] + _FILT + [_i, _i])
_PlusZ = _fn("Configuration_PlusZ", [_vp] + _FILT + [_i, _i])
_MinusZ = _fn("Configuration_MinusZ", [_vp] + _FILT + [_i, _i])
_Random = _fn("Configuration_Random", [_vp] + _FILT + [_b, _i, _i])
_Add_Noise = _fn("Configuration_Add_Noise_Temperature", [_vp, _f] + _FILT + [_i, _i])
_Skyrmion = _fn("Configuration_Skyrmion", [_vp, _f, _f, _f, _b, _b, _b] + _FILT + [_i, _i])
_SpinSpiral = _fn("Configuration_SpinSpiral", [_vp, ctypes.c_char_p, _F3, _F3, _f] + _FILT + [_i, _i])
_DW_Skyrmion = _fn("Configuration_DW_Skyrmion", [_vp, _f, _f, _f, _f, _b, _b, _b] + _FILT + [_i, _i])
_Hopfion = _fn("Configuration_Hopfion", [_vp, _f, _i] + _FILT + [_F3, _i, _i])
_To_Clipboard = _fn("Configuration_To_Clipboard", [_vp, _i, _i])
_From_Clipboard = _fn("Configuration_From_Clipboard", [_vp] + _FILT + [_i, _i])


def _filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted):
    return [_F3(*pos), _F3(*border_rectangular), _f(border_cylindrical), _f(border_spherical), _b(inverted)]


def domain(p_state, dir, pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0,
           border_spherical=-1.0, inverted=False, idx_image=-1, idx_chain=-1):
    _Domain(ctypes.c_void_p(p_state), _F3(*dir),
            *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted), _i(idx_image),
            _i(idx_chain))


def plus_z(p_state, pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0,
           border_spherical=-1.0, inverted=False, idx_image=-1, idx_chain=-1):
    _PlusZ(ctypes.c_void_p(p_state), *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted),
           _i(idx_image), _i(idx_chain))


def minus_z(p_state, pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0,
            border_spherical=-1.0, inverted=False, idx_image=-1, idx_chain=-1):
    _MinusZ(ctypes.c_void_p(p_state), *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted),
            _i(idx_image), _i(idx_chain))


def random(p_state, pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0,
           border_spherical=-1.0, inverted=False, external=False, idx_image=-1, idx_chain=-1):
    _Random(ctypes.c_void_p(p_state), *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted),
            _b(external), _i(idx_image), _i(idx_chain))


def add_noise(p_state, temperature, pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0],
              border_cylindrical=-1.0, border_spherical=-1.0, inverted=False, idx_image=-1, idx_chain=-1):
    _Add_Noise(ctypes.c_void_p(p_state), _f(temperature),
               *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted), _i(idx_image),
               _i(idx_chain))


def skyrmion(p_state, radius, order=1, phase=1, up_down=False, achiral=False, right_left=False,
             pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0,
             border_spherical=-1.0, inverted=False, idx_image=-1, idx_chain=-1):
    _Skyrmion(ctypes.c_void_p(p_state), _f(radius), _f(order), _f(phase), _b(up_down), _b(achiral), _b(right_left),
              *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted), _i(idx_image),
              _i(idx_chain))


def dw_skyrmion(p_state, dw_radius, dw_width, order=1, phase=1, up_down=False, achiral=False, right_left=False,
                pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0,
                border_spherical=-1.0, inverted=False, idx_image=-1, idx_chain=-1):
    """A skyrmion bounded by a circular domain wall of radius dw_radius and width dw_width."""
    _DW_Skyrmion(ctypes.c_void_p(p_state), _f(dw_radius), _f(dw_width), _f(order), _f(phase), _b(up_down),
                 _b(achiral), _b(right_left), *_filt(pos, border_rectangular, border_cylindrical, border_spherical,
                                                     inverted), _i(idx_image), _i(idx_chain))


def hopfion(p_state, radius, order=1, pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0],
            border_cylindrical=-1.0, border_spherical=-1.0, inverted=False, normal=[0.0, 0.0, 1.0], idx_image=-1,
            idx_chain=-1):
    """A toroidal hopfion of the given radius about the axis `normal`."""
    _Hopfion(ctypes.c_void_p(p_state), _f(radius), _i(order),
             *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted), _F3(*normal),
             _i(idx_image), _i(idx_chain))


def spin_spiral(p_state, direction_type, q_vector, axis, theta, pos=[0.0, 0.0, 0.0],
                border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0, border_spherical=-1.0,
                inverted=False, idx_image=-1, idx_chain=-1):
    _SpinSpiral(ctypes.c_void_p(p_state), ctypes.c_char_p(direction_type.encode()), _F3(*q_vector), _F3(*axis),
                _f(theta), *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted),
                _i(idx_image), _i(idx_chain))


def to_clipboard(p_state, idx_image=-1, idx_chain=-1):
    _To_Clipboard(ctypes.c_void_p(p_state), _i(idx_image), _i(idx_chain))


def from_clipboard(p_state, pos=[0.0, 0.0, 0.0], border_rectangular=[-1.0, -1.0, -1.0], border_cylindrical=-1.0,
                   border_spherical=-1.0, inverted=False, idx_image=-1, idx_chain=-1):
    _From_Clipboard(ctypes.c_void_p(p_state),
                    *_filt(pos, border_rectangular, border_cylindrical, border_spherical, inverted), _i(idx_image),
                    _i(idx_chain))
