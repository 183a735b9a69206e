"""Log (core/python/spirit/log.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()
_vp, _i, _b = ctypes.c_void_p, ctypes.c_int, ctypes.c_bool

LEVEL_ALL, LEVEL_SEVERE, LEVEL_ERROR, LEVEL_WARNING, LEVEL_PARAMETER, LEVEL_INFO, LEVEL_DEBUG = range(7)
SENDER_ALL, SENDER_IO, SENDER_GNEB, SENDER_LLG, SENDER_MC, SENDER_MMF, SENDER_EMA, SENDER_API, SENDER_UI = range(9)

_Send = _spirit.Log_Send
_Send.argtypes = [_vp, _i, _i, ctypes.c_char_p, _i, _i]
_Send.restype = None
_N_Errors = _spirit.Log_Get_N_Errors
_N_Errors.argtypes = [_vp]
_N_Errors.restype = _i
_N_Warnings = _spirit.Log_Get_N_Warnings
_N_Warnings.argtypes = [_vp]
_N_Warnings.restype = _i
_Set_Console = _spirit.Log_Set_Output_To_Console
_Set_Console.argtypes = [_vp, _b, _i]
_Set_Console.restype = None


def send(p_state, level, sender, message, idx_image=-1, idx_chain=-1):
    _Send(ctypes.c_void_p(p_state), _i(level), _i(sender), ctypes.c_char_p(message.encode()), _i(idx_image),
          _i(idx_chain))


def get_n_errors(p_state):
    return int(_N_Errors(ctypes.c_void_p(p_state)))


def get_n_warnings(p_state):
    return int(_N_Warnings(ctypes.c_void_p(p_state)))


def set_output_to_console(p_state, output, level):
    _Set_Console(ctypes.c_void_p(p_state), _b(output), _i(level))
