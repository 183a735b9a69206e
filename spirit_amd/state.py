"""State: creation and deletion of the opaque simulation state (core/python/spirit/state.py)."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()


class State:
    """``with State(cfg) as p_state:`` creates and deletes a state."""

    def __init__(self, configfile="", quiet=False):
        self.p_state = setup(configfile, quiet)

    def __enter__(self):
        return self.p_state

    def __exit__(self, exc_type, exc_value, traceback):
        delete(self.p_state)


_State_Setup = _spirit.State_Setup
_State_Setup.argtypes = [ctypes.c_char_p, ctypes.c_bool]
_State_Setup.restype = ctypes.c_void_p


def setup(configfile="", quiet=False):
    return _State_Setup(ctypes.c_char_p(configfile.encode("utf-8")), ctypes.c_bool(quiet))


_State_Delete = _spirit.State_Delete
_State_Delete.argtypes = [ctypes.c_void_p]
_State_Delete.restype = None


def delete(p_state):
    return _State_Delete(ctypes.c_void_p(p_state))


_State_Update = _spirit.State_Update
_State_Update.argtypes = [ctypes.c_void_p]
_State_Update.restype = None


def update(p_state):
    return _State_Update(ctypes.c_void_p(p_state))


_State_To_Config = _spirit.State_To_Config
_State_To_Config.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
_State_To_Config.restype = None


def to_config(p_state, filename, comment=""):
    return _State_To_Config(ctypes.c_void_p(p_state), ctypes.c_char_p(filename.encode("utf-8")),
                            ctypes.c_char_p(comment.encode("utf-8")))


_State_DateTime = _spirit.State_DateTime
_State_DateTime.argtypes = [ctypes.c_void_p]
_State_DateTime.restype = ctypes.c_char_p


def date_time(p_state):
    return bytes(_State_DateTime(ctypes.c_void_p(p_state))).decode("utf-8")
