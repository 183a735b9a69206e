"""Loads libSpirit.so from the package directory (core/python/spirit/spiritlib.py:2-43)."""
import ctypes
import os
import threading

_lib = None


def _get_spirit_lib_name():
    return "libSpirit.so"


def load_spirit_library():
    """Return the engine library; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        path = os.path.join(os.path.dirname(os.path.realpath(__file__)), _get_spirit_lib_name())
        if not os.path.exists(path):
            raise OSError("libSpirit.so not found at %s -- build it with `python -m spirit_amd.build`" % path)
        _lib = ctypes.CDLL(path)
    return _lib


def wrap_function(function, arguments):
    """Run a blocking library call on a daemon thread so that Ctrl-C stays responsive."""
    t = threading.Thread(target=function, args=arguments)
    t.daemon = True
    t.start()
    while t.is_alive():
        t.join(0.1)
