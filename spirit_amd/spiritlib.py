"""Loads libSpirit.so from the package directory (core/python/spirit/spiritlib.py:2-43)."""
import ctypes
import os
import threading

_lib = None


def _get_spirit_lib_name():
    return "libSpirit.so"


def _share_torch_runtime():
    """Load torch's HIP runtime before the library, when torch is installed.

    torch ships its own libamdhip64 / libhsa-runtime64, which it opens by a different file
    name than the library's libamdhip64.so.7.  Loaded first, torch's copy is the one the
    library binds to (the dynamic linker matches the soname), so the process has a single
    HIP runtime and torch tensors and the engine share devices, memory and streams.  Loaded
    the other way round, a second HSA runtime in the process cannot open the GPU and torch
    reports no devices.  SPIRIT_AMD_NO_TORCH=1 skips this (the library then uses the system
    ROCm runtime, and torch cannot use the GPU in the same process)."""
    if os.environ.get("SPIRIT_AMD_NO_TORCH", "0") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_spirit_library():
    """Return the engine library; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        path = os.path.join(os.path.dirname(os.path.realpath(__file__)), _get_spirit_lib_name())
        if not os.path.exists(path):
            raise OSError("libSpirit.so not found at %s -- build it with `python -m spirit_amd.build`" % path)
        _share_torch_runtime()
        _lib = ctypes.CDLL(path)
    return _lib


def wrap_function(function, arguments):
    """Run a blocking library call on a daemon thread so that Ctrl-C stays responsive."""
    t = threading.Thread(target=function, args=arguments)
    t.daemon = True
    t.start()
    while t.is_alive():
        t.join(0.1)
