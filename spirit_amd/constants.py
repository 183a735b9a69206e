"""Physical constants (core/python/spirit/constants.py)."""
from spirit_amd import spiritlib
import ctypes

_spirit = spiritlib.load_spirit_library()


def _c(name):
    f = getattr(_spirit, name)
    f.argtypes = []
    f.restype = ctypes.c_double
    return f()


mu_B = _c("Constants_mu_B")
mu_0 = _c("Constants_mu_0")
k_B = _c("Constants_k_B")
hbar = _c("Constants_hbar")
mRy = _c("Constants_mRy")
gamma = _c("Constants_gamma")
g_e = _c("Constants_g_e")
pi = _c("Constants_Pi")
