"""The library's scalar type (the reference generates this module at build time,
core/CMakeLists.txt:531-533): the engine computes in double precision."""
import ctypes

scalar = ctypes.c_double
