"""Build description (the fields of the module the reference generates at build time,
core/CMakeLists.txt:535-551), plus ``hip`` for this engine."""
version_major = 2
version_minor = 2
version_patch = 0
version = "2.2.0"
revision = "hip"
version_full = "2.2.0 (hip)"
compiler = "hipcc"
compiler_version = "ROCm 7.2"
compiler_full = "hipcc (ROCm 7.2), gfx950"
scalartype = "double"
pinning = "OFF"
defects = "OFF"
cuda = "OFF"
openmp = "OFF"
threads = "OFF"
fftw = "OFF"
hip = "ON"
