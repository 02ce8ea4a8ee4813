"""MI355X (gfx950) implementation of the aeros spherical-harmonic transform path.

The native library (libaer_sht.so, C ABI in include/aer_sht.h) holds the HIP
kernels; aeros_amd.fortran holds the Fortran drop-in for the reference's
aeros_spectral module; aeros_amd.spectral is the Python mirror of the same API
used by the tests and the benchmark.
"""
