"""mfa_amd: MI355X implementation of MFA_PROG's force-and-integrate loop.

The compute path is libmfa_hip.so (HIP kernels for gfx950 behind the C ABI
of include/mfa_hip.h); this package holds its sources, the build recipe, the
Fortran driver and a ctypes binding.
"""
