"""MI355X-native pseudospectral Fourier-Chebyshev RK3 substep (TURBID drop-in).

The compute path is the HIP library ``libturbid_amd.so`` (hand-written gfx950
kernels behind a C ABI, declared in include/turbid_amd.h); ``turbid_amd.runtime``
binds it.  ``turbid_amd.workloads`` holds workload definitions and seeded
initial states.
"""
