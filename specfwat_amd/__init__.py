"""MI355X-native engine for SpecFWAT's spectral-element forward/adjoint/kernel
path and its PDE kernel smoother.

The device code is libspecfwat_hip.so (hand-written HIP for gfx950, C ABI in
include/specfwat_hip.h); this package is the host side: mesh/model/source
generation for the synthetic configurations and the run sequence of
SpecFWAT's PrepareFWD (src/preproc/prepare_fwd.f90:297-384) over the C ABI.
"""
__version__ = "0.1.0"
