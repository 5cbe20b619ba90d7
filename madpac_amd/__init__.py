"""MI355X-native D3Q19 block lattice-Boltzmann solver (HIP kernels for gfx950 behind a C-ABI).

The compute path is liblbm_hip.so (madpac_amd/csrc, built by madpac_amd/build.py); the C-ABI is
declared in include/lbm_capi.h and the reference-style C++ operator API in
madpac_amd/csrc/peano_blocklb.h.  ``madpac_amd.lbm`` binds the C-ABI for Python callers.
"""
from . import lbm  # noqa: F401

__all__ = ["lbm"]
