"""MI355X-native lattice-Boltzmann engine for the PV2013 vessel-flow hot path.

The compute path is libpv2013_lbm.so (hand-written HIP for gfx950 behind the C
ABI in include/lbm_capi.h); this package is its Python binding plus seeded
geometry generators.
"""
from .capi import (BC_PRESSURE, BC_VELOCITY, PROFILE_PARABOLIC, PROFILE_UNIFORM, SCHEME_AA,
                   SCHEME_TWO_LATTICE, STORE_F16S, STORE_F32, STORE_F64, WAVE_FOURIER, WAVE_NONE,
                   WAVE_PV2013, Context, Domain, LbmError, autotune, batch, grid_from_bbox, lib, make_bc,
                   make_grid, make_plane, read_stl)

__all__ = [
    "BC_PRESSURE", "BC_VELOCITY", "PROFILE_PARABOLIC", "PROFILE_UNIFORM", "SCHEME_AA",
    "SCHEME_TWO_LATTICE", "STORE_F16S", "STORE_F32", "STORE_F64", "WAVE_FOURIER", "WAVE_NONE",
    "WAVE_PV2013", "Context", "Domain", "LbmError", "autotune", "batch", "grid_from_bbox", "lib", "make_bc", "make_grid",
    "make_plane", "read_stl",
]
