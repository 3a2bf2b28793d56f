"""rayleigh_amd -- MI355X-native engine for Rayleigh's spectral-transform stack.

The compute lives in the HIP library libraysht.so (C ABI in include/raysht.h);
this package loads it and mirrors the reference's Math_Layer interface names.
"""
from .transforms import (BatchGraph, band_interleave, band_storage, ChebyGrid, LegendreBatch, LegendrePlan, RoundTripGraph,  # noqa: F401
                         fft_to_physical, fft_to_spectral, fft_to_spectral_rsc, l_l_plus1, lu_decompose_full,
                         lu_solve_full, m_values_balanced, over_l_l_plus1,
                         lu_decompose_band, lu_solve_band)

__all__ = ["LegendrePlan", "LegendreBatch", "BatchGraph", "ChebyGrid", "RoundTripGraph", "fft_to_physical",
           "fft_to_spectral", "fft_to_spectral_rsc", "lu_decompose_full", "lu_solve_full", "l_l_plus1",
           "over_l_l_plus1", "m_values_balanced", "band_storage", "band_interleave",
           "lu_decompose_band", "lu_solve_band"]
