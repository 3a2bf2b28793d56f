"""Loads libraysht.so (the HIP library) and declares its C ABI for ctypes.

There is no fallback: if the library is missing or fails to load, every entry
point of this package raises.  Build it with `python -m rayleigh_amd.build`.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libraysht.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "raysht.h")

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_long
_D = ctypes.c_double
_S = ctypes.c_size_t

# name: (restype, argtypes)
SIGNATURES = {
    "raysht_strerror": (ctypes.c_char_p, [_I]),
    "raysht_version": (_I, []),
    "raysht_plan_create": (_I, [ctypes.POINTER(_P), _I, _I, _P, _I, _I, _I, _I]),
    "raysht_plan_create_cols": (_I, [ctypes.POINTER(_P), _I, _I, _P, _I, _I, _I]),
    "raysht_plan_destroy": (None, [_P]),
    "raysht_legendre_stp_host": (_I, [_P, _P, _P]),
    "raysht_legendre_pts_host": (_I, [_P, _P, _P]),
    "raysht_fft_to_physical_host": (_I, [_I, _L, _P]),
    "raysht_fft_to_spectral_host": (_I, [_I, _L, _P, _I]),
    "raysht_cheb_to_spectral_host": (_I, [_P, _P, _P, _L]),
    "raysht_cheb_from_spectral_host": (_I, [_P, _P, _P, _L]),
    "raysht_cheb_deriv_host": (_I, [_P, _P, _P, _L, _I]),
    "raysht_spectral_size": (_S, [_P]),
    "raysht_hybrid_size": (_S, [_P]),
    "raysht_legendre_stp": (_I, [_P, _P, _P, _P]),
    "raysht_legendre_pts": (_I, [_P, _P, _P, _P]),
    "raysht_plan_nodes": (_I, [_P, _P, _P, _P, _P]),
    "raysht_plan_nodes_quad": (_I, [_P, _P, _P]),
    "raysht_plan_table_size": (_S, [_P, _I]),
    "raysht_plan_table": (_I, [_P, _I, _P]),
    "raysht_fft_to_physical": (_I, [_I, _L, _P, _P]),
    "raysht_fft_to_spectral": (_I, [_I, _L, _P, _I, _P]),
    "raysht_permute_2a3a": (_I, [_P, _I, _P, _P, _P]),
    "raysht_permute_3b2b": (_I, [_P, _I, _P, _P, _P]),
    "raysht_reform_to_physical": (_I, [_P, _I, _P, _P, _P]),
    "raysht_reform_to_spectral": (_I, [_P, _I, _P, _P, _P]),
    "raysht_spectral_to_physical": (_I, [_P, _I, _P, _P, _P, _P]),
    "raysht_physical_to_spectral": (_I, [_P, _I, _P, _P, _P, _P]),
    "raysht_spectral_to_physical_scratch": (_I, [_P, _I, _P, _P, _P]),
    "raysht_graph_create_round_trip": (_I, [ctypes.POINTER(_P), _P, _I, _P, _P, _P, _P]),
    "raysht_graph_create_batch": (_I, [ctypes.POINTER(_P), _P, _P, _P, _P]),
    "raysht_graph_launch": (_I, [_P, _P]),
    "raysht_batch_create": (_I, [ctypes.POINTER(_P), _P, _I]),
    "raysht_batch_destroy": (None, [_P]),
    "raysht_batch_size": (_I, [_P]),
    "raysht_batch_stp": (_I, [_P, _P, _P, _P]),
    "raysht_batch_pts": (_I, [_P, _P, _P, _P]),
    "raysht_batch_autotune": (_I, [_P, _P, _P, _P, _I, _P, _P]),
    "raysht_batch_schedule": (_I, [_P, _I]),
    "raysht_batch_set_schedule": (_I, [_P, _I, _I]),
    "raysht_graph_destroy": (None, [_P]),
    "raysht_cheb_create": (_I, [ctypes.POINTER(_P), _I, _D, _D, _P]),
    "raysht_cheb_create_domains": (_I, [ctypes.POINTER(_P), _I, _P, _P, _P, _P]),
    "raysht_cheb_destroy": (None, [_P]),
    "raysht_cheb_table": (_I, [_P, _I, _P]),
    "raysht_cheb_grid_data": (_I, [_P, _P, _P, _P, _P]),
    "raysht_cheb_to_spectral": (_I, [_P, _P, _P, _L, _P]),
    "raysht_cheb_from_spectral": (_I, [_P, _P, _P, _L, _P]),
    "raysht_cheb_dealias": (_I, [_P, _P, _L, _P]),
    "raysht_cheb_set_form": (_I, [_P, _I]),
    "raysht_cheb_deriv": (_I, [_P, _P, _P, _L, _I, _P]),
    "raysht_permute_2b1b": (_I, [_P, _P, _P, _P]),
    "raysht_permute_1a2a": (_I, [_P, _P, _P, _P]),
    "raysht_permute_2a3a_host": (_I, [_P, _I, _P, _P]),
    "raysht_permute_3b2b_host": (_I, [_P, _I, _P, _P]),
    "raysht_permute_2b1b_host": (_I, [_P, _P, _P]),
    "raysht_permute_1a2a_host": (_I, [_P, _P, _P]),
    "raysht_lm_count": (_I, [_P]),
    "raysht_lm_values": (_I, [_P, _P, _P]),
    "raysht_d_by_dtheta": (_I, [_P, _I, _P, _I, _I, _P, _I, _I, _P]),
    "raysht_d_by_dphi": (_I, [_P, _P, _I, _I, _P, _I, _I, _P]),
    "raysht_angular_deriv_host": (_I, [_P, _I, _P, _I, _I, _P, _I, _I]),
    "raysht_rlm_spaceb_hydro": (_I, [_P, _P, _I, _I, _I, _I, _P]),
    "raysht_rlm_spacea_hydro": (_I, [_P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "raysht_compute_bandcurlb": (_I, [_P, _P, _I, _P, _P, _P]),
    "raysht_adjust_emf": (_I, [_P, _P, _I, _I, _I, _P]),
    "raysht_hybrid_output_initial": (_I, [_P, _P, _I, _P, _I, _P, _P, _P, _P]),
    "raysht_hybrid_output_final": (_I, [_P, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P]),
    "raysht_lm_scale": (_I, [_P, _P, _I, _I, _P, _I, _I, _P, _P, _P]),
    "raysht_zero_lmax": (_I, [_P, _P, _I, _P]),
    "raysht_lu_decompose_full": (_I, [_I, _I, _P, _P, _P, _P]),
    "raysht_lu_solve_full": (_I, [_I, _I, _P, _P, _P, _P, _P, _I, _P]),
    "raysht_lu_decompose_band": (_I, [_I, _I, _I, _I, _P, _I, _P, _P, _P]),
    "raysht_lu_solve_band": (_I, [_I, _I, _I, _I, _P, _I, _P, _P, _P, _P, _I, _P]),
    "raysht_lu_decompose_host": (_I, [_I, _I, _I, _P, _I, _P, _P]),
    "raysht_lu_solve_host": (_I, [_I, _I, _I, _P, _I, _P, _P, _I, _I]),
}

_lib = None


class RayshtError(RuntimeError):
    def __init__(self, status: int, what: str):
        msg = lib().raysht_strerror(status)
        super().__init__(f"{what}: {msg.decode() if msg else status} (status {status})")
        self.status = status


def header_symbols(path: str = HEADER) -> list[str]:
    """Function names the public header declares."""
    with open(path) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(raysht_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m rayleigh_amd.build`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        raise RayshtError(status, what)
