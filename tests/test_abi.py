"""The C ABI library loads on any machine and exports every entry point the
public header declares; argument validation happens before any device work.
CPU only (no compute calls)."""
import ctypes

import pytest

from rayleigh_amd import _lib


def test_library_exports_every_header_symbol():
    lib = _lib.lib()
    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert set(names) == set(_lib.SIGNATURES), "ctypes declarations out of step with include/raysht.h"


def test_strerror_and_version():
    lib = _lib.lib()
    assert lib.raysht_strerror(0) == b"ok"
    assert lib.raysht_strerror(1) == b"invalid argument"
    assert lib.raysht_version() >= 100


@pytest.mark.parametrize("args", [
    (63, 95, [0, 1], 4, 1, 0),     # odd n_theta
    (63, 96, [0, 64], 4, 1, 0),    # m > l_max
    (63, 96, [3, 3], 4, 1, 0),     # repeated m
    (63, 96, [0], 0, 1, 0),        # no radii
    (63, 96, [0], 4, 1, 7),        # unknown precision
    (-1, 96, [0], 4, 1, 0),        # negative l_max
])
def test_plan_create_rejects_bad_arguments(args):
    lmax, nt, mv, nr, nf, prec = args
    arr = (ctypes.c_int * len(mv))(*mv)
    h = ctypes.c_void_p()
    st = _lib.lib().raysht_plan_create(ctypes.byref(h), lmax, nt, ctypes.cast(arr, ctypes.c_void_p), len(mv),
                                       nr, nf, prec)
    assert st == 1 and not h.value


def test_fft_and_cheb_reject_bad_arguments():
    lib = _lib.lib()
    dummy = ctypes.c_void_p(16)
    assert lib.raysht_fft_to_physical(7, 1, dummy, None) == 1       # odd length
    assert lib.raysht_fft_to_physical(2 * 11, 1, dummy, None) == 4  # a prime factor above 7
    assert lib.raysht_fft_to_spectral(8, -1, dummy, 0, None) == 1
    h = ctypes.c_void_p()
    assert lib.raysht_cheb_create(ctypes.byref(h), 1, 0.0, 1.0, None) == 1
    assert lib.raysht_cheb_create(ctypes.byref(h), 48, 1.0, 0.5, None) == 1
    assert lib.raysht_cheb_set_form(None, 0) == 1
    assert lib.raysht_legendre_stp(None, dummy, dummy, None) == 1
