"""CPU unit test of the library's table algorithm: the double-double,
extended-exponent P_lm recursion the device runs (plm_recursion.h), compiled
for the host and compared with the oracle's quad restatement."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rayleigh_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("plm") / "libplmhost.so")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-ffp-contract=off", "-fPIC", "-shared", "-Wno-unknown-pragmas",
                    "-I", CSRC, os.path.join(ROOT, "tests", "csrc", "plm_host_driver.cpp"),
                    os.path.join(CSRC, "host_quad.cpp"), "-o", out, "-lquadmath"], check=True)
    lib = ctypes.CDLL(out)
    lib.plm_host_tables.argtypes = [ctypes.c_int] * 2 + [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 4
    return lib


def host_tables(lib, nt, lmax, mv):
    mv = np.ascontiguousarray(mv, dtype=np.int32)
    ne = sum((lmax - int(m) + 2) // 2 for m in mv) * (nt // 2)
    no = sum((lmax - int(m) + 1) // 2 for m in mv) * (nt // 2)
    pe, ie = np.zeros(ne), np.zeros(ne)
    po, io = np.zeros(max(no, 1)), np.zeros(max(no, 1))
    P = ctypes.c_void_p
    lib.plm_host_tables(nt, lmax, mv.ctypes.data_as(P), len(mv), pe.ctypes.data_as(P), po.ctypes.data_as(P),
                        ie.ctypes.data_as(P), io.ctypes.data_as(P))
    return pe, po[:no], ie, io[:no]


def to_ref_orientation(tab, nt, lmax, mv, even):
    """[k][j] device layout -> p_lm(1:n, 1:nt2) (l fastest)."""
    out = np.empty_like(tab)
    nt2, off = nt // 2, 0
    for m in mv:
        n = (lmax - int(m) + 2) // 2 if even else (lmax - int(m) + 1) // 2
        out[off: off + n * nt2] = tab[off: off + n * nt2].reshape(n, nt2).T.ravel()
        off += n * nt2
    return out


@pytest.mark.parametrize("nt,lmax", [(20, 12), (96, 63), (384, 255)])
def test_host_recursion_matches_quad_oracle(driver, nt, lmax):
    mv = O.m_order_balanced(lmax)
    pe, po, ie, io = host_tables(driver, nt, lmax, mv)
    ref = O.legendre_tables(nt, lmax, mv)
    got = (to_ref_orientation(pe, nt, lmax, mv, True), to_ref_orientation(po, nt, lmax, mv, False), ie, io)
    for w, (a, b) in enumerate(zip(got, ref)):
        d = a != b
        # rounding of two ~1e-31-accurate values to double agrees except at near-ties
        assert np.count_nonzero(d) <= max(1, b.size // 100000), (w, np.count_nonzero(d))
        if d.any():
            assert np.all(np.abs(a[d] - b[d]) <= np.spacing(np.abs(b[d]))), w
