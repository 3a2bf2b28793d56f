"""The Fortran ISO_C_BINDING layer (rayleigh_amd/fortran) driven by a Test_SHT-style
Fortran program: Legendre_Transform (4-D and 2-D), FFT_To_Physical/Spectral and the
Chebyshev pair, radial and angular derivatives and the dense and band implicit solves through the
host-array entry points, and Adjust_Emf_Device / Zero_Lmax_Device on a device buffer against the
reference's statements over the host derivatives, on the GPU."""
import os
import subprocess

import pytest

from rayleigh_amd import build as B

pytestmark = pytest.mark.gpu


def test_fortran_binding_test_sht():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(B.FORTRAN_TEST):
        pytest.skip("Fortran test program not built (no Fortran compiler at build time)")
    r = subprocess.run([B.FORTRAN_TEST], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if "result =" in ln]
    # fourteen FP64 checks (Legendre, the LT/FFT loop with index loops, with Reform_* and on
    # device buffers, the radial reforms, the Real*16 grid, the 2-D pair, Chebyshev, angular
    # derivatives, implicit solves, the Cheby_Grid object), then the white-noise round trips
    # with INT8X6 and BF16X6
    assert len(lines) == 18 and all("PASSING" in ln for ln in lines), r.stdout
