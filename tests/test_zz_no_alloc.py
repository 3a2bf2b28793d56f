"""raysht_spectral_to_physical_scratch: the allocation-free form of
raysht_spectral_to_physical (hybrid intermediate in the library's static device array),
called through the C ABI and compared bit for bit with the form that takes a caller's
`work` buffer, on two calls in a row (the second on a smaller problem), plus its two
refusals: a plan whose hybrid array exceeds the array, and an FFT length whose twiddle
table has not been made on the device."""
import ctypes

import numpy as np
import pytest

from rayleigh_amd import _lib

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = (32 << 20) // 8


@pytest.fixture(scope="module")
def ra():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rayleigh_amd
    return rayleigh_amd


def _spectral(plan, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(plan.spectral_size, dtype=torch.float64, device="cuda", generator=g)


def _p3(plan, nphi):
    import torch
    return torch.full(((nphi + 2) * plan.n_r * plan.n_theta * plan.n_fields,), float("nan"),
                      dtype=torch.float64, device="cuda")


def _scratch_call(plan, nphi, s, p3):
    import torch
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return _lib.lib().raysht_spectral_to_physical_scratch(plan._h, int(nphi), ctypes.c_void_p(s.data_ptr()),
                                                          ctypes.c_void_p(p3.data_ptr()), stream)


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_scratch_form_equals_work_form(ra, precision):
    import torch
    # (l_max, N_theta, N_phi, N_r, N_f): the l_max = 63 benchmark size, then a smaller one
    cases = [(63, 96, 192, 48, 4), (31, 48, 96, 16, 2)]
    for i, (L, nt, nphi, nr, nf) in enumerate(cases):
        print(f"\n[no-alloc] {precision} call {i}: l_max={L} N_theta={nt} N_phi={nphi} N_r={nr} N_f={nf}",
              flush=True)
        plan = ra.LegendrePlan(nt, L, ra.m_values_balanced(L), nr, nf, precision=precision)
        assert plan.hybrid_size <= SCRATCH_DOUBLES
        s = _spectral(plan, seed=70 + i)
        work = torch.empty(plan.hybrid_size, dtype=torch.float64, device="cuda")
        want = _p3(plan, nphi)
        plan.spectral_to_physical(s, work, want, nphi)  # also makes the twiddle table of nphi
        got = _p3(plan, nphi)
        assert _scratch_call(plan, nphi, s, got) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"call {i}"
        del plan


def test_scratch_form_refusals(ra):
    import torch
    # hybrid array 192 x 512 x 128 doubles > the 4 Mi-double array: refused before any launch
    plan = ra.LegendrePlan(192, 127, ra.m_values_balanced(127), 64, 4)
    assert plan.hybrid_size > SCRATCH_DOUBLES
    s = _spectral(plan, seed=1)
    p3 = _p3(plan, 384)
    assert _scratch_call(plan, 384, s, p3) == 1
    torch.cuda.synchronize()
    assert bool(torch.isnan(p3).all())
    # N_phi = 630 (315 = 9 * 5 * 7), a length no other test transforms: its twiddle table
    # does not exist yet, and this form may not make it
    small = ra.LegendrePlan(48, 31, ra.m_values_balanced(31), 8, 1)
    s2 = _spectral(small, seed=2)
    p3s = _p3(small, 630)
    assert _scratch_call(small, 630, s2, p3s) == 5
    assert bool(torch.isnan(p3s).all())
    # once any other call has made the table, the same call runs and matches
    work = torch.empty(small.hybrid_size, dtype=torch.float64, device="cuda")
    want = _p3(small, 630)
    small.spectral_to_physical(s2, work, want, 630)
    assert _scratch_call(small, 630, s2, p3s) == 0
    torch.cuda.synchronize()
    assert torch.equal(p3s, want)
    assert np.isfinite(p3s.cpu().numpy()).all()
