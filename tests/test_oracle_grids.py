"""The oracle's Real*16 Gauss-Legendre nodes/weights (as the reference stores them) and
its Chebyshev radial integration weights, against the compiled reference's fixtures
(bit for bit) and what the mathematics fixes: the binary128 values round to the stored
head doubles, sum w = 2, x symmetric; the Chebyshev weights integrate r^2 dr over the
shell to 1 (Gauss-Chebyshev quadrature of 3 r^2 / (rmax^3 - rmin^3))."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.make_fixtures import GRID_CHEBS, GRID_THETAS


def quad_to_float(b):
    """binary128 bytes (little endian) -> nearest double and the exact value as a
    Python fraction-free pair (sign, exponent, mantissa) for comparisons."""
    b = np.asarray(b, dtype=np.uint8)
    lo = b[:, :8].copy().view("<u8").ravel().astype(object)
    hi = b[:, 8:].copy().view("<u8").ravel().astype(object)
    out = []
    for h, l in zip(hi, lo):
        sign = -1 if (h >> 63) else 1
        e = (h >> 48) & 0x7FFF
        mant = ((h & ((1 << 48) - 1)) << 64) | l
        out.append((sign, e, mant))
    return out


def to_double(q):
    sign, e, mant = q
    if e == 0 and mant == 0:
        return 0.0
    from fractions import Fraction
    v = Fraction((1 << 112) + mant, 1 << 112) * (Fraction(2) ** (e - 16383))
    return sign * float(v)  # float(Fraction) rounds to nearest


@pytest.mark.parametrize("nt", GRID_THETAS)
def test_binary128_nodes_bit_exact_vs_reference(golden, nt):
    g = golden("grid_nodes_weights.npz")
    x, w = O.gauss_legendre_binary128(nt)
    assert np.array_equal(x, g[f"coloc_q_{nt}"]) and np.array_equal(w, g[f"gl_weights_q_{nt}"])


@pytest.mark.parametrize("nt", [64, 96])
def test_binary128_nodes_round_to_stored_doubles(nt):
    x, w = O.gauss_legendre_binary128(nt)
    xh, _, wh, _ = O.gauss_legendre(nt)
    assert [to_double(q) for q in quad_to_float(x)] == list(xh)
    assert [to_double(q) for q in quad_to_float(w)] == list(wh)
    assert abs(sum(to_double(q) for q in quad_to_float(w)) - 2.0) < 1e-13
    xs = [to_double(q) for q in quad_to_float(x)]
    assert all(a == -b for a, b in zip(xs, xs[::-1])) and xs[0] < 0  # ascending from -1


@pytest.mark.parametrize("i", range(len(GRID_CHEBS)))
def test_cheby_weights_bit_exact_vs_reference(golden, i):
    g = golden("grid_nodes_weights.npz")
    n, rmin, rmax = GRID_CHEBS[i]
    assert np.array_equal(O.cheby_grid(n, rmin, rmax), g[f"cheb_grid_{i}"])
    assert np.array_equal(O.cheby_weights(n, rmin, rmax), g[f"cheb_weights_{i}"])


def test_cheby_weights_integrate_the_shell_volume():
    """sum_i w_i f(r_i) approximates 3 / (rmax^3 - rmin^3) int f r^2 dr: 1 for f = 1 (up
    to the halved end points and the quadrature error of the sqrt(1 - x^2) factor) and
    the mean of r^2 weighted f = r for f = r."""
    n, rmin, rmax = 256, 0.35, 1.0
    w = O.cheby_weights(n, rmin, rmax)
    r = O.cheby_grid(n, rmin, rmax)
    assert abs(w.sum() - 1.0) < 1e-3
    want = 3.0 / (rmax ** 3 - rmin ** 3) * (rmax ** 4 - rmin ** 4) / 4.0
    assert abs((w * r).sum() - want) < 1e-3
    # a dropped sqrt(1 - x^2) factor (or r^2) moves the sum far from 1
    x = np.cos((np.arange(n) + 0.5) * np.pi / n)
    assert abs((w / np.sqrt(1 - x * x)).sum() - 1.0) > 0.1
