"""CPU checks of numerical identities the HIP kernels rely on for bit-exact parity.

The kernels divide texture coordinates by the image size, and the 5-point stencils by 12, as
float(double(a) * (1/double(n))) instead of the IEEE float division a / float(n). The two agree
for every float a and integer 1 <= n <= 2^24: a power of two divides exactly; otherwise a / n is
never a float midpoint and stays at least 2^-25 / n (relative) away from one, while the double
product is within 2^-52 of a / n. These tests exercise that claim on the values the kernels see
and on random floats over a wide exponent range.
"""
import numpy as np


def _div_n(a, n):
    return (a.astype(np.float64) * (1.0 / np.float64(n))).astype(np.float32)


def test_texel_centre_coordinates_all_sizes():
    for n in list(range(1, 4097)) + [6000, 4000, 3000, 2000, 1500, 1000, 750, 500, 12000, 8000, 4032, 3024]:
        x = np.arange(n, dtype=np.float32) + np.float32(0.5)
        np.testing.assert_array_equal(_div_n(x, n), x / np.float32(n))


def test_warped_coordinates():
    rng = np.random.default_rng(1)
    for n in (6000, 4000, 3001, 1024, 777, 13):
        x = (np.arange(n, dtype=np.float32) + np.float32(0.5))[None, :]
        f = rng.uniform(-70, 70, (64, 1)).astype(np.float32)
        a = (x + f).astype(np.float32)
        np.testing.assert_array_equal(_div_n(a, n), a / np.float32(n))


def test_random_floats_and_divisors():
    rng = np.random.default_rng(2)
    a = (rng.uniform(1, 2, 2_000_000) * np.exp2(rng.integers(-60, 60, 2_000_000))).astype(np.float32)
    a *= np.where(rng.random(a.size) < 0.5, -1, 1).astype(np.float32)
    for n in list(range(1, 64)) + list(rng.integers(64, 1 << 24, 40)):
        np.testing.assert_array_equal(_div_n(a, int(n)), a / np.float32(n))


def test_stencil_division_by_12():
    rng = np.random.default_rng(3)
    t = rng.standard_normal(4_000_000).astype(np.float32) * np.float32(3.0)
    np.testing.assert_array_equal(_div_n(t, 12), t / np.float32(12.0))
