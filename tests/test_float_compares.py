"""The float forms of the reference's double Epsilon compares (qmc_math.h: kEpsF, used by the
ray/parallelogram tests, the DDA's abs_divide and the cylinder test) decide exactly as the double
compares do, for every float near the boundaries (and trivially for all others)."""
import numpy as np

EPS = 1e-5                      # the reference's Epsilon, a double
EPS_F = np.float32(EPS)         # kEpsF = (float)kEps, rounded to nearest


def _floats_around(v, n=4096):
    """The 2n + 1 consecutive floats centred on float v."""
    b = np.array([v], np.float32).view(np.int32)[0]
    k = np.arange(b - n, b + n + 1, dtype=np.int64)
    return k.astype(np.int32).view(np.float32)


def test_eps_f_is_the_float_below_eps():
    assert float(EPS_F) < EPS
    assert float(np.nextafter(EPS_F, np.float32(1.0))) > EPS


def test_float_compares_equal_double_compares():
    for centre in (EPS_F, -EPS_F):
        x = _floats_around(centre)
        xd = x.astype(np.float64)
        np.testing.assert_array_equal(xd < EPS, x <= EPS_F)                 # t < Epsilon, |det| < Epsilon
        np.testing.assert_array_equal(xd <= EPS, x <= EPS_F)                # abs_divide's num <= Epsilon
        np.testing.assert_array_equal(xd >= -EPS, x >= -EPS_F)              # abs_divide's num >= -Epsilon
        np.testing.assert_array_equal(xd < -EPS, x < -EPS_F)                # cylinder tp < -Epsilon
        ax = np.abs(x)
        np.testing.assert_array_equal(ax.astype(np.float64) < EPS, ax <= EPS_F)
