"""Exact rewrites of the reference's float/double arithmetic in the tracer (qmc_math.h,
qmc_trace_core.h): the float forms of the double Epsilon compares (kEpsF, used by the ray/parallelogram
tests, the DDA's abs_divide and the cylinder test) decide exactly as the double compares do, for every
float near the boundaries (and trivially for all others); the cylinder's v = float(double(acos) / 2 pi)
equals float(double(acos) * (1 / 2 pi)) for every float the arc cosine can return."""
import math

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


def test_cylinder_v_by_reciprocal_for_every_float_in_0_pi():
    """Every float a in [0, float(pi)] (1,078,530,012 of them, zero and subnormals included): float(a / c)
    == float(a * (1 / c)) in double, c the double nearest 2 pi. Both sides round twice (to double, then
    to float); a float boundary between the two double results would separate them, and there is none."""
    c = 2 * math.pi
    inv = 1.0 / c
    top = int(np.array([np.float32(math.pi)], np.float32).view(np.uint32)[0])
    step = 1 << 25
    for s in range(0, top + 1, step):
        a = np.arange(s, min(s + step, top + 1), dtype=np.uint32).view(np.float32).astype(np.float64)
        q1 = (a / c).astype(np.float32).view(np.uint32)
        q2 = (a * inv).astype(np.float32).view(np.uint32)
        assert np.array_equal(q1, q2), f"mismatch in the block from bit pattern {s:#x}"


def test_cylinder_column_lookup_equals_the_double_arc_cosine():
    """The trace's cylinder column (an arc cosine estimate corrected by the receiver's breakpoint table,
    qmc_math.h: cylinder_column) equals (int)(v * W) from the double arc cosine, for both signs of the
    hit's z offset, on every 61st float cosine of [-1, 1] and the 513 floats around each breakpoint, for
    maps 1, 7, 512 (the config-3 field) and 32768 (the largest with a table) columns wide. Every float
    of [-1, 1] was checked once with stride 1 for these widths (4.26e9 pairs each, no mismatch)."""
    from solarenergy_chier_amd.scene import check_cylinder_columns

    for width in (1, 7, 512, 32768):
        n, bad, err = check_cylinder_columns(width, 61)
        assert n > 6.9e7 and bad == 0, (width, n, bad)
        assert err < 6.8e-5, err  # Abramowitz and Stegun 4.4.45: 6.7e-5, plus float rounding
