"""The Python checker of the exact image sums, pinned on the CPU: where the
float64 sum of float32 terms is exact (terms within 29 binary orders of each
other), numpy's float64 -> float32 conversion is one correctly rounded
operation, and the checker must agree with it; ties, subnormals, overflow and
the special values by hand."""
import math

import numpy as np

from exact_sum import exact_sum_f32, round_units, units


def test_units_round_trip_every_class():
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.standard_normal(2000).astype(np.float32) * np.float32(2.0) ** rng.integers(-120, 120, 2000).astype(np.float32),
                           np.array([1.4e-45, -1.4e-45, 1.17549435e-38, 3.4028235e38, -3.4028235e38, 0.0, 1.0], np.float32)])
    vals = vals[np.isfinite(vals)]
    for v in vals:
        assert round_units(units(v)) == v or (v == 0 and round_units(units(v)) == 0)


def test_agrees_with_numpy_where_float64_is_exact():
    rng = np.random.default_rng(2)
    for trial in range(300):
        n = int(rng.integers(1, 40))
        scale = np.float32(2.0) ** np.float32(rng.integers(-100, 100))
        col = (rng.standard_normal(n).astype(np.float32) * np.float32(2.0) ** rng.integers(0, 12, n).astype(np.float32)) * scale
        exact64 = math.fsum(float(v) for v in col)   # exact: spread < 29 orders
        assert exact_sum_f32(col) == np.float32(exact64), (trial, col)


def test_ties_subnormals_overflow_and_specials():
    f = np.float32
    assert exact_sum_f32([f(16777216.0), f(1.0)]) == f(16777216.0)      # tie to even (down)
    assert exact_sum_f32([f(16777218.0), f(1.0)]) == f(16777220.0)      # tie to even (up)
    assert exact_sum_f32([f(1e30), f(1.0), f(-1e30)]) == f(1.0)         # cancellation
    assert exact_sum_f32([f(1.4e-45)] * 3) == np.float32(3) * f(1.4e-45)  # subnormal, exact
    assert exact_sum_f32([f(3.4028235e38), f(3.4028235e38)]) == np.inf    # overflow
    assert exact_sum_f32([f(-3.4028235e38), f(-2.0e38)]) == -np.inf
    assert np.isnan(exact_sum_f32([f(1.0), f(np.nan)]))
    assert np.isnan(exact_sum_f32([f(np.inf), f(-np.inf)]))
    assert exact_sum_f32([f(np.inf), f(-1e38)]) == np.inf
    assert exact_sum_f32([]) == 0.0
