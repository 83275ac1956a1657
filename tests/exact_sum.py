"""Correctly rounded float32 sums in Python integers: the checker of the
exact image accumulation. Every float32 is an integer multiple of 2^-149, so a
sum of them is one too; it is rounded to the nearest float32, ties to even."""
import math

import numpy as np


def units(v) -> int:
    """float32 v as an integer number of 2^-149 (finite v only)."""
    m, e = math.frexp(float(np.float32(v)))   # exact: v = m * 2^e, m has <= 24 bits
    return int(m * (1 << 24)) << (e - 24 + 149) if e - 24 + 149 >= 0 else int(m * (1 << 24)) >> (24 - e - 149)


def round_units(n: int) -> np.float32:
    """The float32 nearest n * 2^-149 (ties to even; +-inf past the range)."""
    if n == 0:
        return np.float32(0.0)
    sign = -1.0 if n < 0 else 1.0
    a = abs(n)
    b = a.bit_length() - 1
    if b <= 23:
        mant, ex = a, 0
    else:
        sh = b - 23
        mant = a >> sh
        rem = a - (mant << sh)
        half = 1 << (sh - 1)
        if rem > half or (rem == half and mant & 1):
            mant += 1
        ex = sh
        if mant >> 24:
            mant >>= 1
            ex += 1
    if mant >= (1 << 23) and ex + 1 >= 255:
        return np.float32(sign * np.inf)
    return np.float32(sign * math.ldexp(mant, ex - 149))


def exact_sum_f32(col) -> np.float32:
    """The float32 nearest the exact sum of a float32 vector (NaN for a NaN
    or for both infinities, +-inf for one infinity)."""
    col = np.asarray(col, np.float32)
    if np.isnan(col).any() or (np.isposinf(col).any() and np.isneginf(col).any()):
        return np.float32(np.nan)
    if np.isposinf(col).any():
        return np.float32(np.inf)
    if np.isneginf(col).any():
        return np.float32(-np.inf)
    return round_units(sum(units(v) for v in col))
