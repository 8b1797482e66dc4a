"""Every input nrand() can hand to logf and sinf (src/nemo/RNG.cpp:25-33).

logf sees x = 1 - float(u)*2^-32 for a 32-bit u; in float32 these are exactly
the multiples m * 2^-24, m = 0..2^24 (b's spacing near 1 is 2^-24, and below
0.5 every x is already on that grid).

sinf sees a = float(u) * 1.4629180792671596810513378043098e-9f where float(u)
is one of the distinct float32 values of a 32-bit unsigned: all integers
below 2^24, then 2^23 mantissas for each exponent 24..31, then 2^32.
"""
import numpy as np

C_SIN = np.float32(1.4629180792671596810513378043098e-9)
LOG_COUNT = (1 << 24) + 1


def log_inputs(first, count):
    m = np.arange(first, first + count, dtype=np.float64)
    return (m * 2.0 ** -24).astype(np.float32)


def sin_float_u(start, count):
    """Distinct float(u) values with ordinal start..start+count-1."""
    k = np.arange(start, start + count, dtype=np.int64)
    out = np.empty(count, np.float32)
    small = k < (1 << 24)
    out[small] = k[small].astype(np.float32)
    big = ~small
    j = k[big] - (1 << 24)
    e = 24 + (j >> 23)
    mant = (j & ((1 << 23) - 1)) | (1 << 23)
    out[big] = np.ldexp(mant.astype(np.float64), (e - 23).astype(np.int64)).astype(np.float32)
    return out


SIN_COUNT = (1 << 24) + 8 * (1 << 23) + 1


def sin_inputs(start, count):
    f = sin_float_u(start, count)
    return (f * C_SIN).astype(np.float32)
