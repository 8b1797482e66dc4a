"""Pins of the oracle's scalar pieces: fixed point, RNG, seeding.

Each check is against something the mathematics or a standard fixes, not a
retyping of the oracle: integer arithmetic done in Python, the C++
standard's known answer for mt19937, the xorshift128 recurrence evaluated
with Python integers, and the host C library called directly."""
import ctypes
import ctypes.util

import numpy as np
import pytest

import netutil


def test_fx_toFix_truncates_toward_zero(oracle_mod):
    # src/nemo/fixedpoint.cpp:18-29: (int)(f * 2^20), truncation
    for w in [0.0, 0.5, -0.5, 0.1, -0.1, 0.3333333, -0.9999999, 1e-7, -1e-7, 2047.999, -2047.999,
              np.float32(1.0) / 3, 123.456]:
        assert oracle_mod.fx_toFix(float(w)) == netutil.fix20(w), w


def test_fx_toFix_overflow_throws(oracle_mod):
    # |int(f)| >= 2^11 does not fit Q11.20
    for w in [2048.0, -2048.0, 1e6, -3000.5]:
        with pytest.raises(oracle_mod.OracleError) as e:
            oracle_mod.fx_toFix(w)
        assert e.value.status == 6
    assert oracle_mod.fx_toFix(2047.9) == netutil.fix20(2047.9)


def test_wfx_saturation(oracle_mod):
    for v in [0, 1, -1, 2**31 - 1, 2**31, 2**40, -2**31, -2**31 - 1, -2**45, 123456789, -987654321]:
        assert oracle_mod.wfx_toFloat(v) == netutil.sat_to_float(v), v


def test_fx_mul_arithmetic_shift(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        a = int(rng.integers(-2**31, 2**31))
        b = int(rng.integers(-2**31, 2**31))
        expect = (a * b) >> 20                       # Python >> floors, as an arithmetic shift
        expect = ((expect + 2**31) % 2**32) - 2**31  # truncation to int32
        assert oracle_mod.fx_mul(a, b) == expect


def test_mt19937_known_answer(oracle_mod):
    # C++ standard [rand.predef]: the 10000th output of a default-constructed
    # mt19937 is 4123659995
    assert oracle_mod.mt19937_nth(10000) == 4123659995


def _mt19937(n):
    return netutil.mt19937(n)


def test_initialise_rng_closed_form(oracle_mod):
    # src/nemo/RNG.cpp:38-63 with Boost >= 1.47: seed word = mt() >> 1, and a
    # range [lo, hi] skips the 4*lo draws of the lower neurons
    draws = _mt19937(4 * 40)
    got = oracle_mod.init_rng(0, 39)
    assert np.array_equal(got.reshape(-1), np.array([d >> 1 for d in draws], np.uint32))
    got2 = oracle_mod.init_rng(25, 39)
    assert np.array_equal(got2, got[25:40])


def test_urand_xorshift128(oracle_mod):
    state = [123456789, 362436069, 521288629, 88675123]  # Marsaglia's example seed
    s = list(state)
    expect = []
    for _ in range(1000):
        t = (s[0] ^ ((s[0] << 11) & 0xffffffff)) & 0xffffffff
        s = [s[1], s[2], s[3], (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8))]
        expect.append(s[3])
    got = oracle_mod.urand_seq(state, 1000)
    assert np.array_equal(got, np.array(expect, np.uint32))
    # Marsaglia (2003), xorshift128 with this seed: first output 3701687786
    assert expect[0] == 3701687786


def test_nrand_composition(oracle_mod):
    """nrand = sinf(u1*C) * sqrtf(-2*logf(1 - u2*2^-32)) in float32, with the
    host libm functions called directly through ctypes."""
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    for name in ("logf", "sinf", "sqrtf"):
        getattr(libm, name).restype = ctypes.c_float
        getattr(libm, name).argtypes = [ctypes.c_float]
    state = [5, 7, 11, 13]
    u = oracle_mod.urand_seq(state, 400)
    got = oracle_mod.nrand_seq(state, 200)
    C = np.float32(1.4629180792671596810513378043098e-9)
    S = np.float32(0.00000000023283064365386962890625)
    for k in range(200):
        a = np.float32(np.float32(u[2 * k]) * C)
        b = np.float32(np.float32(u[2 * k + 1]) * S)
        lg = np.float32(libm.logf(float(np.float32(np.float32(1) - b))))
        r = np.float32(libm.sqrtf(float(np.float32(np.float32(-2) * lg))))
        expect = np.float32(np.float32(libm.sinf(float(a))) * r)
        assert got[k] == expect or (np.isnan(got[k]) and np.isnan(expect)), k
