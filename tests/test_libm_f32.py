"""The backend's logf / sinf (nemosim_amd/csrc/device/libm_f32.h) against the
host C library on every input NeMo's nrand() can produce, sinf also on the
wider domain of the Kuramoto model's phase differences, and the Kuramoto
phase wrap against the reference's expression with the host fmodf.

CPU half: the host build of the same source (nemo_hip_eval_libm with
on_device=0) -- no GPU needed. The GPU half is in test_gpu_parity.py and
test_gpu_kuramoto.py."""
import numpy as np
import pytest

import nrand_domain as dom


def test_domain_enumeration():
    # the enumeration covers float(u) for sampled u exactly
    rng = np.random.default_rng(0)
    u = rng.integers(0, 2**32, size=200000, dtype=np.uint64)
    f = u.astype(np.float32)
    allf = dom.sin_float_u(0, dom.SIN_COUNT)
    assert np.all(np.isin(f, allf))
    assert allf[-1] == np.float32(2**32)
    assert len(np.unique(allf)) == dom.SIN_COUNT
    # 1 - b lies on the 2^-24 grid
    b = (f * np.float32(2.0**-32)).astype(np.float32)
    x = (np.float32(1) - b).astype(np.float32)
    m = x.astype(np.float64) * 2.0**24
    assert np.all(m == np.round(m))


@pytest.mark.parametrize("kind", ["logf", "sinf", "sinf_any"])
def test_host_build_matches_libc_exhaustively(nemo_mod, oracle_mod, kind):
    """logf and sinf as the RNG uses them, and the any-argument (fused) sinf
    of the Kuramoto model on the same arguments: both forms agree there."""
    chunk = 1 << 23
    total = dom.LOG_COUNT if kind == "logf" else dom.SIN_COUNT
    bad = 0
    for start in range(0, total, chunk):
        cnt = min(chunk, total - start)
        if kind == "logf":
            x = dom.log_inputs(start, cnt)
            ref = oracle_mod.libm_logf_grid(start, cnt)
            got = nemo_mod.eval_libm(0, x, on_device=False)
        else:
            x = dom.sin_inputs(start, cnt)
            ref = oracle_mod.libm_sinf(x)
            got = nemo_mod.eval_libm(1 if kind == "sinf" else 3, x, on_device=False)
        bad += int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
    assert bad == 0


def wide_sin_inputs(count, seed):
    """Arguments beyond the RNG's: every float bit pattern class (random
    32-bit words, non-finite ones dropped), uniform in [-130, 130] (both
    reduction paths and their boundary at 120), and the 2^12 floats nearest
    each multiple k*pi/2 for |k| <= 200 (where the reduction cancels)."""
    rng = np.random.default_rng(seed)
    raw = rng.integers(0, 2**32, size=count, dtype=np.uint64).astype(np.uint32).view(np.float32)
    uni = rng.uniform(-130.0, 130.0, size=count).astype(np.float32)
    near = []
    for k in range(-200, 201):
        c = np.float32(k * np.pi / 2)
        b = np.int64(np.float32(c).view(np.int32))
        near.append((np.arange(-2048, 2048) + b).astype(np.int32).view(np.float32))
    x = np.concatenate([raw, uni, np.concatenate(near)])
    return x[np.isfinite(x)]


def test_inv_pio4_table_is_two_over_pi():
    """The 2/pi bit windows of the large-argument reduction, recomputed from
    pi (Machin's formula in integers)."""
    import os
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    text = open(os.path.join(here, "..", "nemosim_amd", "csrc", "device", "libm_f32.h")).read()
    body = re.search(r"INV_PIO4_BITS\[24\] = \{(.*?)\};", text, re.S).group(1)
    table = [int(v.rstrip("u"), 16) for v in re.findall(r"0x[0-9a-fA-F]+u", body)]

    def arctan_inv(x, one):
        s = t = one // x
        n, sign, x2 = 1, -1, x * x
        while t:
            t //= x2
            n += 2
            s += sign * (t // n)
            sign = -sign
        return s

    bits = 1024
    one = 1 << bits
    pi = 4 * (4 * arctan_inv(5, one) - arctan_inv(239, one))
    two_over_pi = (2 * one * one) // pi
    expect = [(two_over_pi >> (bits - 8 * (i + 1))) & 0xffffffff for i in range(24)]
    assert table == expect


def test_host_build_sinf_wide_domain(nemo_mod, oracle_mod):
    x = wide_sin_inputs(1 << 21, seed=11)
    ref = oracle_mod.libm_sinf(x)
    got = nemo_mod.eval_libm(3, x, on_device=False)
    bad = np.nonzero(got.view(np.uint32) != ref.view(np.uint32))[0]
    assert len(bad) == 0, (x[bad[:5]], ref[bad[:5]], got[bad[:5]])


def test_host_build_kuramoto_wrap(nemo_mod, oracle_mod):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-20, 20, 1 << 20), rng.uniform(-1e5, 1e5, 1 << 18),
                        [-0.0, 0.0, -1e-38, 2 * np.pi, -2 * np.pi]]).astype(np.float32)
    ref = oracle_mod.kuramoto_wrap(x)
    got = nemo_mod.eval_libm(2, x, on_device=False)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
