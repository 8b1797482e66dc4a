"""The backend's logf / sinf (nemosim_amd/csrc/device/libm_f32.h) against the
host C library on every input NeMo's nrand() can produce.

CPU half: the host build of the same source (nemo_hip_eval_libm with
on_device=0) -- no GPU needed. The GPU half is in test_gpu_parity.py."""
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


@pytest.mark.parametrize("kind", ["logf", "sinf"])
def test_host_build_matches_libc_exhaustively(nemo_mod, oracle_mod, kind):
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
            got = nemo_mod.eval_libm(1, x, on_device=False)
        bad += int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
    assert bad == 0
