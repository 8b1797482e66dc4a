"""nemo_hip_run_static, the allocation-free form of nemo_hip_run, through the
C ABI: its raster (written to a device buffer the test allocates) and its
delivered-synapse count equal nemo_hip_run's on an identical simulation, on
two calls in a row -- the second with fewer cycles and a smaller stimulus
schedule -- and a schedule larger than the scratch area is refused before
anything runs."""
import ctypes

import numpy as np
import pytest

from nemosim_amd import _lib, examples
from parity import hip_sim

pytestmark = pytest.mark.gpu

HIP_MEMCPY_D2H = 2


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMalloc.restype = ctypes.c_int
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipFree.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    return hip


def _schedule(nsteps, fstim, istim):
    """A nemo_schedule_t over per-cycle lists; returns (struct, arrays kept alive)."""
    fptr = np.zeros(nsteps + 1, np.uint32)
    fidx = []
    iptr = np.zeros(nsteps + 1, np.uint32)
    iidx, ival = [], []
    for k in range(nsteps):
        fidx.extend(fstim[k])
        fptr[k + 1] = len(fidx)
        for n, c in istim[k]:
            iidx.append(n)
            ival.append(c)
        iptr[k + 1] = len(iidx)
    fidx = np.ascontiguousarray(fidx if fidx else [0], np.uint32)
    iidx = np.ascontiguousarray(iidx if iidx else [0], np.uint32)
    ival = np.ascontiguousarray(ival if ival else [0.0], np.float32)
    s = _lib.ScheduleT()
    s.fstim_ptr = fptr.ctypes.data_as(_lib.c_uint_p)
    s.fstim_idx = fidx.ctypes.data_as(_lib.c_uint_p)
    s.istim_ptr = iptr.ctypes.data_as(_lib.c_uint_p)
    s.istim_idx = iidx.ctypes.data_as(_lib.c_uint_p)
    s.istim_val = ival.ctypes.data_as(_lib.c_float_p)
    return s, (fptr, fidx, iptr, iidx, ival)


def test_static_run_equals_run(nemo_mod):
    lib = _lib.load()
    hip = _hip()
    spec = examples.izhikevich_80_20(3000, 120, seed=41)
    plain = hip_sim(nemo_mod, spec)
    static = hip_sim(nemo_mod, spec)
    assert static.info()["fused"] == 1
    wps = ctypes.c_size_t()
    assert lib.nemo_hip_read_raster(static._h, None, 0, ctypes.byref(wps)) == 0
    wps = wps.value
    first = 400
    buf = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(buf), first * wps * 8) == 0
    try:
        rng = np.random.default_rng(5)
        for nsteps, nf, ni in [(first, 3, 4), (150, 1, 2)]:
            fstim = [sorted(set(rng.integers(0, spec.n, size=nf).tolist())) for _ in range(nsteps)]
            istim = [[(int(i), float(np.float32(rng.uniform(0, 12)))) for i in rng.integers(0, spec.n, size=ni)]
                     for _ in range(nsteps)]
            ev_plain = plain.run(nsteps, fstim=fstim, istim=istim)
            want = plain.read_raster()
            sched, keep = _schedule(nsteps, fstim, istim)
            ev = ctypes.c_ulonglong()
            st = lib.nemo_hip_run_static(static._h, nsteps, ctypes.byref(sched), buf, ctypes.byref(ev))
            assert st == 0, lib.nemo_strerror().decode()
            got = np.zeros(nsteps * wps, np.uint64)
            assert hip.hipMemcpy(got.ctypes.data, buf, got.nbytes, HIP_MEMCPY_D2H) == 0
            assert np.array_equal(got.reshape(nsteps, wps), want), "raster differs (%d cycles)" % nsteps
            assert ev.value == ev_plain
            del keep
        # the raster of a static run is in the caller's buffer, not the library's
        total = ctypes.c_size_t()
        assert lib.nemo_hip_read_fired(static._h, None, None, 0, ctypes.byref(total)) != 0
        # a schedule larger than the scratch area is refused before anything runs
        n_big = 600
        fstim = [list(range(2000)) for _ in range(n_big)]
        istim = [[] for _ in range(n_big)]
        sched, keep = _schedule(n_big, fstim, istim)
        before = static.elapsed_simulation()
        st = lib.nemo_hip_run_static(static._h, n_big, ctypes.byref(sched), buf, None)
        assert st == 6, (st, lib.nemo_strerror().decode())   # NEMO_INVALID_INPUT
        assert b"scratch" in lib.nemo_strerror()
        assert static.elapsed_simulation() == before
        # and both simulations still agree afterwards
        plain.run(50)
        want = plain.read_raster()
        assert lib.nemo_hip_run_static(static._h, 50, None, buf, None) == 0
        got = np.zeros(50 * wps, np.uint64)
        assert hip.hipMemcpy(got.ctypes.data, buf, got.nbytes, HIP_MEMCPY_D2H) == 0
        assert np.array_equal(got.reshape(50, wps), want)
    finally:
        hip.hipFree(buf)
