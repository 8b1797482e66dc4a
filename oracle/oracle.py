"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of oracle/libnemo_oracle.so, the
CPU restatement of NeMo's CPU backend. Imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
product (nemosim_amd/)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnemo_oracle.so")

_c_uint_p = ctypes.POINTER(ctypes.c_uint)
_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_u64_p = ctypes.POINTER(ctypes.c_uint64)
_c_u8_p = ctypes.POINTER(ctypes.c_ubyte)
_c_i32_p = ctypes.POINTER(ctypes.c_int32)
_c_size_p = ctypes.POINTER(ctypes.c_size_t)

MODEL_IDS = {"Izhikevich": 0, "IF_curr_exp": 1, "IF_curr_exp_noise": 2, "PoissonSource": 3, "Input": 4,
             "Kuramoto": 5}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "oracle_strerror": (ctypes.c_char_p, []),
        "oracle_new": (ctypes.c_void_p, []),
        "oracle_delete": (None, [ctypes.c_void_p]),
        "oracle_num_threads": (ctypes.c_int, []),
        "oracle_set_threads": (None, [ctypes.c_int]),
        "oracle_add_neurons": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, _c_uint_p, ctypes.c_size_t,
                                              ctypes.c_uint, _c_float_p]),
        "oracle_add_synapses": (ctypes.c_int, [ctypes.c_void_p, _c_uint_p, _c_uint_p, _c_uint_p, _c_float_p,
                                               _c_u8_p, ctypes.c_size_t, _c_u64_p]),
        "oracle_add_synapses_csr": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _c_u64_p, _c_uint_p,
                                                   _c_u8_p, _c_float_p, _c_u8_p]),
        "oracle_set_stdp": (ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_size_t, _c_float_p,
                                           ctypes.c_size_t, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_float]),
        "oracle_finalize": (ctypes.c_int, [ctypes.c_void_p]),
        "oracle_neuron_count": (ctypes.c_size_t, [ctypes.c_void_p]),
        "oracle_step": (ctypes.c_int, [ctypes.c_void_p, _c_uint_p, ctypes.c_size_t, _c_uint_p, _c_float_p,
                                       ctypes.c_size_t, _c_uint_p, _c_size_p]),
        "oracle_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, _c_uint_p, _c_uint_p, ctypes.c_size_t,
                                      _c_size_p]),
        "oracle_apply_stdp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
        "oracle_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, _c_float_p]),
        "oracle_get_param": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, _c_float_p]),
        "oracle_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
        "oracle_set_param": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
        "oracle_set_neuron": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, _c_float_p]),
        "oracle_get_membrane_potential": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, _c_float_p]),
        "oracle_get_state_all": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, _c_float_p]),
        "oracle_get_rng_all": (ctypes.c_int, [ctypes.c_void_p, _c_uint_p]),
        "oracle_get_synapse_weight": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _c_float_p]),
        "oracle_get_synapse_target": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _c_uint_p]),
        "oracle_get_synapse_delay": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _c_uint_p]),
        "oracle_get_synapse_plastic": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _c_u8_p]),
        "oracle_get_weights_fx": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _c_i32_p]),
        "oracle_elapsed_simulation": (ctypes.c_ulong, [ctypes.c_void_p]),
        "oracle_elapsed_wall_ms": (ctypes.c_double, [ctypes.c_void_p]),
        "oracle_reset_timer": (None, [ctypes.c_void_p]),
        "oracle_fx_toFix": (ctypes.c_int, [ctypes.c_float, _c_i32_p]),
        "oracle_wfx_toFloat": (ctypes.c_float, [ctypes.c_int64]),
        "oracle_fx_mul": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
        "oracle_urand_seq": (None, [_c_uint_p, ctypes.c_size_t, _c_uint_p]),
        "oracle_nrand_seq": (None, [_c_uint_p, ctypes.c_size_t, _c_float_p]),
        "oracle_libm_logf_grid": (None, [ctypes.c_uint32, ctypes.c_size_t, _c_float_p]),
        "oracle_libm_sinf": (None, [_c_float_p, ctypes.c_size_t, _c_float_p]),
        "oracle_kuramoto_wrap": (None, [_c_float_p, ctypes.c_size_t, _c_float_p]),
        "oracle_init_rng": (None, [ctypes.c_uint, ctypes.c_uint, _c_uint_p]),
        "oracle_mt19937_nth": (ctypes.c_uint32, [ctypes.c_uint]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class OracleError(RuntimeError):
    def __init__(self, status, message):
        RuntimeError.__init__(self, "%s (status %d)" % (message, status))
        self.status = status


def _check(st):
    if st != 0:
        raise OracleError(st, load().oracle_strerror().decode())


def _f(a):
    return a.ctypes.data_as(_c_float_p)


def _u(a):
    return a.ctypes.data_as(_c_uint_p)


class Oracle(object):
    """Network + simulation in one object: add neurons/synapses, set STDP,
    finalize(), then step()."""

    def __init__(self):
        self.lib = load()
        self.h = self.lib.oracle_new()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_delete(self.h)
            self.h = None

    def add_neuron(self, model, idx, *args):
        a = np.ascontiguousarray(args, np.float32)
        i = np.array([idx], np.uint32)
        _check(self.lib.oracle_add_neurons(self.h, MODEL_IDS[model], _u(i), 1, len(a), _f(a)))

    def add_neurons(self, model, idx, args):
        idx = np.ascontiguousarray(idx, np.uint32)
        args = np.ascontiguousarray(args, np.float32)
        _check(self.lib.oracle_add_neurons(self.h, MODEL_IDS[model], _u(idx), len(idx), args.shape[1], _f(args)))

    def add_synapse(self, s, t, d, w, plastic=0):
        sid = np.zeros(1, np.uint64)
        _check(self.lib.oracle_add_synapses(self.h, _u(np.array([s], np.uint32)), _u(np.array([t], np.uint32)),
                                            _u(np.array([d], np.uint32)), _f(np.array([w], np.float32)),
                                            np.array([plastic], np.uint8).ctypes.data_as(_c_u8_p), 1,
                                            sid.ctypes.data_as(_c_u64_p)))
        return int(sid[0])

    def add_synapses_csr(self, rowptr, targets, delays, weights, plastic=None):
        rowptr = np.ascontiguousarray(rowptr, np.uint64)
        targets = np.ascontiguousarray(targets, np.uint32)
        delays = np.ascontiguousarray(delays, np.uint8)
        weights = np.ascontiguousarray(weights, np.float32)
        pl = None if plastic is None else np.ascontiguousarray(plastic, np.uint8)
        _check(self.lib.oracle_add_synapses_csr(self.h, len(rowptr) - 1, rowptr.ctypes.data_as(_c_u64_p),
                                                _u(targets), delays.ctypes.data_as(_c_u8_p), _f(weights),
                                                None if pl is None else pl.ctypes.data_as(_c_u8_p)))

    def set_stdp_function(self, pre, post, min_e, max_e, min_i, max_i):
        pre = np.ascontiguousarray(pre, np.float32)
        post = np.ascontiguousarray(post, np.float32)
        _check(self.lib.oracle_set_stdp(self.h, _f(pre), len(pre), _f(post), len(post),
                                        min_e, max_e, min_i, max_i))

    def finalize(self):
        _check(self.lib.oracle_finalize(self.h))
        self.n = int(self.lib.oracle_neuron_count(self.h))

    def step(self, fstim=None, istim=None):
        fs = np.ascontiguousarray(fstim if fstim is not None else [], np.uint32)
        if istim:
            ii = np.ascontiguousarray([p[0] for p in istim], np.uint32)
            iv = np.ascontiguousarray([p[1] for p in istim], np.float32)
        else:
            ii = np.zeros(0, np.uint32)
            iv = np.zeros(0, np.float32)
        out = np.zeros(max(self.n, 1), np.uint32)
        cnt = ctypes.c_size_t()
        _check(self.lib.oracle_step(self.h, _u(fs), len(fs), _u(ii), _f(iv), len(ii), _u(out), ctypes.byref(cnt)))
        return out[:cnt.value].tolist()

    def run(self, nsteps, chunk=100):
        """nsteps cycles without stimulus: (fired count per cycle, fired ids)."""
        counts = np.zeros(nsteps, np.uint32)
        parts = []
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            cap = max(1, self.n * k)
            ids = np.zeros(cap, np.uint32)
            total = ctypes.c_size_t()
            c = np.zeros(k, np.uint32)
            _check(self.lib.oracle_run(self.h, k, _u(c), _u(ids), cap, ctypes.byref(total)))
            counts[done:done + k] = c
            parts.append(ids[:total.value].copy())
            done += k
        return counts, (np.concatenate(parts) if parts else np.zeros(0, np.uint32))

    def apply_stdp(self, reward):
        _check(self.lib.oracle_apply_stdp(self.h, reward))

    def get_neuron_state(self, g, var):
        v = ctypes.c_float()
        _check(self.lib.oracle_get_state(self.h, g, var, ctypes.byref(v)))
        return v.value

    def get_neuron_parameter(self, g, p):
        v = ctypes.c_float()
        _check(self.lib.oracle_get_param(self.h, g, p, ctypes.byref(v)))
        return v.value

    def set_neuron_state(self, g, var, val):
        _check(self.lib.oracle_set_state(self.h, g, var, val))

    def set_neuron_parameter(self, g, p, val):
        _check(self.lib.oracle_set_param(self.h, g, p, val))

    def set_neuron(self, g, *args):
        a = np.ascontiguousarray(args, np.float32)
        _check(self.lib.oracle_set_neuron(self.h, g, len(a), _f(a)))

    def get_membrane_potential(self, g):
        v = ctypes.c_float()
        _check(self.lib.oracle_get_membrane_potential(self.h, g, ctypes.byref(v)))
        return v.value

    def state_all(self, var):
        out = np.zeros(self.n, np.float32)
        _check(self.lib.oracle_get_state_all(self.h, var, _f(out)))
        return out

    def rng_all(self):
        out = np.zeros(4 * self.n, np.uint32)
        _check(self.lib.oracle_get_rng_all(self.h, _u(out)))
        return out.reshape(self.n, 4)

    def get_synapse_weight(self, sid):
        v = ctypes.c_float()
        _check(self.lib.oracle_get_synapse_weight(self.h, sid, ctypes.byref(v)))
        return v.value

    def get_synapse_target(self, sid):
        v = ctypes.c_uint()
        _check(self.lib.oracle_get_synapse_target(self.h, sid, ctypes.byref(v)))
        return v.value

    def get_synapse_delay(self, sid):
        v = ctypes.c_uint()
        _check(self.lib.oracle_get_synapse_delay(self.h, sid, ctypes.byref(v)))
        return v.value

    def get_synapse_plastic(self, sid):
        v = ctypes.c_ubyte()
        _check(self.lib.oracle_get_synapse_plastic(self.h, sid, ctypes.byref(v)))
        return v.value

    def weights_fx(self, nsrc, nsyn):
        out = np.zeros(nsyn, np.int32)
        _check(self.lib.oracle_get_weights_fx(self.h, nsrc, out.ctypes.data_as(_c_i32_p)))
        return out

    def elapsed_wall_ms(self):
        return self.lib.oracle_elapsed_wall_ms(self.h)

    def reset_timer(self):
        self.lib.oracle_reset_timer(self.h)


def from_spec(spec, stdp=None):
    """Oracle loaded with a nemosim_amd.examples.NetSpec; stdp = (pre, post,
    minE, maxE, minI, maxI) or None. Returns the finalized oracle."""
    o = Oracle()
    for name, idx, args in spec.groups:
        o.add_neurons(name, idx, args)
    o.add_synapses_csr(spec.rowptr, spec.targets, spec.delays, spec.weights, spec.plastic)
    if stdp is not None:
        o.set_stdp_function(*stdp)
    o.finalize()
    return o


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops when called from the
    calling Python thread (ctypes releases the GIL, so several oracles can
    run side by side, one per thread, each with n threads)."""
    load().oracle_set_threads(int(n))


def map_parallel(fn, items, workers=None, threads_each=1, on_done=None):
    """[fn(x) for x in items], run on a pool of host threads, each of whose
    oracles uses threads_each OpenMP threads. For many small networks: one
    network per thread is faster than one at a time across all cores.
    on_done(done_count, total), if given, is called on the calling thread as
    results come in."""
    import concurrent.futures
    items = list(items)
    workers = workers or max(1, min(16, os.cpu_count() or 1) // max(1, threads_each))

    def call(x):
        set_threads(threads_each)
        return fn(x)

    out = [None] * len(items)
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        futs = {pool.submit(call, x): i for i, x in enumerate(items)}
        done = 0
        for f in concurrent.futures.as_completed(futs):
            out[futs[f]] = f.result()
            done += 1
            if on_done is not None:
                on_done(done, len(items))
    return out


def first_divergence(ref_counts, ref_ids, got_counts, got_ids):
    """(cycle, detail) of the first cycle whose fired set differs, or None.
    Rasters are per-cycle fired counts plus the fired ids back to back."""
    ref_counts = np.asarray(ref_counts, np.int64)
    got_counts = np.asarray(got_counts, np.int64)
    ro = np.concatenate([[0], np.cumsum(ref_counts)]).astype(np.int64)
    go = np.concatenate([[0], np.cumsum(got_counts)]).astype(np.int64)
    ref_ids = np.asarray(ref_ids, np.int64)
    got_ids = np.asarray(got_ids, np.int64)
    if len(ref_counts) != len(got_counts):
        return min(len(ref_counts), len(got_counts)), "raster lengths differ: %d and %d cycles" % (
            len(ref_counts), len(got_counts))
    if np.array_equal(ref_counts, got_counts) and np.array_equal(ref_ids, got_ids):
        return None
    for t in range(len(ref_counts)):
        a = ref_ids[ro[t]:ro[t + 1]]
        b = got_ids[go[t]:go[t + 1]]
        if len(a) != len(b) or not np.array_equal(a, b):
            extra = sorted(set(b.tolist()) - set(a.tolist()))[:5]
            missing = sorted(set(a.tolist()) - set(b.tolist()))[:5]
            return t, "ref %d fired, got %d; extra %s missing %s" % (len(a), len(b), extra, missing)
    return None


def fx_toFix(f):
    v = ctypes.c_int32()
    _check(load().oracle_fx_toFix(f, ctypes.byref(v)))
    return v.value


def wfx_toFloat(v):
    return load().oracle_wfx_toFloat(int(v))


def fx_mul(a, b):
    return load().oracle_fx_mul(int(a), int(b))


def urand_seq(state4, n):
    s = np.ascontiguousarray(state4, np.uint32)
    out = np.zeros(n, np.uint32)
    load().oracle_urand_seq(_u(s), n, _u(out))
    return out


def nrand_seq(state4, n):
    s = np.ascontiguousarray(state4, np.uint32)
    out = np.zeros(n, np.float32)
    load().oracle_nrand_seq(_u(s), n, _f(out))
    return out


def libm_logf_grid(first, n):
    out = np.zeros(n, np.float32)
    load().oracle_libm_logf_grid(first, n, _f(out))
    return out


def libm_sinf(a):
    a = np.ascontiguousarray(a, np.float32)
    out = np.zeros_like(a)
    load().oracle_libm_sinf(_f(a), len(a), _f(out))
    return out


def kuramoto_wrap(a):
    """The Kuramoto phase wrap with the host C library's fmodf."""
    a = np.ascontiguousarray(a, np.float32)
    out = np.zeros_like(a)
    load().oracle_kuramoto_wrap(_f(a), len(a), _f(out))
    return out


def init_rng(lo, hi):
    out = np.zeros(4 * (hi - lo + 1), np.uint32)
    load().oracle_init_rng(lo, hi, _u(out))
    return out.reshape(-1, 4)


def mt19937_nth(n):
    return int(load().oracle_mt19937_nth(n))
