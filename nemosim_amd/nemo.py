"""Python host interface of the HIP backend, mirroring NeMo's Python module
(src/api/python/nemo_py.cpp:752-817): ``Network``, ``Configuration`` and
``Simulation`` with the same method names and argument meaning, over the
nemo.h C ABI of libnemo_hip.so (loaded with ctypes).

Errors raise :class:`NemoError` carrying NeMo's status code and message.
"""
import ctypes

import numpy as np

from . import _lib

NEMO_OK = 0
NEMO_API_UNSUPPORTED = 5
NEMO_INVALID_INPUT = 6
NEMO_BUFFER_UNDERFLOW = 8
NEMO_LOGIC_ERROR = 9
NEMO_DL_ERROR = 12

STEP_MODES = {"auto": 0, "fused": 1, "plugins": 2}


def add_plugin_path(path):
    """Prepend a directory to the neuron-model plugin path (nemo_add_plugin_path)."""
    _check(_lib_checked().nemo_add_plugin_path(path.encode()))


class NemoError(RuntimeError):
    def __init__(self, status, message):
        RuntimeError.__init__(self, "%s (status %d)" % (message, status))
        self.status = status


def _lib_checked():
    return _lib.load()


def _check(status):
    if status != NEMO_OK:
        lib = _lib_checked()
        raise NemoError(status, lib.nemo_strerror().decode())


def _fptr(a):
    return a.ctypes.data_as(_lib.c_float_p)


def _uptr(a):
    return a.ctypes.data_as(_lib.c_uint_p)


class Configuration(object):
    def __init__(self):
        self._lib = _lib_checked()
        self._h = self._lib.nemo_new_configuration()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nemo_delete_configuration(self._h)
            self._h = None

    def set_stdp_function(self, prefire, postfire, min_weight=None, max_weight=None,
                          min_e=None, max_e=None, min_i=None, max_i=None):
        """Either (prefire, postfire, min_weight, max_weight) -- NeMo's
        deprecated form, meaning minE=0, maxE=max_weight, minI=0,
        maxI=min_weight (src/nemo/Configuration.cpp:81-88) -- or the explicit
        min_e/max_e/min_i/max_i keywords."""
        pre = np.ascontiguousarray(prefire, dtype=np.float32)
        post = np.ascontiguousarray(postfire, dtype=np.float32)
        if min_e is None:
            min_e, max_e, min_i, max_i = 0.0, max_weight, 0.0, min_weight
        _check(self._lib.nemo_set_stdp_function(self._h, _fptr(pre), len(pre), _fptr(post), len(post),
                                                float(min_e), float(max_e), float(min_i), float(max_i)))

    def set_cuda_backend(self, device=-1):
        _check(self._lib.nemo_set_cuda_backend(self._h, int(device)))

    set_hip_backend = set_cuda_backend

    def set_cpu_backend(self):
        _check(self._lib.nemo_set_cpu_backend(self._h))

    def backend_description(self):
        s = ctypes.c_char_p()
        _check(self._lib.nemo_backend_description(self._h, ctypes.byref(s)))
        return s.value.decode()

    def set_write_only_synapses(self):
        _check(self._lib.nemo_set_write_only_synapses(self._h))

    def set_partition_size(self, neurons):
        _check(self._lib.nemo_set_partition_size(self._h, int(neurons)))

    def set_delivery_lanes(self, lanes):
        """Lanes sharing one fired source in delivery (power of two <= 64; 0 = automatic)."""
        _check(self._lib.nemo_set_delivery_lanes(self._h, int(lanes)))

    def set_stdp_lds_staging(self, enable):
        """STDP count prefixes and reverse records staged in LDS (default on; same results)."""
        _check(self._lib.nemo_set_stdp_lds_staging(self._h, 1 if enable else 0))

    def set_exchange_waves(self, waves):
        """Waves per step-kernel workgroup running the exchange and delay-1
        delivery (< 0 = none; 0 = automatic, the default: 3 of 8 for STDP
        networks of at least 8 waves per workgroup; same results)."""
        _check(self._lib.nemo_set_exchange_waves(self._h, int(waves)))

    def set_device_construction(self, enable):
        """Synapse structures built on the device (default) or on the host; same results."""
        _check(self._lib.nemo_set_device_construction(self._h, 1 if enable else 0))

    def set_step_mode(self, mode):
        """'auto', 'fused' or 'plugins' (see nemo_set_step_mode)."""
        _check(self._lib.nemo_set_step_mode(self._h, STEP_MODES[mode]))

    def enable_logging(self):
        _check(self._lib.nemo_log_stdout(self._h))


class Network(object):
    def __init__(self):
        self._lib = _lib_checked()
        self._h = self._lib.nemo_new_network()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nemo_delete_network(self._h)
            self._h = None

    def add_neuron_type(self, name):
        t = ctypes.c_uint()
        _check(self._lib.nemo_add_neuron_type(self._h, name.encode(), ctypes.byref(t)))
        return t.value

    def add_neuron(self, type, idx, *args):
        """add_neuron(type, idx, param..., state...) for one neuron."""
        a = np.ascontiguousarray(args, dtype=np.float32)
        _check(self._lib.nemo_add_neuron(self._h, int(type), int(idx), len(a), _fptr(a)))

    def add_neurons(self, type, idx, args):
        """Bulk form: idx (n,) and args (n, nargs)."""
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        args = np.ascontiguousarray(args, dtype=np.float32)
        n = len(idx)
        nargs = args.shape[1] if args.ndim == 2 else 0
        _check(self._lib.nemo_add_neurons(self._h, int(type), n, _uptr(idx), nargs, _fptr(args)))

    def add_synapse(self, source, target, delay, weight, plastic):
        sid = ctypes.c_uint64()
        _check(self._lib.nemo_add_synapse(self._h, int(source), int(target), int(delay), float(weight),
                                          int(bool(plastic)), ctypes.byref(sid)))
        return sid.value

    def add_synapses_csr(self, rowptr, targets, delays, weights, plastic=None, sources=None):
        """Bulk form: source k (sources[k], or k) owns synapses rowptr[k]..rowptr[k+1]."""
        rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
        targets = np.ascontiguousarray(targets, dtype=np.uint32)
        delays = np.ascontiguousarray(delays, dtype=np.uint8)
        weights = np.ascontiguousarray(weights, dtype=np.float32)
        pl = None if plastic is None else np.ascontiguousarray(plastic, dtype=np.uint8)
        src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
        _check(self._lib.nemo_add_synapses_csr(
            self._h, len(rowptr) - 1, None if src is None else _uptr(src),
            rowptr.ctypes.data_as(_lib.c_u64_p), _uptr(targets), delays.ctypes.data_as(_lib.c_u8_p),
            _fptr(weights), None if pl is None else pl.ctypes.data_as(_lib.c_u8_p)))

    def neuron_count(self):
        n = ctypes.c_uint()
        _check(self._lib.nemo_neuron_count(self._h, ctypes.byref(n)))
        return n.value

    def set_neuron(self, idx, *args):
        a = np.ascontiguousarray(args, dtype=np.float32)
        _check(self._lib.nemo_set_neuron_n(self._h, int(idx), len(a), _fptr(a)))

    def get_neuron_state(self, idx, var):
        v = ctypes.c_float()
        _check(self._lib.nemo_get_neuron_state_n(self._h, int(idx), int(var), ctypes.byref(v)))
        return v.value

    def get_neuron_parameter(self, idx, p):
        v = ctypes.c_float()
        _check(self._lib.nemo_get_neuron_parameter_n(self._h, int(idx), int(p), ctypes.byref(v)))
        return v.value

    def set_neuron_state(self, idx, var, val):
        _check(self._lib.nemo_set_neuron_state_n(self._h, int(idx), int(var), float(val)))

    def set_neuron_parameter(self, idx, p, val):
        _check(self._lib.nemo_set_neuron_parameter_n(self._h, int(idx), int(p), float(val)))

    def get_synapses_from(self, source):
        ptr = _lib.c_u64_p()
        n = ctypes.c_size_t()
        _check(self._lib.nemo_get_synapses_from_n(self._h, int(source), ctypes.byref(ptr), ctypes.byref(n)))
        return [ptr[i] for i in range(n.value)]

    def _syn(self, fn, ctype, sid):
        v = ctype()
        _check(fn(self._h, int(sid), ctypes.byref(v)))
        return v.value

    def get_synapse_source(self, sid):
        return self._syn(self._lib.nemo_get_synapse_source_n, ctypes.c_uint, sid)

    def get_synapse_target(self, sid):
        return self._syn(self._lib.nemo_get_synapse_target_n, ctypes.c_uint, sid)

    def get_synapse_delay(self, sid):
        return self._syn(self._lib.nemo_get_synapse_delay_n, ctypes.c_uint, sid)

    def get_synapse_weight(self, sid):
        return self._syn(self._lib.nemo_get_synapse_weight_n, ctypes.c_float, sid)

    def get_synapse_plastic(self, sid):
        return self._syn(self._lib.nemo_get_synapse_plastic_n, ctypes.c_ubyte, sid)


class Simulation(object):
    """A simulation on the HIP device (NeMo's nemo.Simulation(net, conf))."""

    def __init__(self, net, conf):
        self._lib = _lib_checked()
        self._h = self._lib.nemo_new_simulation(net._h, conf._h)
        if not self._h:
            raise NemoError(_status_of_last(self._lib), self._lib.nemo_strerror().decode())
        self._last_steps = 0

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nemo_delete_simulation(self._h)
            self._h = None

    def step(self, fstim=None, istim=None):
        """One cycle. fstim: neuron indices forced to fire; istim: list of
        (neuron, current) pairs. Returns the fired neuron indices."""
        fs = np.ascontiguousarray(fstim if fstim is not None else [], dtype=np.uint32)
        if istim:
            ii = np.ascontiguousarray([p[0] for p in istim], dtype=np.uint32)
            iv = np.ascontiguousarray([p[1] for p in istim], dtype=np.float32)
        else:
            ii = np.zeros(0, np.uint32)
            iv = np.zeros(0, np.float32)
        fired = _lib.c_uint_p()
        n = ctypes.c_size_t()
        _check(self._lib.nemo_step(self._h, _uptr(fs), len(fs), _uptr(ii), _fptr(iv), len(ii),
                                   ctypes.byref(fired), ctypes.byref(n)))
        return [fired[i] for i in range(n.value)]

    # reference binding names
    def step_noinput(self):
        return self.step()

    def step_f(self, fstim):
        return self.step(fstim)

    def step_i(self, istim):
        return self.step(None, istim)

    def step_fi(self, fstim, istim):
        return self.step(fstim, istim)

    def apply_stdp(self, reward):
        _check(self._lib.nemo_apply_stdp(self._h, float(reward)))

    def run(self, nsteps, fstim=None, istim=None):
        """nsteps cycles on the device. fstim / istim: optional per-cycle lists
        (length nsteps) as for step(). Returns the delivered synapse count."""
        sched = None
        keep = []
        if fstim is not None or istim is not None:
            sched = _lib.ScheduleT()
            if fstim is not None:
                ptr = np.zeros(nsteps + 1, np.uint32)
                idx = []
                for k in range(nsteps):
                    idx.extend(fstim[k])
                    ptr[k + 1] = len(idx)
                idx = np.ascontiguousarray(idx, dtype=np.uint32)
                keep += [ptr, idx]
                sched.fstim_ptr = _uptr(ptr)
                sched.fstim_idx = _uptr(idx) if len(idx) else _uptr(np.zeros(1, np.uint32))
            if istim is not None:
                ptr = np.zeros(nsteps + 1, np.uint32)
                ii, iv = [], []
                for k in range(nsteps):
                    for (n, c) in istim[k]:
                        ii.append(n)
                        iv.append(c)
                    ptr[k + 1] = len(ii)
                ii = np.ascontiguousarray(ii if ii else [0], dtype=np.uint32)
                iv = np.ascontiguousarray(iv if iv else [0.0], dtype=np.float32)
                keep += [ptr, ii, iv]
                sched.istim_ptr = _uptr(ptr)
                sched.istim_idx = _uptr(ii)
                sched.istim_val = _fptr(iv)
        ev = ctypes.c_ulonglong()
        _check(self._lib.nemo_hip_run(self._h, int(nsteps), ctypes.byref(sched) if sched else None,
                                      ctypes.byref(ev)))
        self._last_steps = nsteps
        return ev.value

    def read_fired(self):
        """(counts per cycle, concatenated fired indices) of the last run."""
        counts = np.zeros(self._last_steps, np.uint32)
        total = ctypes.c_size_t()
        _check(self._lib.nemo_hip_read_fired(self._h, _uptr(counts), None, 0, ctypes.byref(total)))
        ids = np.zeros(max(total.value, 1), np.uint32)
        _check(self._lib.nemo_hip_read_fired(self._h, _uptr(counts), _uptr(ids), len(ids), ctypes.byref(total)))
        return counts, ids[:total.value]

    def read_raster(self):
        wps = ctypes.c_size_t()
        _check(self._lib.nemo_hip_read_raster(self._h, None, 0, ctypes.byref(wps)))
        words = np.zeros(max(self._last_steps * wps.value, 1), np.uint64)
        _check(self._lib.nemo_hip_read_raster(self._h, words.ctypes.data_as(_lib.c_u64_p), len(words),
                                              ctypes.byref(wps)))
        return words[:self._last_steps * wps.value].reshape(self._last_steps, wps.value)

    def info(self):
        i = _lib.InfoT()
        _check(self._lib.nemo_hip_get_info(self._h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _lib.InfoT._fields_}

    def last_kernel_ms(self):
        """Step-kernel time of the last run: HIP events for the simulation's
        own launches; in a batched run, from the first start to the last end
        of its workgroups on the device clock."""
        i = _lib.InfoT()
        _check(self._lib.nemo_hip_get_info(self._h, ctypes.byref(i)))
        return float(i.last_kernel_ms)

    def get_membrane_potential(self, idx):
        v = ctypes.c_float()
        _check(self._lib.nemo_get_membrane_potential(self._h, int(idx), ctypes.byref(v)))
        return v.value

    def get_neuron_state(self, idx, var):
        v = ctypes.c_float()
        _check(self._lib.nemo_get_neuron_state_s(self._h, int(idx), int(var), ctypes.byref(v)))
        return v.value

    def get_neuron_parameter(self, idx, p):
        v = ctypes.c_float()
        _check(self._lib.nemo_get_neuron_parameter_s(self._h, int(idx), int(p), ctypes.byref(v)))
        return v.value

    def set_neuron(self, idx, *args):
        a = np.ascontiguousarray(args, dtype=np.float32)
        _check(self._lib.nemo_set_neuron_s(self._h, int(idx), len(a), _fptr(a)))

    def set_neuron_state(self, idx, var, val):
        _check(self._lib.nemo_set_neuron_state_s(self._h, int(idx), int(var), float(val)))

    def set_neuron_parameter(self, idx, p, val):
        _check(self._lib.nemo_set_neuron_parameter_s(self._h, int(idx), int(p), float(val)))

    def get_synapses_from(self, source):
        ptr = _lib.c_u64_p()
        n = ctypes.c_size_t()
        _check(self._lib.nemo_get_synapses_from_s(self._h, int(source), ctypes.byref(ptr), ctypes.byref(n)))
        return [ptr[i] for i in range(n.value)]

    def _syn(self, fn, ctype, sid):
        v = ctype()
        _check(fn(self._h, int(sid), ctypes.byref(v)))
        return v.value

    def get_synapse_source(self, sid):
        return self._syn(self._lib.nemo_get_synapse_source_s, ctypes.c_uint, sid)

    def get_synapse_target(self, sid):
        return self._syn(self._lib.nemo_get_synapse_target_s, ctypes.c_uint, sid)

    def get_synapse_delay(self, sid):
        return self._syn(self._lib.nemo_get_synapse_delay_s, ctypes.c_uint, sid)

    def get_synapse_weight(self, sid):
        return self._syn(self._lib.nemo_get_synapse_weight_s, ctypes.c_float, sid)

    def get_synapse_plastic(self, sid):
        return self._syn(self._lib.nemo_get_synapse_plastic_s, ctypes.c_ubyte, sid)

    def elapsed_wallclock(self):
        v = ctypes.c_ulong()
        _check(self._lib.nemo_elapsed_wallclock(self._h, ctypes.byref(v)))
        return v.value

    def elapsed_simulation(self):
        v = ctypes.c_ulong()
        _check(self._lib.nemo_elapsed_simulation(self._h, ctypes.byref(v)))
        return v.value

    def reset_timer(self):
        _check(self._lib.nemo_reset_timer(self._h))


def _status_of_last(lib):
    return lib.nemo_hip_last_status()


def run_batch(sims, nsteps):
    """Run nsteps[i] cycles of every simulation in batched persistent launches.
    Returns the delivered synapses of each."""
    lib = _lib_checked()
    n = len(sims)
    handles = (ctypes.c_void_p * n)(*[s._h for s in sims])
    steps = np.ascontiguousarray(nsteps, dtype=np.uint32)
    ev = np.zeros(n, np.uint64)
    _check(lib.nemo_hip_run_batch(handles, n, _uptr(steps), ev.ctypes.data_as(_lib.c_ull_p)))
    for s, k in zip(sims, steps):
        s._last_steps = int(k)
    return ev


def apply_stdp_batch(sims, rewards):
    """apply_stdp(rewards[i]) on every sims[i] under one hold of the device
    lock (nemo_hip_apply_stdp_batch)."""
    lib = _lib_checked()
    n = len(sims)
    if n == 0:
        return
    handles = (ctypes.c_void_p * n)(*[s._h for s in sims])
    r = np.ascontiguousarray(rewards, dtype=np.float32)
    _check(lib.nemo_hip_apply_stdp_batch(handles, n, _fptr(r)))


def device_count():
    lib = _lib_checked()
    c = ctypes.c_uint()
    st = lib.nemo_cuda_device_count(ctypes.byref(c))
    return c.value if st == NEMO_OK else 0


def eval_libm(kind, x, on_device):
    lib = _lib_checked()
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    _check(lib.nemo_hip_eval_libm(int(kind), _fptr(x), _fptr(out), len(x), int(on_device)))
    return out


def nrand_seq(state4, n):
    lib = _lib_checked()
    s = np.ascontiguousarray(state4, dtype=np.uint32)
    out = np.zeros(n, np.float32)
    _check(lib.nemo_hip_nrand_seq(s.ctypes.data_as(_lib.c_u32_p), _fptr(out), n))
    return out
