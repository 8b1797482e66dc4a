"""The C ABI without a GPU: every function declared in include/*.h is
exported by libnemo_hip.so, and the host-side parts of the API (network
construction, network queries, configuration validation, error codes)
behave like NeMo's."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("nemo.h", "nemo_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(nemo_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(nemo_mod):
    from nemosim_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) > 60
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert missing == []
    # and the Python binding knows every one of them
    assert declared <= set(_lib.SIGNATURES)


def test_version(nemo_mod):
    from nemosim_amd import _lib
    assert _lib.load().nemo_version().decode().startswith("0.7.2")


def test_network_construction_and_queries(nemo_mod):
    net = nemo_mod.Network()
    iz = net.add_neuron_type("Izhikevich")
    assert net.add_neuron_type("Izhikevich") == iz
    for i in range(5):
        net.add_neuron(iz, i, 0.02, 0.2, -65.0, 8.0, 5.0, -13.0, -65.0)
    assert net.neuron_count() == 5
    s0 = net.add_synapse(0, 1, 3, 0.25, True)
    s1 = net.add_synapse(0, 2, 1, -0.5, False)
    s2 = net.add_synapse(4, 0, 7, 0.125, False)
    assert s0 == (0 << 32) | 0 and s1 == (0 << 32) | 1 and s2 == (4 << 32) | 0
    assert net.get_synapses_from(0) == [s0, s1]
    assert net.get_synapse_target(s1) == 2
    assert net.get_synapse_delay(s0) == 3
    assert net.get_synapse_weight(s1) == np.float32(-0.5)
    assert net.get_synapse_plastic(s0) == 1
    assert net.get_synapse_source(s2) == 4
    assert net.get_neuron_parameter(3, 4) == np.float32(5.0)
    assert net.get_neuron_state(3, 1) == np.float32(-65.0)
    net.set_neuron_state(3, 0, -14.0)
    assert net.get_neuron_state(3, 0) == np.float32(-14.0)
    net.set_neuron_parameter(3, 0, 0.1)
    assert net.get_neuron_parameter(3, 0) == np.float32(0.1)


@pytest.mark.parametrize("start,step", [(0, 1), (1000, 1), (0, 5), (1000, 5)])
def test_neuron_count(nemo_mod, start, step):
    """src/test/test.cpp:150-174: neuronCount counts neurons, whatever their
    indices (contiguous or not, from zero or not)."""
    net = nemo_mod.Network()
    iz = net.add_neuron_type("Izhikevich")
    for i in range(2000):
        net.add_neuron(iz, start + i * step, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7)
    assert net.neuron_count() == 2000


def test_invalid_neuron_type(nemo_mod):
    """src/test/test.cpp:966-977: a type id that was never returned by
    add_neuron_type is rejected."""
    net = nemo_mod.Network()
    iz = net.add_neuron_type("Izhikevich")
    for t in range(100):
        if t != iz:
            with pytest.raises(nemo_mod.NemoError):
                net.add_neuron(t, iz, 0, 0, 0, 0, 0, 0, 0)


def test_bulk_and_single_synapse_ids_agree(nemo_mod):
    net = nemo_mod.Network()
    iz = net.add_neuron_type("Izhikevich")
    net.add_neurons(iz, np.arange(4), np.tile(np.array([0.02, 0.2, -65, 8, 0, -13, -65], np.float32), (4, 1)))
    first = net.add_synapse(1, 2, 1, 0.5, False)
    net.add_synapses_csr([0, 0, 2], [3, 0], [2, 5], [0.1, 0.2], sources=[0, 1])
    ids = net.get_synapses_from(1)
    assert ids[0] == first and ids[1:] == [(1 << 32) | 1, (1 << 32) | 2]
    assert net.get_synapse_delay(ids[2]) == 5


def test_errors(nemo_mod):
    net = nemo_mod.Network()
    with pytest.raises(nemo_mod.NemoError) as e:
        net.add_neuron_type("NoSuchModel")
    assert e.value.status == 12           # NEMO_DL_ERROR
    iz = net.add_neuron_type("Izhikevich")
    net.add_neuron(iz, 0, 0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0)
    with pytest.raises(nemo_mod.NemoError) as e:
        net.add_neuron(iz, 0, 0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0)
    assert e.value.status == 6
    with pytest.raises(nemo_mod.NemoError) as e:
        net.add_neuron(iz, 1, 0.02, 0.2)
    assert e.value.status == 6
    with pytest.raises(nemo_mod.NemoError) as e:
        net.add_synapse(0, 0, 0, 1.0, False)   # delay < 1
    assert e.value.status == 6
    with pytest.raises(nemo_mod.NemoError) as e:
        net.get_neuron_state(7, 0)
    assert e.value.status == 6
    conf = nemo_mod.Configuration()
    with pytest.raises(nemo_mod.NemoError) as e:
        conf.set_cpu_backend()
    assert e.value.status == 5


@pytest.mark.parametrize("args", [
    dict(min_e=0, max_e=1, min_i=0, max_i=0.5),     # positive inhibitory limit
    dict(min_e=-1, max_e=1, min_i=0, max_i=-1),     # negative excitatory limit
    dict(min_e=0.5, max_e=0.2, min_i=0, max_i=-1),  # maxE < minE
    dict(min_e=0, max_e=1, min_i=-1, max_i=-0.5),   # |maxI| < |minI|
])
def test_stdp_function_validation(nemo_mod, args):
    """src/nemo/StdpFunction.cpp:12-50 (src/test/stdp.cpp:278-355)"""
    conf = nemo_mod.Configuration()
    with pytest.raises(nemo_mod.NemoError) as e:
        conf.set_stdp_function([0.1] * 5, [-0.1] * 5, **args)
    assert e.value.status == 6


def test_stdp_window_too_long(nemo_mod):
    conf = nemo_mod.Configuration()
    with pytest.raises(nemo_mod.NemoError):
        conf.set_stdp_function([0.1] * 40, [-0.1] * 30, min_e=0, max_e=1, min_i=0, max_i=-1)


def test_delivery_lanes_validation(nemo_mod):
    conf = nemo_mod.Configuration()
    for ok in (0, 1, 8, 64):
        conf.set_delivery_lanes(ok)
    for bad in (3, 12, 128):
        with pytest.raises(nemo_mod.NemoError) as e:
            conf.set_delivery_lanes(bad)
        assert e.value.status == 6   # NEMO_INVALID_INPUT


def test_batch_rejects_a_simulation_listed_twice(nemo_mod):
    """nemo_hip_run_batch refuses a batch naming one simulation twice before
    touching it (the two copies would share state and exchange buffers in
    one launch); checked on the host with a stand-in handle that is never
    dereferenced."""
    from nemosim_amd import _lib
    lib = _lib.load()
    token = ctypes.create_string_buffer(64)
    handles = (ctypes.c_void_p * 2)(ctypes.addressof(token), ctypes.addressof(token))
    steps = (ctypes.c_uint * 2)(10, 10)
    fn = lib.nemo_hip_run_batch
    assert fn(handles, 2, steps, None) == 6
    assert b"more than once" in lib.nemo_strerror()
