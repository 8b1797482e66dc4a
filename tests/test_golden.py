"""Regression fixtures (tests/golden/, written by tests/golden/make_golden.py
from the oracle): the oracle must keep producing them, and the HIP backend
must produce the same numbers."""
import os

import numpy as np
import pytest

from backends import first_divergence
from nemosim_amd import examples

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _stdp_run(sim, spec):
    for _ in range(5):
        sim.run(100)
        last = sim.read_fired()
        sim.apply_stdp(1.0)
    return last


def test_oracle_c1_raster(oracle_mod):
    g = _load("c1_raster.npz")
    counts, ids = oracle_mod.from_spec(examples.izhikevich_80_20(1000, 100, seed=1)).run(1000)
    d = first_divergence(g["counts"], g["ids"], counts, ids)
    assert d is None, "oracle vs fixture, first diverging cycle %d: %s" % d


def test_oracle_stdp_weights(oracle_mod):
    g = _load("stdp_weights.npz")
    spec = examples.izhikevich_80_20(2000, 100, seed=13, stdp=True)
    pre, post = examples.stdp_window()
    o = oracle_mod.from_spec(spec, stdp=(pre, post, 0.0, 1.0, 0.0, -1.0))
    for _ in range(5):
        counts, ids = o.run(100)
        o.apply_stdp(1.0)
    w = o.weights_fx(spec.n, spec.nsyn)
    assert np.array_equal(w[g["index"].astype(np.int64)], g["weight_fx"])
    assert first_divergence(g["last_counts"], g["last_ids"], counts, ids) is None


@pytest.mark.gpu
def test_hip_c1_raster(nemo_mod):
    g = _load("c1_raster.npz")
    spec = examples.izhikevich_80_20(1000, 100, seed=1)
    sim = nemo_mod.Simulation(spec.to_nemo(nemo_mod), nemo_mod.Configuration())
    sim.run(1000)
    counts, ids = sim.read_fired()
    d = first_divergence(g["counts"], g["ids"], counts, ids)
    assert d is None, "HIP vs fixture, first diverging cycle %d: %s" % d


@pytest.mark.gpu
def test_hip_stdp_weights(nemo_mod):
    g = _load("stdp_weights.npz")
    spec = examples.izhikevich_80_20(2000, 100, seed=13, stdp=True)
    pre, post = examples.stdp_window()
    conf = nemo_mod.Configuration()
    conf.set_stdp_function(pre, post, min_e=0.0, max_e=1.0, min_i=0.0, max_i=-1.0)
    sim = nemo_mod.Simulation(spec.to_nemo(nemo_mod), conf)
    counts, ids = _stdp_run(sim, spec)
    assert first_divergence(g["last_counts"], g["last_ids"], counts, ids) is None
    for k, wfx in zip(g["index"][:1000], g["weight_fx"][:1000]):
        k = int(k)
        src = int(np.searchsorted(spec.rowptr, k, side="right") - 1)
        sid = (src << 32) | int(k - spec.rowptr[src])
        assert sim.get_synapse_weight(sid) == np.float32(np.float32(wfx) / np.float32(1 << 20)), k
