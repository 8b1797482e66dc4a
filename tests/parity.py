"""Comparisons of the HIP backend with the oracle shared by the GPU parity
tests: rasters bit-exact per cycle (first diverging cycle reported), neuron
state bit-exact (non-finite values compare by class), synapse weights
bit-exact."""
import numpy as np

from backends import first_divergence


def hip_sim(nemo_mod, spec, stdp=None, partition_size=None):
    net = spec.to_nemo(nemo_mod)
    conf = nemo_mod.Configuration()
    if stdp is not None:
        pre, post, min_e, max_e, min_i, max_i = stdp
        conf.set_stdp_function(pre, post, min_e=min_e, max_e=max_e, min_i=min_i, max_i=max_i)
    if partition_size:
        conf.set_partition_size(partition_size)
    return nemo_mod.Simulation(net, conf)


def same_float(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return np.all((a.view(np.uint32) == b.view(np.uint32)) | both_nan)


def compare_state(sim, orc, n, nstate, sample=None):
    idx = range(n) if sample is None else sample
    for var in range(nstate):
        ref = orc.state_all(var)
        got = np.array([sim.get_neuron_state(i, var) for i in idx], np.float32)
        exp = ref[list(idx)]
        assert same_float(got, exp), "state %d differs at neurons %s" % (
            var, np.nonzero(got.view(np.uint32) != exp.view(np.uint32))[0][:10])


def compare_run(sim, orc, steps):
    sim.run(steps)
    gc, gi = sim.read_fired()
    rc, ri = orc.run(steps)
    d = first_divergence(rc, ri, gc, gi)
    assert d is None, "first diverging cycle %d: %s" % d
    return rc, ri


def compare_weights(sim, orc, spec, nsample, seed=0):
    ref_w = orc.weights_fx(spec.n, spec.nsyn)
    sample = np.random.default_rng(seed).integers(0, spec.nsyn, size=nsample)
    for k in sample:
        src = int(np.searchsorted(spec.rowptr, k, side="right") - 1)
        sid = (src << 32) | int(k - spec.rowptr[src])
        got = sim.get_synapse_weight(sid)
        exp = np.float32(np.float32(ref_w[k]) / np.float32(1 << 20))
        assert got == exp, (k, got, exp)
