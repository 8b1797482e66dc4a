"""One builder interface over the oracle and the HIP backend, so that every
scenario test runs unchanged against both (as NeMo's TEST_ALL_BACKENDS does,
src/test/test.cpp:33-61)."""
import numpy as np


class OracleBuilder(object):
    name = "oracle"

    def __init__(self, oracle_mod):
        self.o = oracle_mod.Oracle()

    def add_type(self, model):
        """Types need no registration in the oracle"""

    def add_neuron(self, model, idx, *args):
        self.o.add_neuron(model, idx, *args)

    def add_synapse(self, s, t, d, w, plastic=0):
        return self.o.add_synapse(s, t, d, w, plastic)

    def set_stdp(self, pre, post, min_e, max_e, min_i, max_i):
        self.o.set_stdp_function(pre, post, min_e, max_e, min_i, max_i)

    def load_spec(self, spec):
        for name, idx, args in spec.groups:
            self.o.add_neurons(name, idx, args)
        self.o.add_synapses_csr(spec.rowptr, spec.targets, spec.delays, spec.weights, spec.plastic)

    def simulation(self, partition_size=None, step_mode=None):
        self.o.finalize()
        return self.o


class HipBuilder(object):
    name = "hip"

    def __init__(self, nemo_mod):
        self.nemo = nemo_mod
        self.net = nemo_mod.Network()
        self.conf = nemo_mod.Configuration()
        self.types = {}

    def add_type(self, model):
        if model not in self.types:
            self.types[model] = self.net.add_neuron_type(model)
        return self.types[model]

    def add_neuron(self, model, idx, *args):
        self.net.add_neuron(self.add_type(model), idx, *args)

    def add_synapse(self, s, t, d, w, plastic=0):
        return self.net.add_synapse(s, t, d, w, plastic)

    def set_stdp(self, pre, post, min_e, max_e, min_i, max_i):
        self.conf.set_stdp_function(pre, post, min_e=min_e, max_e=max_e, min_i=min_i, max_i=max_i)

    def load_spec(self, spec):
        for name, idx, args in spec.groups:
            if name not in self.types:
                self.types[name] = self.net.add_neuron_type(name)
            self.net.add_neurons(self.types[name], idx, args)
        self.net.add_synapses_csr(spec.rowptr, spec.targets, spec.delays, spec.weights, spec.plastic)

    def simulation(self, partition_size=None, step_mode=None):
        if partition_size:
            self.conf.set_partition_size(partition_size)
        if step_mode:
            self.conf.set_step_mode(step_mode)
        return self.nemo.Simulation(self.net, self.conf)


def first_divergence(ref_counts, ref_ids, got_counts, got_ids):
    """(cycle, detail) of the first cycle whose fired set differs, or None."""
    import oracle
    return oracle.first_divergence(ref_counts, ref_ids, got_counts, got_ids)
