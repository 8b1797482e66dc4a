"""NeMo's closed-form step tests on the HIP backend, at several partition
sizes (the reference's "different partition size => identical raster"
pattern, src/test/test.cpp:104-132)."""
import numpy as np
import pytest

import scenarios
from backends import HipBuilder

pytestmark = pytest.mark.gpu


@pytest.fixture
def make(nemo_mod):
    return lambda: HipBuilder(nemo_mod)


@pytest.mark.parametrize("ncount,delay,ps", [(1000, 1, None), (1000, 1, 64), (1000, 3, 128), (2000, 20, None)])
def test_ring(make, ncount, delay, ps):
    scenarios.ring(make, ncount, delay, partition_size=ps)


@pytest.mark.parametrize("ps", [None, 64])
def test_double_ring(make, ps):
    scenarios.double_ring(make, partition_size=ps)


def test_firing_stimulus(make):
    scenarios.firing_stimulus(make, cycles=200)


def test_current_stimulus(make):
    scenarios.current_stimulus(make)


def test_invalid_current_stimulus(make, nemo_mod):
    scenarios.invalid_current_stimulus(make, nemo_mod.NemoError)


def test_fixed_point_saturation(make, oracle_mod):
    from backends import OracleBuilder
    got = scenarios.fixed_point_saturation(make)
    ref = scenarios.fixed_point_saturation(lambda: OracleBuilder(oracle_mod))
    assert got == ref


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_integrator_brute_force(make, seed):
    scenarios.integrator_brute_force(make, seed=seed)


def test_integrator_brute_force_saturating(make):
    scenarios.integrator_brute_force(make, seed=11, saturating=True, nsrc=30)


@pytest.mark.parametrize("reward,noise", [(1.0, False), (0.8, False), (0.9, True)])
def test_stdp_closed_form(make, reward, noise):
    scenarios.stdp_closed_form(make, reward=reward, noise=noise)


@pytest.mark.parametrize("delay", [20, 25])
@pytest.mark.parametrize("step_mode", ["fused", "plugins"])
def test_stdp_in_flight(make, delay, step_mode):
    scenarios.stdp_in_flight(make, delay=delay, step_mode=step_mode)


@pytest.mark.parametrize("step_mode", [None, "plugins"])
def test_poisson_and_input(nemo_mod, step_mode):
    from backends import HipBuilder

    class Make(HipBuilder):
        def simulation(self, partition_size=None, step_mode=step_mode):
            return HipBuilder.simulation(self, partition_size, step_mode)
    scenarios.poisson_and_input(lambda: Make(nemo_mod))


def test_no_parameter_neuron_type(nemo_mod):
    """src/test/test.cpp:981-1001 (there with IzhikevichRS, a test model of
    NeMo's CUDA backend): a type without parameters, here Input, simulates;
    its neurons fire exactly when forced."""
    net = nemo_mod.Network()
    t = net.add_neuron_type("Input")
    for n in range(1000):
        net.add_neuron(t, n)
    sim = nemo_mod.Simulation(net, nemo_mod.Configuration())
    for c in range(100):
        forced = [c, 999 - c] if c % 3 == 0 else []
        assert sorted(sim.step(forced)) == sorted(set(forced))


@pytest.mark.parametrize("n0,nstep", [(1, 1), (1000000, 1), (1, 4), (1000000, 4)])
def test_non_contiguous_indices(make, n0, nstep):
    scenarios.non_contiguous_indices(make, n0, nstep)


@pytest.mark.parametrize("contiguous", [True, False])
@pytest.mark.parametrize("stdp", [False, True])
def test_no_outgoing(make, oracle_mod, contiguous, stdp):
    """src/test/networks.cpp:15-78 against the oracle: fired lists of every
    cycle and, with STDP, every weight after applyStdp"""
    from backends import OracleBuilder
    for first, last in ((0, 10), (45, 55), (90, 100)):
        got = scenarios.no_outgoing(make, 100, first, last, contiguous, stdp)
        ref = scenarios.no_outgoing(lambda: OracleBuilder(oracle_mod), 100, first, last, contiguous, stdp)
        assert got == ref, (first, last)


@pytest.mark.parametrize("start,count", [(0, 1), (0, 4), (1, 4)])
def test_simple_network(make, oracle_mod, start, count):
    """Unary network, network without synapses, one-based indices
    (src/test/test.cpp:178-201) against the oracle"""
    from backends import OracleBuilder
    got = scenarios.simple_network(make, start, count)
    ref = scenarios.simple_network(lambda: OracleBuilder(oracle_mod), start, count)
    assert got == ref


def test_set_neuron(nemo_mod):
    """src/test/test.cpp:767-962: the neuron setters and getters of the
    network and of the simulation; invalid neurons and indices throw; a value
    set on the simulation is used from the next step on and then evolves
    (it is written once, not again every step); a membrane potential set
    before the first step is the one the step starts from. With
    src/test/test.cpp:736-750: the membrane potential of a new simulation is
    the initial v."""
    f32 = np.float32
    a, b = f32(0.02), f32(0.2)
    c, d = f32(-65.0) + f32(15.0) * f32(0.25), f32(8.0) - f32(6.0) * f32(0.25)
    v = f32(-65.0)
    u = b * v
    sigma, e = f32(5.0), f32(0.1)
    Err = nemo_mod.NemoError
    net = nemo_mod.Network()
    iz = net.add_neuron_type("Izhikevich")
    with pytest.raises(Err):
        net.set_neuron(0, a, b, c, d, sigma, u, v)
    net.add_neuron(iz, 0, a, b, c - f32(0.1), d, sigma, u, v - f32(1.0))
    for fn, args in ((net.get_neuron_parameter, (1, 0)), (net.get_neuron_state, (1, 0)),
                     (net.get_neuron_parameter, (0, 5)), (net.get_neuron_state, (0, 2))):
        with pytest.raises(Err):
            fn(*args)
    net.set_neuron(0, a - e, b - e, c - e, d - e, sigma - e, u - e, v - e)
    assert [net.get_neuron_parameter(0, k) for k in range(5)] == [x - e for x in (a, b, c, d, sigma)]
    assert [net.get_neuron_state(0, k) for k in range(2)] == [u - e, v - e]
    for k, x in enumerate((a, b, c, d, sigma)):
        net.set_neuron_parameter(0, k, x)
        assert net.get_neuron_parameter(0, k) == x
    for k, x in enumerate((u, v)):
        net.set_neuron_state(0, k, x)
        assert net.get_neuron_state(0, k) == x
    for fn, args in ((net.set_neuron_parameter, (1, 0, 0.0)), (net.set_neuron_state, (1, 0, 0.0)),
                     (net.set_neuron_parameter, (0, 5, 0.0)), (net.set_neuron_state, (0, 2, 0.0))):
        with pytest.raises(Err):
            fn(*args)
    conf = nemo_mod.Configuration()

    sim = nemo_mod.Simulation(net, conf)
    sim.step()
    sim.set_neuron_state(0, 0, u - e)
    assert sim.get_neuron_state(0, 0) == u - e
    sim.set_neuron_state(0, 1, v - e)
    assert sim.get_neuron_state(0, 1) == v - e
    u0, v0 = sim.get_neuron_state(0, 0), sim.get_neuron_state(0, 1)
    for k, x in enumerate((a, b, c, d, sigma)):
        sim.set_neuron_parameter(0, k, x - e)
        assert sim.get_neuron_parameter(0, k) == x - e
    sim.step()
    assert [sim.get_neuron_parameter(0, k) for k in range(5)] == [x - e for x in (a, b, c, d, sigma)]
    assert sim.get_neuron_state(0, 0) != u0
    assert sim.get_neuron_state(0, 1) != v0
    for fn, args in ((sim.set_neuron_parameter, (1, 0, 0.0)), (sim.set_neuron_state, (1, 0, 0.0)),
                     (sim.get_neuron_parameter, (1, 0)), (sim.get_neuron_state, (1, 0)),
                     (sim.set_neuron_parameter, (0, 5, 0.0)), (sim.set_neuron_state, (0, 2, 0.0)),
                     (sim.get_neuron_parameter, (0, 5)), (sim.get_neuron_state, (0, 2))):
        with pytest.raises(Err):
            fn(*args)
    del sim

    sim = nemo_mod.Simulation(net, conf)
    sim.step()
    v0 = sim.get_membrane_potential(0)
    del sim

    sim = nemo_mod.Simulation(net, conf)
    assert sim.get_neuron_state(0, 0) == u
    assert sim.get_neuron_state(0, 1) == v
    assert sim.get_membrane_potential(0) == v
    # c is used only when the neuron fires, which it does not in this step
    sim.set_neuron(0, a, b, c + f32(1.0), d, sigma, u, v)
    sim.step()
    assert sim.get_membrane_potential(0) == v0
    assert [sim.get_neuron_parameter(0, k) for k in range(5)] == [a, b, c + f32(1.0), d, sigma]
    del sim

    net0 = nemo_mod.Network()
    net0.add_neuron(net0.add_neuron_type("Izhikevich"), 0, a, b, c, d, 0.0, u, v)
    sim0 = nemo_mod.Simulation(net0, conf)
    sim0.step()
    sim0.step()
    sim1 = nemo_mod.Simulation(net0, conf)
    sim1.step()
    sim1.set_neuron(0, a, b, c, d, 0.0, u, v)
    sim1.step()
    sim1.step()
    assert sim0.get_membrane_potential(0) == sim1.get_membrane_potential(0)
    del sim0, sim1

    net1 = nemo_mod.Network()
    net1.add_neuron(net1.add_neuron_type("Izhikevich"), 0, a, b, c, d, sigma, u, v - f32(1.0))
    sim = nemo_mod.Simulation(net1, conf)
    sim.set_neuron(0, a, b, c, d, sigma, u, v)
    assert sim.get_membrane_potential(0) == v
    sim.step()
    assert sim.get_membrane_potential(0) == v0


def test_synapse_queries_write_only_and_unconnected(nemo_mod):
    """src/test/test.cpp:605-646: with write-only synapses every synapse
    query throws; a neuron without synapses has an empty list, and a
    neuron that does not exist throws."""
    from backends import HipBuilder
    Err = nemo_mod.NemoError
    b = HipBuilder(nemo_mod)
    scenarios.create_ring(b, 10, plastic=1)
    b.conf.set_write_only_synapses()
    sim = b.simulation()
    sim.step()
    with pytest.raises(Err):
        sim.get_synapses_from(0)
    with pytest.raises(Err):
        sim.get_synapse_weight((1 << 32) | 0)
    b = HipBuilder(nemo_mod)
    for k in range(4):
        b.add_neuron(scenarios.IZ, k, 0.1, 0.2, 0.3, 0.4, 0.7, 0.5, 0.6)
    sim = b.simulation()
    sim.step()
    with pytest.raises(Err):
        sim.get_synapses_from(4)
    assert list(sim.get_synapses_from(3)) == []


@pytest.mark.parametrize("iz_first", [True, False])
def test_neuron_type_mixture(make, oracle_mod, iz_first):
    """src/test/test.cpp:366-412, and every cycle's fired list against the
    oracle's"""
    from backends import OracleBuilder
    got = scenarios.neuron_type_mixture(make, 1024, iz_first)
    assert got == scenarios.neuron_type_mixture(lambda: OracleBuilder(oracle_mod), 1024, iz_first)


def _both(oracle_mod, make, fn, *args):
    from backends import OracleBuilder
    got = fn(make, *args)
    assert got == fn(lambda: OracleBuilder(oracle_mod), *args)


@pytest.mark.parametrize("iz_first", [True, False])
def test_neuron_type_mixture_unused_type(make, oracle_mod, iz_first):
    """src/test/test.cpp:1041-1043: a neuron type added and left without
    neurons has no effect"""
    _both(oracle_mod, make, scenarios.neuron_type_mixture, 0, iz_first)


@pytest.mark.parametrize("duration,other", [(1000, False), (10000, False), (1000, True), (10000, True)])
def test_poisson_rate(make, oracle_mod, duration, other):
    """src/test/PoissonSource.cpp:6-41, every cycle against the oracle"""
    _both(oracle_mod, make, scenarios.poisson_rate, duration, other)


@pytest.mark.parametrize("ncount,iz", [(1, False), (1000, False), (1, True), (1000, True)])
def test_input_neurons(make, ncount, iz):
    """src/test/Input.cpp:6-77"""
    scenarios.input_neurons(make, ncount, iz)


@pytest.mark.parametrize("multiple", [False, True])
def test_if_curr_exp_create(make, oracle_mod, multiple):
    """src/test/IF_curr_exp.cpp:16-49, raster and state against the oracle"""
    _both(oracle_mod, make, scenarios.if_curr_exp_create, 1000, multiple)
