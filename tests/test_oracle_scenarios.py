"""NeMo's closed-form step tests, run on the oracle (pins the oracle's step
semantics: delivery delays, forced firing, current injection, saturation,
STDP window arithmetic)."""
import pytest

import scenarios
from backends import OracleBuilder


@pytest.fixture
def make(oracle_mod):
    return lambda: OracleBuilder(oracle_mod)


@pytest.mark.parametrize("ncount,delay", [(1000, 1), (1000, 3), (2000, 20)])
def test_ring(make, ncount, delay):
    scenarios.ring(make, ncount, delay)


def test_double_ring(make):
    scenarios.double_ring(make)


def test_firing_stimulus(make):
    scenarios.firing_stimulus(make, cycles=200)


def test_current_stimulus(make):
    scenarios.current_stimulus(make)


def test_invalid_current_stimulus(make, oracle_mod):
    scenarios.invalid_current_stimulus(make, oracle_mod.OracleError)


def test_fixed_point_saturation(make):
    scenarios.fixed_point_saturation(make)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_integrator_brute_force(make, seed):
    scenarios.integrator_brute_force(make, seed=seed)


def test_integrator_brute_force_saturating(make):
    scenarios.integrator_brute_force(make, seed=11, saturating=True, nsrc=30)


@pytest.mark.parametrize("reward,noise", [(1.0, False), (0.8, False), (0.9, True)])
def test_stdp_closed_form(make, reward, noise):
    scenarios.stdp_closed_form(make, reward=reward, noise=noise)


def test_invalid_synapse_target(oracle_mod):
    o = oracle_mod.Oracle()
    o.add_neuron("Izhikevich", 0, 0.02, 0.2, -65, 8, 0, -13, -65)
    o.add_synapse(0, 5, 1, 1.0)
    with pytest.raises(oracle_mod.OracleError) as e:
        o.finalize()
    assert e.value.status == 6


def test_duplicate_neuron(oracle_mod):
    o = oracle_mod.Oracle()
    o.add_neuron("Izhikevich", 0, 0.02, 0.2, -65, 8, 0, -13, -65)
    with pytest.raises(oracle_mod.OracleError):
        o.add_neuron("Izhikevich", 0, 0.02, 0.2, -65, 8, 0, -13, -65)


def test_apply_stdp_without_stdp(oracle_mod):
    o = oracle_mod.Oracle()
    o.add_neuron("Izhikevich", 0, 0.02, 0.2, -65, 8, 0, -13, -65)
    o.finalize()
    with pytest.raises(oracle_mod.OracleError) as e:
        o.apply_stdp(1.0)
    assert e.value.status == 9


@pytest.mark.parametrize("delay", [20, 25])
def test_stdp_in_flight(make, delay):
    scenarios.stdp_in_flight(make, delay=delay)


def test_poisson_and_input(make):
    scenarios.poisson_and_input(make)


@pytest.mark.parametrize("n0,nstep", [(1, 1), (1000000, 1), (1, 4), (1000000, 4)])
def test_non_contiguous_indices(make, n0, nstep):
    scenarios.non_contiguous_indices(make, n0, nstep)


@pytest.mark.parametrize("contiguous", [True, False])
@pytest.mark.parametrize("stdp", [False, True])
def test_no_outgoing(make, contiguous, stdp):
    for first, last in ((0, 10), (45, 55), (90, 100)):
        fired, weights = scenarios.no_outgoing(make, 100, first, last, contiguous, stdp)
        assert len(fired) == (60 if stdp else 20)
        assert len(weights) == (90 * 100 if stdp else 0)
        assert all(w != 2.0 for w in weights)


@pytest.mark.parametrize("start,count", [(0, 1), (0, 4), (1, 4)])
def test_simple_network(make, start, count):
    fired, v = scenarios.simple_network(make, start, count)
    assert len(fired) == 100 and len(v) == count


@pytest.mark.parametrize("iz_first", [True, False])
def test_neuron_type_mixture(make, iz_first):
    scenarios.neuron_type_mixture(make, 1024, iz_first)


@pytest.mark.parametrize("iz_first", [True, False])
def test_neuron_type_mixture_unused_type(make, iz_first):
    scenarios.neuron_type_mixture(make, 0, iz_first)


@pytest.mark.parametrize("duration,other", [(1000, False), (10000, False), (1000, True), (10000, True)])
def test_poisson_rate(make, duration, other):
    scenarios.poisson_rate(make, duration, other)


@pytest.mark.parametrize("ncount,iz", [(1, False), (1000, False), (1, True), (1000, True)])
def test_input_neurons(make, ncount, iz):
    scenarios.input_neurons(make, ncount, iz)


@pytest.mark.parametrize("multiple", [False, True])
def test_if_curr_exp_create(make, multiple):
    fired, state = scenarios.if_curr_exp_create(make, 1000, multiple)
    assert len(fired) == 1000
