"""NeMo's closed-form step tests on the HIP backend, at several partition
sizes (the reference's "different partition size => identical raster"
pattern, src/test/test.cpp:104-132)."""
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
