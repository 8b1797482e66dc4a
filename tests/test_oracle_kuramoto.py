"""Pins of the oracle's Kuramoto model (oracle/nemo_oracle.cpp, restating
src/nemo/cpu/plugins/Kuramoto.cpp) by checks that do not reuse its code:

* the reference's own test cases (src/test/Kuramoto.cpp:46-180): history
  initialisation, uncoupled oscillators, one-to-one and n-to-1 coupling,
  each with the reference's tolerance of 0.001 percent;
* a float64 integration of the delayed model (fourth-order Runge-Kutta with
  the source phases held over the four stages; the phases before cycle 0
  from running the free oscillators backwards), which a wrong delay offset,
  a history run in the wrong direction, wrong stage weights or a swapped
  sine argument would leave by far more than single-precision rounding --
  each of those mistakes is planted in a copy of the float64 model and shown
  to fail the same check;
* the bits of one step: the coupling sum in the order of the network's
  synapse iterator (sources ascending, each source's synapses in insertion
  order) and the wrap of a negative phase, whose 2*pi is added in double
  precision; the alternatives (insertion order, a float-only wrap) are shown
  to give other bits on the same inputs.
* the plane bookkeeping of the state history (src/nemo/cpu/Neurons.cpp:
  66-139): getNeuronState reads the current plane, setNeuron writes plane 0,
  and the update after resetTimer reads the plane of timer cycle 0.
"""
import math

import numpy as np
import pytest

import oracle as orc_mod

TWO_PI_F = np.float32(2 * math.pi)
TWO_PI_D = 2 * math.pi


def close_pct(a, b, pct=0.001):
    """BOOST_REQUIRE_CLOSE: relative difference within pct percent of both"""
    a = float(a)
    b = float(b)
    if a == b:
        return True
    d = abs(a - b)
    return d <= abs(a) * pct / 100 and d <= abs(b) * pct / 100


def oscillators(freq, phase, synapses=()):
    o = orc_mod.Oracle()
    for i, (f, p) in enumerate(zip(freq, phase)):
        o.add_neuron("Kuramoto", i, np.float32(f), np.float32(p))
    for s, t, d, w in synapses:
        o.add_synapse(s, t, d, w)
    o.finalize()
    return o


# ---------------------------------------------------- the reference's cases

def test_init():
    o = oscillators([0.1], [0.0])
    assert o.get_neuron_state(0, 0) == 0.0
    o.step()
    assert close_pct(o.get_neuron_state(0, 0), 0.1)


def test_uncoupled():
    n = 1200
    o = oscillators([0.1] * n, [0.0] * n)
    phase = np.float32(0.0)
    for _ in range(1000):
        o.step()
        phase = np.fmod(np.float32(phase + np.float32(0.1)), TWO_PI_F)
        assert close_pct(o.get_neuron_state(n // 2, 0), phase)


def test_one_to_one():
    freq = [np.float32(0.1), np.float32(0.1)]
    phase = [np.float32(0.0), np.float32(1.57)]
    o = oscillators(freq, phase, [(0, 1, 1, 1.0)])
    for _ in range(100):
        o.step()
        f = freq[0]
        k1 = f + math.sin(phase[0] - phase[1])
        k2 = f + math.sin(phase[0] - (phase[1] + 0.5 * k1))
        k3 = f + math.sin(phase[0] - (phase[1] + 0.5 * k2))
        k4 = f + math.sin(phase[0] - (phase[1] + k3))
        phase[1] = np.float32(phase[1] + np.float32((k1 + 2 * k2 + 2 * k3 + k4) / 6))
        phase[0] = np.float32(phase[0] + f)
        for i in range(2):
            phase[i] = np.fmod(phase[i], TWO_PI_F)
            assert close_pct(o.get_neuron_state(i, 0), phase[i])


@pytest.mark.parametrize("ncount,noise", [(2, False), (100, False), (100, True), (256, False),
                                          (257, False), (1000, False)])
def test_n_to_1(ncount, noise):
    f = np.float32(0.1)
    strength = np.float32(1.0 / ncount)
    syn = []
    for n in range(ncount):
        syn.append((n + 1, 0, 1, strength))
        if noise:
            syn += [(n + 1, tgt, 1, 0.0) for tgt in range(1, ncount) if tgt != n + 1]
    o = oscillators([f] * (ncount + 1), [0.0] * (ncount + 1), syn)
    p0 = np.float32(0.0)
    pn = np.float32(0.0)
    for _ in range(50):
        o.step()
        s = float(ncount * strength)
        k1 = f + s * math.sin(pn - p0)
        k2 = f + s * math.sin(pn - (p0 + 0.5 * k1))
        k3 = f + s * math.sin(pn - (p0 + 0.5 * k2))
        k4 = f + s * math.sin(pn - (p0 + k3))
        p0 = np.fmod(np.float32(p0 + np.float32((k1 + 2 * k2 + 2 * k3 + k4) / 6)), TWO_PI_F)
        pn = np.fmod(np.float32(pn + f), TWO_PI_F)
        assert close_pct(o.get_neuron_state(0, 0), p0)
        for n in (1, ncount // 2 + 1, ncount):
            assert close_pct(o.get_neuron_state(n, 0), pn)


# ------------------------------------------------ delayed model in float64

def model_f64(freq, phase, syn, steps, delay_offset=1, backwards=True, weights=(1, 2, 2, 1), swap=False):
    """Phases of cycles 1..steps by a float64 integration; the variations
    plant the mistakes the test must be able to see."""
    freq = np.asarray(freq, np.float64)
    hist = {0: np.asarray(phase, np.float64) % TWO_PI_D}
    for t in range(31):
        prev = hist[-t] - freq if backwards else hist[-t] + freq
        hist[-(t + 1)] = prev % TWO_PI_D
    src = np.array([s for s, _, _, _ in syn])
    tgt = np.array([t for _, t, _, _ in syn])
    dly = np.array([d for _, _, d, _ in syn])
    w = np.array([x for _, _, _, x in syn], np.float64)
    out = []
    for t in range(steps):
        theta = hist[t]
        sp = np.array([hist[t - (d - delay_offset)][s] for s, d in zip(src, dly)])

        def rhs(x):
            arg = (x[tgt] - sp) if swap else (sp - x[tgt])
            return freq + np.bincount(tgt, weights=w * np.sin(arg), minlength=len(freq))

        k0 = rhs(theta)
        k1 = rhs(theta + 0.5 * k0)
        k2 = rhs(theta + 0.5 * k1)
        k3 = rhs(theta + k2)
        a, b, c, d = weights
        hist[t + 1] = (theta + (a * k0 + b * k1 + c * k2 + d * k3) / sum(weights)) % TWO_PI_D
        out.append(hist[t + 1])
    return np.array(out)


def angular(a, b):
    return np.abs((np.asarray(a, np.float64) - b + math.pi) % TWO_PI_D - math.pi)


@pytest.fixture(scope="module")
def delayed_case():
    rng = np.random.default_rng(42)
    n = 60
    freq = rng.uniform(0.0, 0.2, n).astype(np.float32)
    phase = rng.uniform(0.0, 2 * math.pi, n).astype(np.float32)
    syn = [(int(rng.integers(n)), int(t), int(rng.integers(1, 32)), float(np.float32(rng.normal(0, 0.15))))
           for t in range(n) for _ in range(8)]
    o = oscillators(freq, phase, syn)
    steps = 40
    got = []
    for _ in range(steps):
        o.step()
        got.append(o.state_all(0))
    return freq, phase, syn, np.array(got)


def test_delayed_coupling_matches_float64(delayed_case):
    freq, phase, syn, got = delayed_case
    ref = model_f64(freq.astype(np.float64), phase.astype(np.float64), syn, len(got))
    # single-precision rounding, amplified by the coupled dynamics: 1.6e-4
    assert angular(got, ref).max() < 5e-4


@pytest.mark.parametrize("mistake", [dict(delay_offset=0), dict(backwards=False),
                                     dict(weights=(1, 1, 1, 1)), dict(swap=True)])
def test_delayed_coupling_check_sees_mistakes(delayed_case, mistake):
    freq, phase, syn, got = delayed_case
    wrong = model_f64(freq.astype(np.float64), phase.astype(np.float64), syn, len(got), **mistake)
    assert angular(got, wrong).max() > 2e-2


# --------------------------------------------------------- one step, bits

def replay_step(freq, theta, incoming, wrap_double=True):
    """One step of one oscillator in float32 with the host sinf, couplings
    (source phase, weight) summed in the given order."""
    f = np.float32(freq)
    x0 = np.float32(theta)
    sp = np.array([p for p, _ in incoming], np.float32)
    w = np.array([x for _, x in incoming], np.float32)

    def coupled(x):
        s = np.float32(0.0)
        sines = orc_mod.libm_sinf((sp - np.float32(x)).astype(np.float32))
        for k in range(len(w)):
            s = np.float32(s + np.float32(w[k] * sines[k]))
        return s

    k0 = np.float32(f + coupled(x0))
    k1 = np.float32(f + coupled(np.float32(x0 + np.float32(k0 * np.float32(0.5)))))
    k2 = np.float32(f + coupled(np.float32(x0 + np.float32(k1 * np.float32(0.5)))))
    k3 = np.float32(f + coupled(np.float32(x0 + k2)))
    acc = np.float32(np.float32(np.float32(k0 + np.float32(2) * k1) + np.float32(2) * k2) + k3)
    nxt = np.float32(x0 + np.float32(acc / np.float32(6.0)))
    r = np.fmod(nxt, TWO_PI_F)
    if wrap_double:
        return np.float32(np.float64(r) + (TWO_PI_D if nxt < 0 else 0.0))
    return np.float32(r + (TWO_PI_F if nxt < 0 else np.float32(0)))


def bits(x):
    return int(np.float32(x).view(np.uint32))


def test_sum_order_is_the_synapse_iterator_order():
    """Target 0 gets couplings added as sources 7, 3, 12, 3, 7, 5: the sum
    runs over sources 3, 3, 5, 7, 7, 12 (each source's in insertion order)."""
    rng = np.random.default_rng(3)
    for trial in range(200):
        phases = rng.uniform(0, 2 * math.pi, 13).astype(np.float32)
        adds = [(7, 0.9), (3, -0.4), (12, 0.05), (3, 0.61), (7, -0.33), (5, 0.27)]
        adds = [(s, float(np.float32(w * rng.uniform(0.5, 2.0)))) for s, w in adds]
        syn = [(s, 0, 1, w) for s, w in adds]
        o = oscillators([0.1] * 13, phases, syn)
        o.step()
        got = o.get_neuron_state(0, 0)
        order = sorted(range(len(adds)), key=lambda k: adds[k][0])   # stable: insertion within a source
        sorted_in = [(phases[adds[k][0]], adds[k][1]) for k in order]
        insert_in = [(phases[s], w) for s, w in adds]
        a = replay_step(0.1, phases[0], sorted_in)
        b = replay_step(0.1, phases[0], insert_in)
        assert bits(got) == bits(a)
        if bits(a) != bits(b):
            return   # the orders give different bits and the oracle has the iterator's
    pytest.fail("no input separated the two summation orders")


def test_negative_phase_wrap_in_double():
    rng = np.random.default_rng(5)
    for trial in range(2000):
        p0 = np.float32(rng.uniform(0.0, 0.05))
        p1 = np.float32(rng.uniform(1.0, 2.0))
        w = float(np.float32(-rng.uniform(0.5, 3.0)))
        o = oscillators([0.0, 0.0], [p0, p1], [(1, 0, 1, w)])
        o.step()
        got = o.get_neuron_state(0, 0)
        a = replay_step(0.0, p0, [(p1, w)])
        b = replay_step(0.0, p0, [(p1, w)], wrap_double=False)
        assert bits(got) == bits(a)
        if bits(a) != bits(b):
            return
    pytest.fail("no input separated the double and float wraps")


def test_host_wrap_is_the_reference_expression():
    x = np.concatenate([np.linspace(-20, 20, 200001, dtype=np.float32),
                        np.array([-0.0, 0.0, -1e-30, TWO_PI_F, -TWO_PI_F, 1e6, -1e6], np.float32)])
    got = orc_mod.kuramoto_wrap(x)
    r = np.fmod(x, TWO_PI_F)
    # the conditional adds 0.0 otherwise, which turns -0.0 into +0.0
    exp = (r.astype(np.float64) + np.where(x < 0, TWO_PI_D, 0.0)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


# ---------------------------------------------------- state planes

def test_planes_set_state_set_neuron_and_timer_reset():
    f = np.float32(0.1)
    o = oscillators([f, f], [0.0, 1.0])
    for _ in range(5):
        o.step()
    p5 = o.get_neuron_state(0, 0)
    assert close_pct(p5, 0.5)
    # setNeuronState writes the current plane: the next step starts from it
    o.set_neuron_state(0, 0, 2.0)
    o.step()
    assert o.get_neuron_state(0, 0) == np.float32(np.float32(2.0) + np.float32(
        np.float32(np.float32(np.float32(f + np.float32(2) * f) + np.float32(2) * f) + f) / np.float32(6.0)))
    # setNeuron writes plane 0, not the current plane (6): the state read
    # back is unchanged, the parameter is the new one
    o.set_neuron(1, 0.2, 3.0)
    assert o.get_neuron_state(1, 0) != np.float32(3.0)
    assert o.get_neuron_parameter(1, 0) == np.float32(0.2)
    # after resetTimer the update is that of timer cycle 0: it reads plane 0
    # (where setNeuron put 3.0) and writes plane 1
    o.reset_timer()
    o.step()
    assert close_pct(o.get_neuron_state(1, 0), 3.2)


def test_errors():
    o = orc_mod.Oracle()
    o.add_neuron("Kuramoto", 0, 0.1, 0.0)
    o.add_neuron("Izhikevich", 1, 0.02, 0.2, -65.0, 8.0, 0.0, 0.2 * -65.0, -65.0)
    with pytest.raises(orc_mod.OracleError) as e:
        o.finalize()
    assert e.value.status == 5   # NEMO_API_UNSUPPORTED
    o = orc_mod.Oracle()
    o.add_neuron("Kuramoto", 0, 0.1, 0.0)
    o.add_neuron("Kuramoto", 1, 0.1, 0.0)
    o.add_synapse(0, 1, 32, 0.5)
    with pytest.raises(orc_mod.OracleError) as e:
        o.finalize()
    assert e.value.status == 6   # NEMO_INVALID_INPUT
    o = orc_mod.Oracle()
    o.add_neuron("Kuramoto", 0, 0.1, 0.0)
    o.add_neuron("Kuramoto", 1, 0.1, 0.0)
    o.add_synapse(0, 1, 1, 0.5)
    o.set_stdp_function([0.1], [-0.1], 0.0, 1.0, 0.0, -1.0)
    with pytest.raises(orc_mod.OracleError) as e:
        o.finalize()
    assert e.value.status == 5
