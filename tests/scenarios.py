"""Closed-form step scenarios from NeMo's own test suite (src/test/test.cpp,
src/test/stdp.cpp), written once against the builder interface of
backends.py. Each returns nothing and asserts the reference's expectation;
test_oracle_scenarios.py runs them on the oracle, test_gpu_scenarios.py on the
HIP backend."""
import numpy as np

import netutil

IZ = "Izhikevich"
IF = "IF_curr_exp"


def create_ring(b, ncount, n0=0, delay=1, plastic=0, nstep=1):
    """src/test/utils.cpp:125-145: neuron k of the ring has index n0 + k*nstep"""
    for i in range(ncount):
        v, bb, r2 = -65.0, 0.2, 0.25
        src = n0 + i * nstep
        b.add_neuron(IZ, src, 0.02, bb, v + 15.0 * r2, 8.0 - 6.0 * r2, 0.0, np.float32(bb) * np.float32(v), v)
        b.add_synapse(src, n0 + ((i + 1) % ncount) * nstep, delay, 1000.0, plastic)


def ring(make, ncount, delay, partition_size=None):
    """src/test/test.cpp:299-325"""
    b = make()
    create_ring(b, ncount, delay=delay)
    sim = b.simulation(partition_size)
    duration = ncount * 5 // 2
    sim.step([0])
    for ms in range(1, duration):
        fired = sim.step()
        if delay == 1:
            assert fired == [ms % ncount], (ms, fired)
        elif ms % delay == 0:
            assert fired == [(ms // delay) % ncount], (ms, fired)
        else:
            assert fired == [], (ms, fired)


def double_ring(make, partition_size=None):
    """src/test/test.cpp:329-364"""
    ncount = 512
    b = make()
    create_ring(b, ncount, 0, delay=1)
    create_ring(b, ncount, ncount, delay=2)
    sim = b.simulation(partition_size)
    sim.step([0, ncount])
    for ms in range(1, ncount * 5 // 2):
        fired = sim.step()
        if ms % 2 == 0:
            assert fired == [ms % ncount, ncount + (ms // 2) % ncount], (ms, fired)
        else:
            assert fired == [ms % ncount], (ms, fired)


def add_excitatory_neuron(b, idx, sigma=0.0):
    """src/test/utils.cpp:107-113"""
    r = 0.5
    bb = np.float32(0.25) - np.float32(0.05) * np.float32(r)
    v = -65.0
    b.add_neuron(IZ, idx, np.float32(0.02) + np.float32(0.08) * np.float32(r), bb, v, 2.0, sigma,
                 np.float32(bb) * np.float32(v), v)


def firing_stimulus(make, ncount=3000, cycles=300, seed=0):
    """src/test/test.cpp:205-242: the fired set equals the forced set"""
    b = make()
    for n in range(ncount):
        add_excitatory_neuron(b, n)
    sim = b.simulation()
    rng = np.random.default_rng(seed)
    p = 10.0 / ncount
    for t in range(cycles):
        fstim = np.nonzero(rng.random(ncount) < p)[0].tolist()
        fired = sim.step(fstim)
        assert sorted(fired) == fstim, t


def current_stimulus(make, ncount=1500):
    """src/test/test.cpp:245-271: injected current starts a ring"""
    b = make()
    create_ring(b, ncount)
    sim = b.simulation()
    sim.step(None, [(5, 0.001), (8, 0.001), (0, 1000.0), (100, 0.001), (1, 0.001)])
    for ms in range(1, ncount * 2):
        fired = sim.step()
        assert fired == [ms % ncount], (ms, fired)


def invalid_current_stimulus(make, error_type):
    """src/test/test.cpp:274-288"""
    ncount = 1000
    b = make()
    create_ring(b, ncount)
    sim = b.simulation()
    sim.step()
    sim.step()
    try:
        sim.step(None, [(ncount + 1, 0.5)])
    except error_type as e:
        assert e.status == 6
        return
    raise AssertionError("invalid current stimulus did not raise")


def fixed_point_saturation(make):
    """src/test/test.cpp:414-467: a target whose input saturates the Q11.20
    range behaves like the reference target receiving fx_max."""
    INPUT0, INPUT1, OUT_SAT, OUT_REF, OUT_WEAK, NCOUNT = range(6)
    b = make()
    for n in range(NCOUNT):
        b.add_neuron(IZ, n, 0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0)
    b.add_synapse(INPUT0, OUT_REF, 1, 1024.0, 0)
    b.add_synapse(INPUT1, OUT_REF, 1, 1024.0, 0)
    b.add_synapse(INPUT0, OUT_WEAK, 1, 1000.0, 0)
    b.add_synapse(INPUT1, OUT_WEAK, 1, 1000.0, 0)
    b.add_synapse(INPUT0, OUT_SAT, 1, 1100.0, 0)
    b.add_synapse(INPUT1, OUT_SAT, 1, 1100.0, 0)
    sim = b.simulation()
    assert len(sim.step([INPUT0, INPUT1])) == 2
    assert len(sim.step()) == 3
    u_ref = sim.get_neuron_state(OUT_REF, 0)
    u_weak = sim.get_neuron_state(OUT_WEAK, 0)
    u_sat = sim.get_neuron_state(OUT_SAT, 0)
    assert abs(u_ref - u_weak) > 1e-5
    assert abs(u_ref - u_sat) < 1e-5
    return u_ref, u_weak, u_sat


def integrator_brute_force(make, seed=1, nsrc=24, ntgt=8, cycles=60, dmax=7, saturating=False):
    """Delivery checked against a brute-force sum: IF_curr_exp targets that
    integrate their input exactly (tau_syn=1, no leak, no firing) receive
    spikes from sources forced to fire on a random schedule. Each cycle's
    E and I input is the Q11.20 sum, computed here with Python integers, of
    the weights whose (source fired at t-d, delay d) arrive at t; v(t+1) =
    v(t) + ((E+I)+0) in float32."""
    rng = np.random.default_rng(seed)
    b = make()
    for s in range(nsrc):
        netutil.if_source(b, s)
    for t in range(ntgt):
        netutil.if_integrator(b, nsrc + t)
    syn = []
    for s in range(nsrc):
        for _ in range(int(rng.integers(0, 6))):
            t = nsrc + int(rng.integers(0, ntgt))
            d = int(rng.integers(1, dmax + 1))
            if saturating:
                w = float(np.float32(rng.uniform(-2047.0, 2047.0)))
            else:
                w = float(np.float32(rng.uniform(-1.0, 1.0)))
            if rng.random() < 0.05:
                w = 0.0
            syn.append((s, t, d, w))
            b.add_synapse(s, t, d, w, 0)
    sim = b.simulation()
    sched = [np.nonzero(rng.random(nsrc) < 0.3)[0].tolist() for _ in range(cycles)]
    v = {t: np.float32(-65.0) for t in range(nsrc, nsrc + ntgt)}
    for c in range(cycles):
        fired = sim.step(sched[c])
        assert sorted(fired) == sched[c], c
        E = {t: 0 for t in v}
        I = {t: 0 for t in v}
        for (s, t, d, w) in syn:
            if c - d >= 0 and s in sched[c - d]:
                fx = netutil.fix20(w)
                if fx >= 0:
                    E[t] += fx
                else:
                    I[t] += fx
        for t in v:
            cur = np.float32(netutil.sat_to_float(E[t]) + netutil.sat_to_float(I[t]))
            cur = np.float32(cur + np.float32(0.0))
            cur = np.float32(cur + np.float32(0.0))
            v[t] = np.float32(v[t] + np.float32(np.float32(cur / np.float32(1.0)) + np.float32(0.0)))
            got = sim.get_neuron_state(t, 0)
            assert got == v[t], (c, t, got, v[t])


def stdp_closed_form(make, reward=1.0, noise=False, epochs=4):
    """src/test/stdp.cpp:117-265 in miniature: a one-to-one group of
    pre neurons 0..n-1 projecting to post neurons n..2n-1 with delays
    1 + (i % 20); both groups forced to fire dt_fire cycles apart; after each
    epoch + applyStdp the weight equals w0 + epoch*reward*dw(dt) where dw is
    the window value at the arrival/post-firing offset."""
    n = 40
    w0 = 0.5
    pre_fn = [np.float32(1.0) * np.float32(np.exp(-dt / 20.0)) for dt in range(1, 21)]
    post_fn = [np.float32(-0.8) * np.float32(np.exp(-dt / 20.0)) for dt in range(1, 21)]
    b = make()
    b.set_stdp(pre_fn, post_fn, 0.0, 100.0, 0.0, -100.0)
    ids = []
    for i in range(2 * n):
        b.add_neuron(IZ, i, 0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0)
    for i in range(n):
        ids.append(b.add_synapse(i, n + i, 1 + (i % 20), w0, 1))
    if noise:
        rng = np.random.default_rng(7)
        for i in range(n):
            for _ in range(3):
                b.add_synapse(i, n + int(rng.integers(0, n)), 1 + int(rng.integers(0, 20)), 0.0, 0)
    sim = b.simulation()
    dt_fire = 10   # post fires dt_fire cycles after pre
    period = 100
    expect = []
    for e in range(epochs):
        for t in range(period):
            fs = []
            if t == 10:
                fs = list(range(n))
            if t == 10 + dt_fire:
                fs = list(range(n, 2 * n))
            sim.step(fs)
        sim.apply_stdp(reward)
        expect.append([sim.get_synapse_weight(sid) for sid in ids])
    # closed form: arrival at 10 + d; post at 20: dt = 20 - (10 + d) = 10 - d
    fx_reward = netutil.fix20(reward)
    w = [netutil.fix20(w0)] * n
    for e in range(epochs):
        for i in range(n):
            d = 1 + (i % 20)
            arrival = 10 + d
            post = 10 + dt_fire
            if arrival <= post:
                dwf = pre_fn[post - arrival]          # pre-post, dt = 0..: fnPre[dt]
            else:
                dwf = post_fn[arrival - post - 1]     # post-pre, first cycle after: fnPost[0]
            acc = netutil.fix20(dwf)
            delta = (fx_reward * acc) >> 20
            w[i] = min(netutil.fix20(100.0), max(w[i] + delta, 0))
        got = expect[e]
        for i in range(n):
            assert got[i] == np.float32(np.float32(w[i]) / np.float32(1 << 20)), (e, i, got[i], w[i])


def stdp_in_flight(make, delay=20, reward=1.0, **sim_args):
    """A spike in flight across applyStdp arrives with the new weight: NeMo
    reads a synapse's weight when the spike is delivered
    (src/nemo/cpu/Simulation.cpp:258-307), not when the source fires.

    Source S forced at cycle 10 (arrival 10+delay), integrator X forced at
    cycle 35 (so the pre-post pair updates the plastic synapse S->X by
    reward*fnPre[35-(10+delay)]), S forced again at cycle 90, applyStdp after
    cycle 99, and the second spike arrives at cycle 90+delay > 100. X
    integrates exactly (tau_syn=1, no leak, tau_refrac=0), so its potential
    steps by the converted new weight at the arrival cycle."""
    assert 10 + delay <= 35 and 90 + delay >= 100
    S, X = 0, 1
    w0 = 0.5
    pre_fn = [np.float32(1.0) * np.float32(np.exp(-dt / 20.0)) for dt in range(1, 21)]
    post_fn = [np.float32(-0.8) * np.float32(np.exp(-dt / 20.0)) for dt in range(1, 21)]
    b = make()
    b.set_stdp(pre_fn, post_fn, 0.0, 100.0, 0.0, -100.0)
    netutil.if_source(b, S)
    netutil.if_integrator(b, X)
    sid = b.add_synapse(S, X, delay, w0, 1)
    sim = b.simulation(**sim_args)
    arrive = 90 + delay
    v_before = None
    for c in range(arrive + 1):
        fs = []
        if c in (10, 90):
            fs = [S]
        if c == 35:
            fs = [X]
        if c == 100:
            sim.apply_stdp(reward)
            w1 = sim.get_synapse_weight(sid)
        sim.step(fs)
        if c == arrive - 1:
            v_before = sim.get_neuron_state(X, 0)
    # closed form of the updated weight (StdpProcess.cpp:87-113, ConnectivityMatrix.cpp:247-276)
    dwf = pre_fn[35 - (10 + delay)]
    delta = (netutil.fix20(reward) * netutil.fix20(dwf)) >> 20
    w1_fx = min(netutil.fix20(100.0), max(netutil.fix20(w0) + delta, 0))
    assert w1 == np.float32(np.float32(w1_fx) / np.float32(1 << 20)), (w1, w1_fx)
    assert w1_fx != netutil.fix20(w0)
    cur = np.float32(netutil.sat_to_float(w1_fx) + netutil.sat_to_float(0))
    cur = np.float32(np.float32(cur + np.float32(0.0)) + np.float32(0.0))
    expect = np.float32(v_before + np.float32(np.float32(cur / np.float32(1.0)) + np.float32(0.0)))
    got = sim.get_neuron_state(X, 0)
    assert got == expect, (got, expect, v_before)


def poisson_and_input(make, npois=64, ninput=16, cycles=300, seed=5):
    """PoissonSource (src/nemo/cpu/plugins/PoissonSource.cpp:22-29) and Input
    (src/nemo/cpu/plugins/Input.cpp:22-27) in closed form. Neuron k of the
    Poisson group (simulation index k: it is the first type) draws from
    xorshift128 seeded with MT19937 draws 4k..4k+3 shifted right by one
    (src/nemo/RNG.cpp:38-63) and fires when the low 16 bits of its draw are
    below rate * 2^16 -- rate 0 never, rate 1 always -- or when forced.
    Input neurons fire exactly when forced; synapses into either kind (here
    strong ones, from the Poisson sources) change nothing."""
    b = make()
    rng = np.random.default_rng(seed)
    rates = np.concatenate([[0.0, 1.0, 0.5], rng.uniform(0.0, 0.2, npois - 3)]).astype(np.float32)
    for k in range(npois):
        b.add_neuron("PoissonSource", k, float(rates[k]))
    for k in range(ninput):
        b.add_neuron("Input", npois + k)
    for k in range(npois):
        for j in range(4):
            b.add_synapse(k, int(rng.integers(0, npois + ninput)), int(rng.integers(1, 6)), 1000.0, 0)
    sim = b.simulation()
    draws = netutil.mt19937(4 * npois)
    states = [[d >> 1 for d in draws[4 * k:4 * k + 4]] for k in range(npois)]
    thresholds = [int(np.float32(r) * np.float32(65536.0)) for r in rates]
    for t in range(cycles):
        forced = sorted(rng.choice(npois + ninput, size=3, replace=False).tolist()) if t % 7 == 0 else []
        expect = [k for k in range(npois)
                  if (netutil.xorshift128(states[k]) & 0xffff) < thresholds[k] or k in forced]
        expect += [k for k in forced if k >= npois]
        got = sim.step(forced)
        assert sorted(got) == sorted(expect), (t, sorted(got), sorted(expect))


def non_contiguous_indices(make, n0, nstep, ncount=1000, seconds=2):
    """src/test/test.cpp:517-548: a ring whose indices start at n0 and are
    nstep apart fires as the same ring from index 0 does, one neuron per
    cycle, with every index offset by n0."""
    runs = []
    for first in (0, n0):
        b = make()
        create_ring(b, ncount, n0=first, nstep=nstep)
        sim = b.simulation()
        fired = [sim.step([first])]
        fired += [sim.step() for _ in range(seconds * 1000 - 1)]
        assert sum(len(f) for f in fired) == seconds * ncount
        runs.append(fired)
    for t, (f0, f1) in enumerate(zip(*runs)):
        assert [x + n0 for x in f0] == f1, (t, f0, f1)


def reference_stdp_config():
    """The STDP function of NeMo's test configuration (src/test/utils.cpp:84-97):
    pre 0.1 exp(-dt/20), post -0.08 exp(-dt/20) for dt = 1..20, weights in
    [-10, 10] through the four-argument form (minE=0, maxE=10, minI=0,
    maxI=-10, src/nemo/Configuration.cpp:81-88)."""
    dt = np.arange(1, 21, dtype=np.float32)
    pre = (np.float32(0.1) * np.exp(-dt / np.float32(20.0))).astype(np.float32)
    post = (np.float32(-0.08) * np.exp(-dt / np.float32(20.0))).astype(np.float32)
    return pre, post, 0.0, 10.0, 0.0, -10.0


def no_outgoing(make, ncount, first, last, contiguous, stdp):
    """src/test/networks.cpp:15-78: ncount excitatory neurons (indices i, or
    50 + 3i), all-to-all synapses of delay 1 and weight 2 except from the
    neurons [first, last), which have none; neurons with index / 8 == t are
    forced to fire at cycle t, for 20 cycles. NeMo only asks that this runs;
    the fired lists and, with STDP, the weights after applyStdp(1.0) are
    returned for a comparison between backends. With STDP the run is 60
    cycles, so that the postfire window of the first firings has closed and
    every weight has moved."""
    cycles = 60 if stdp else 20
    def index(i):
        return i if contiguous else 50 + 3 * i
    b = make()
    if stdp:
        b.set_stdp(*reference_stdp_config())
    sids = []
    for i in range(ncount):
        add_excitatory_neuron(b, index(i))
        if not (first <= i < last):
            for j in range(ncount):
                sids.append(b.add_synapse(index(i), index(j), 1, 2.0, int(stdp)))
    sim = b.simulation()
    fired = []
    for t in range(cycles):
        fired.append(sim.step([index(i) for i in range(ncount) if index(i) // 8 == t]))
    weights = []
    if stdp:
        sim.apply_stdp(1.0)
        weights = [sim.get_synapse_weight(sid) for sid in sids]
    return fired, weights


def simple_network(make, start, count, cycles=100):
    """src/test/test.cpp:138-148 (unary network, network without synapses,
    one-based indices): count Izhikevich neurons with indices start.. and
    NeMo's placeholder parameters (a..d = 0.1..0.4, u=0.5, v=0.6, sigma=0.7)
    and no synapses step without error. The fired lists and the final
    membrane potentials are returned for a comparison between backends."""
    b = make()
    for k in range(count):
        b.add_neuron(IZ, start + k, 0.1, 0.2, 0.3, 0.4, 0.7, 0.5, 0.6)
    sim = b.simulation()
    fired = [sim.step() for _ in range(cycles)]
    return fired, [sim.get_membrane_potential(start + k) for k in range(count)]


def neuron_type_mixture(make, sz_others, iz_first, sz_ring=1024, cycles=1000):
    """src/test/test.cpp:366-412: a ring of sz_ring Izhikevich neurons and
    sz_others unconnected PoissonSource neurons (rate 0.001, indices from
    sz_ring on), the ring's type added first or second, so that simulation
    indices and neuron indices differ for one of them. Once started, the
    ring fires neuron t mod sz_ring at cycle t, and anything else that fires
    is a Poisson source. The fired lists are returned for a comparison
    between backends."""
    b = make()
    if iz_first:
        create_ring(b, sz_ring)
    b.add_type("PoissonSource")    # with sz_others = 0: a type with no neurons
    for n in range(sz_ring, sz_ring + sz_others):
        b.add_neuron("PoissonSource", n, 0.001)
    if not iz_first:
        create_ring(b, sz_ring)
    sim = b.simulation()
    out = [sim.step([0])]
    for ms in range(1, cycles):
        fired = sorted(sim.step())
        assert fired and fired[0] == ms % sz_ring, (ms, fired[:4])
        assert all(x >= sz_ring for x in fired[1:]), (ms, fired[:4])
        out.append(fired)
    if sz_others:
        assert sum(len(f) for f in out) > cycles   # some Poisson firing
    return out


def poisson_rate(make, duration, other_neurons):
    """src/test/PoissonSource.cpp:6-30: one PoissonSource of rate 0.01
    (index 0), optionally after a silent 1024-neuron ring from index 1, fires
    at least duration * rate times and fewer than three times that. The
    fired lists are returned for a comparison between backends."""
    b = make()
    if other_neurons:
        create_ring(b, 1024, n0=1)
    b.add_neuron("PoissonSource", 0, 0.01)
    sim = b.simulation()
    out = [sim.step() for _ in range(duration)]
    nfired = sum(len(f) for f in out)
    expected = int(abs(np.float32(duration) * np.float32(0.01)))
    assert expected <= nfired < 3 * expected, (nfired, expected)
    return out


def input_neurons(make, ncount, izhikevich, cycles=1000, seed=3):
    """src/test/Input.cpp:6-65: ncount Input neurons (no parameters, no
    state), optionally followed by ncount Izhikevich neurons, fire exactly
    the neurons forced in each cycle (a random fraction p ~ U[0, 0.5) of
    them)."""
    b = make()
    b.add_type("Input")
    for i in range(ncount):
        b.add_neuron("Input", i)
    if izhikevich:
        for i in range(ncount, 2 * ncount):
            b.add_neuron(IZ, i, 0.02, 0.2, -65.0, 8.0, 0.0, np.float32(0.2) * np.float32(-65.0), -65.0)
    sim = b.simulation()
    rng = np.random.default_rng(seed)
    for t in range(cycles):
        p = 0.5 * rng.random()
        fstim = [n for n in range(ncount) if rng.random() < p]
        assert sim.step(fstim) == fstim, t


def if_curr_exp_create(make, duration, multiple):
    """src/test/IF_curr_exp.cpp:16-42: one IF_curr_exp neuron with the
    arguments NeMo's test passes (in PyNN order, so here v_reset=1, c_m=20,
    tau_m=5, ...: the header's order is the one the model reads), optionally
    with a silent 1024-neuron ring from index 1, steps duration cycles. The
    fired lists and the final state are returned for a comparison between
    backends."""
    b = make()
    b.add_neuron(IF, 0, -65.0, 1.0, 20.0, 5.0, 2.0, 5.0, 0.1, -70.0, -51.0, -65.0, 0.0, 0.0, 1000.0)
    if multiple:
        create_ring(b, 1024, n0=1)
    sim = b.simulation()
    out = [sim.step() for _ in range(duration)]
    return out, [sim.get_neuron_state(0, k) for k in range(4)]
