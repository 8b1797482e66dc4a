"""HIP backend vs the oracle on seeded networks: firing rasters bit-exact per
cycle (first diverging cycle reported), neuron state bit-exact (the FP
arithmetic keeps NeMo's op order; non-finite values compare by class), and
synapse weights bit-exact under STDP."""

import time

import numpy as np
import pytest

import nrand_domain as dom
from backends import first_divergence
from parity import compare_run, compare_state, compare_weights, hip_sim, same_float
from nemosim_amd import examples

pytestmark = pytest.mark.gpu


def test_device_libm_exhaustive(nemo_mod, oracle_mod):
    chunk = 1 << 24
    for kind in (0, 1):
        total = dom.LOG_COUNT if kind == 0 else dom.SIN_COUNT
        bad = 0
        for start in range(0, total, chunk):
            cnt = min(chunk, total - start)
            if kind == 0:
                x = dom.log_inputs(start, cnt)
                ref = oracle_mod.libm_logf_grid(start, cnt)
            else:
                x = dom.sin_inputs(start, cnt)
                ref = oracle_mod.libm_sinf(x)
            got = nemo_mod.eval_libm(kind, x, on_device=True)
            bad += int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
        assert bad == 0, "kind %d: %d inputs differ from the host C library" % (kind, bad)


def test_device_nrand_sequence(nemo_mod, oracle_mod):
    for state in ([1, 2, 3, 4], [0x7fffffff, 0x12345678, 0x0badf00d, 0x3eadbeef], [5489, 0, 0, 1]):
        ref = oracle_mod.nrand_seq(state, 4000)
        got = nemo_mod.nrand_seq(state, 4000)
        assert same_float(got, ref)


@pytest.mark.parametrize("ps", [None, 64, 256])
def test_c1_raster_and_state(nemo_mod, oracle_mod, ps):
    """Izhikevich 80/20, 1000 neurons x 100 synapses, 1000 cycles."""
    spec = examples.izhikevich_80_20(1000, 100, seed=11)
    sim = hip_sim(nemo_mod, spec, partition_size=ps)
    orc = oracle_mod.from_spec(spec)
    rc, _ = compare_run(sim, orc, 1000)
    assert rc.sum() > 1000
    compare_state(sim, orc, 1000, 2)


def test_step_api_equals_run_api(nemo_mod, oracle_mod):
    spec = examples.izhikevich_80_20(1000, 100, seed=5)
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    rng = np.random.default_rng(2)
    for t in range(300):
        fs = sorted(set(rng.integers(0, 1000, size=int(rng.integers(0, 3))).tolist()))
        ist = [(int(i), float(np.float32(rng.uniform(0, 20)))) for i in rng.integers(0, 1000, size=int(rng.integers(0, 3)))]
        a = sim.step(fs, ist)
        b = orc.step(fs, ist)
        assert a == b, t
    # then continue with a device-resident run
    compare_run(sim, orc, 500)
    compare_state(sim, orc, 1000, 2)


def test_run_with_schedule(nemo_mod, oracle_mod):
    spec = examples.izhikevich_80_20(2000, 100, seed=8)
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    rng = np.random.default_rng(4)
    steps = 200
    fst = [sorted(set(rng.integers(0, 2000, size=3).tolist())) for _ in range(steps)]
    ist = [[(int(i), float(np.float32(rng.uniform(0, 10)))) for i in rng.integers(0, 2000, size=4)] for _ in range(steps)]
    sim.run(steps, fstim=fst, istim=ist)
    gc, gi = sim.read_fired()
    rc, ri = [], []
    for t in range(steps):
        f = orc.step(fst[t], ist[t])
        rc.append(len(f))
        ri.extend(f)
    d = first_divergence(np.array(rc), np.array(ri, np.uint32), gc, gi)
    assert d is None, "first diverging cycle %d: %s" % d


def test_if_curr_exp_noise_parity(nemo_mod, oracle_mod):
    spec = examples.if_curr_exp_80_20(3000, 100, seed=3)
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    rc, _ = compare_run(sim, orc, 800)
    assert rc.sum() > 0
    compare_state(sim, orc, 3000, 4)


def test_if_curr_exp_plain_parity(nemo_mod, oracle_mod):
    spec = examples.if_curr_exp_80_20(2000, 100, seed=4, noise=False)
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    rng = np.random.default_rng(1)
    for t in range(300):
        ist = [(int(i), 20.0) for i in rng.integers(0, 2000, size=30)]
        assert sim.step(None, ist) == orc.step(None, ist), t
    compare_state(sim, orc, 2000, 4)


def test_mixed_types(nemo_mod, oracle_mod):
    """Izhikevich and IF_curr_exp_noise neurons in one network, interleaved
    user indices (output in simulation-local order, src/nemo/cpu/Simulation.cpp:199-208)."""
    a = examples.izhikevich_80_20(1500, 60, seed=21)
    b = examples.if_curr_exp_80_20(900, 60, seed=22)
    n = 2400
    rng = np.random.default_rng(9)
    perm = rng.permutation(n).astype(np.uint32)
    groups = [(a.groups[0][0], perm[:1500], a.groups[0][2]), (b.groups[0][0], perm[1500:], b.groups[0][2])]
    m = 60
    rowptr = np.arange(0, n * m + 1, m, dtype=np.uint64)
    targets = rng.integers(0, n, size=n * m, dtype=np.uint32)
    delays = rng.integers(1, 21, size=n * m, dtype=np.uint8)
    weights = np.concatenate([a.weights, b.weights]).astype(np.float32)
    spec = examples.NetSpec(groups, rowptr, targets, delays, weights, np.zeros(n * m, np.uint8))
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    compare_run(sim, orc, 600)


def test_stdp_parity(nemo_mod, oracle_mod):
    """STDP on: excitatory synapses plastic, window A+=0.1, A-=-0.12, tau=20,
    weights clamped to [0,1], applyStdp(1.0) every 100 cycles: weights and
    rasters bit-exact."""
    spec = examples.izhikevich_80_20(2000, 100, seed=13, stdp=True)
    pre, post = examples.stdp_window()
    stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    sim = hip_sim(nemo_mod, spec, stdp=stdp)
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for epoch in range(5):
        compare_run(sim, orc, 100)
        sim.apply_stdp(1.0)
        orc.apply_stdp(1.0)
    ref_w = orc.weights_fx(2000, spec.nsyn)
    sample = np.random.default_rng(0).integers(0, spec.nsyn, size=3000)
    for k in sample:
        src = int(np.searchsorted(spec.rowptr, k, side="right") - 1)
        sid = (src << 32) | int(k - spec.rowptr[src])
        got = sim.get_synapse_weight(sid)
        exp = np.float32(np.float32(ref_w[k]) / np.float32(1 << 20))
        assert got == exp, (k, got, exp)
    assert np.any(ref_w != np.array([int(np.trunc(np.float64(w) * (1 << 20))) for w in spec.weights], np.int64))


@pytest.mark.parametrize("stdp_on", [False, True])
@pytest.mark.parametrize("waves", [-1, 1, 3, 7])
def test_exchange_waves(nemo_mod, oracle_mod, waves, stdp_on):
    """Exchange waves (the exchange and the delay-1 records on `waves` waves
    of each 512-neuron workgroup, the previous cycle's far records on the
    others; -1: none), with and without STDP, in launches of uneven length
    and with applyStdp between them, against the oracle: rasters, state and
    weights."""
    spec = examples.izhikevich_80_20(6000, 300, seed=29, stdp=stdp_on)
    stdp = None
    if stdp_on:
        pre, post = examples.stdp_window()
        stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    net = spec.to_nemo(nemo_mod)
    conf = nemo_mod.Configuration()
    if stdp:
        conf.set_stdp_function(stdp[0], stdp[1], min_e=0.0, max_e=1.0, min_i=0.0, max_i=-1.0)
    conf.set_partition_size(512)
    conf.set_exchange_waves(waves)
    sim = nemo_mod.Simulation(net, conf)
    info = sim.info()
    assert info["deferred"] == 1
    assert info["exchange_waves"] == max(waves, 0), info
    if waves > 0:
        assert info["far_list_sources"] >= 256, info
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for k in [1, 37, 250, 112]:
        compare_run(sim, orc, k)
        if stdp_on:
            sim.apply_stdp(1.0)
            orc.apply_stdp(1.0)
    compare_state(sim, orc, spec.n, 2, sample=range(0, spec.n, 11))
    if stdp_on:
        compare_weights(sim, orc, spec, 5000, seed=3)


def test_automatic_exchange_waves_single_and_batched(nemo_mod, oracle_mod):
    """An STDP network of 512-neuron workgroups gets exchange waves by
    default (3 of 8) in its own launches, and none in batched launches,
    whose LDS it would otherwise raise for every member. Launches of both
    kinds alternate, with applyStdp between some, against the oracle:
    rasters, state and weights."""
    spec = examples.izhikevich_80_20(6000, 300, seed=37, stdp=True)
    pre, post = examples.stdp_window()
    stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    sim = hip_sim(nemo_mod, spec, stdp=stdp, partition_size=512)
    info = sim.info()
    assert info["exchange_waves"] == 3 and info["far_list_sources"] >= 256, info
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for k, batched, apply in [(120, False, True), (90, True, False), (75, False, False), (150, True, True),
                              (60, True, False), (100, False, False)]:
        if batched:
            nemo_mod.run_batch([sim], [k])
            gc, gi = sim.read_fired()
            rc, ri = orc.run(k)
            d = first_divergence(rc, ri, gc, gi)
            assert d is None, "first diverging cycle %d: %s" % d
        else:
            compare_run(sim, orc, k)
        if apply:
            sim.apply_stdp(1.0)
            orc.apply_stdp(1.0)
    compare_state(sim, orc, spec.n, 2, sample=range(0, spec.n, 13))
    compare_weights(sim, orc, spec, 4000, seed=5)


@pytest.mark.parametrize("exchange", [0, 1])
@pytest.mark.parametrize("staging", [True, False])
def test_stdp_work_lists(nemo_mod, oracle_mod, staging, exchange):
    """STDP work lists with and without LDS staging (count-prefix ring and
    reverse records in LDS), in one-cycle launches (each launch rebuilds
    the ring from the counts of the last postLen cycles) and with bursts of
    3000 forced neurons: postLen cycles later each of the 32 partitions has
    about 1200 chunks of 64 synapses to accumulate, more than its list
    holds (1024), so the list is rebuilt after delivery and the excess is
    accumulated inline. With one exchange wave of the two, the other wave
    builds the list and counts itself in LDS in place of a barrier. Rasters
    and weights bit-exact."""
    n = 4000
    spec = examples.izhikevich_80_20(n, 1000, seed=21, stdp=True)
    pre, post = examples.stdp_window()
    net = spec.to_nemo(nemo_mod)
    conf = nemo_mod.Configuration()
    conf.set_stdp_function(pre, post, min_e=0.0, max_e=1.0, min_i=0.0, max_i=-1.0)
    conf.set_partition_size(128)
    conf.set_stdp_lds_staging(staging)
    if exchange:
        conf.set_exchange_waves(exchange)
    sim = nemo_mod.Simulation(net, conf)
    info = sim.info()
    assert info["partitions"] == 32
    assert info["exchange_waves"] == exchange, info
    assert info["stdp_prefix_ring"] == (1 if staging else 0), info
    assert (info["stdp_staged_chunks"] > 0) == staging, info
    orc = oracle_mod.from_spec(spec, stdp=(pre, post, 0.0, 1.0, 0.0, -1.0))
    compare_run(sim, orc, 50)
    rng = np.random.default_rng(8)
    for t in range(60):
        fs = sorted(rng.choice(n, size=3000, replace=False).tolist()) if t % 30 == 0 else []
        assert sim.step(fs) == orc.step(fs), t
    compare_run(sim, orc, 100)
    sim.apply_stdp(1.0)
    orc.apply_stdp(1.0)
    compare_run(sim, orc, 50)
    compare_weights(sim, orc, spec, 3000)


def _construction_pair(nemo_mod, fill, stdp=None, step_mode=None):
    """The same network built with device and with host construction."""
    from backends import HipBuilder
    sims = []
    for dev in (True, False):
        b = HipBuilder(nemo_mod)
        fill(b)
        if stdp is not None:
            b.set_stdp(*stdp)
        b.conf.set_device_construction(dev)
        sims.append(b.simulation(step_mode=step_mode))
    return sims


def _same_simulation(sims, n, cycles, sources, nsyn_check=True):
    dev, host = sims
    idev, ihost = dev.info(), host.info()
    for k in ("partitions", "partition_size", "saturating", "deferred", "fused", "synapses", "plastic_synapses"):
        assert idev[k] == ihost[k], (k, idev, ihost)
    dev.run(cycles)
    host.run(cycles)
    dc, di = dev.read_fired()
    hc, hi = host.read_fired()
    d = first_divergence(hc, hi, dc, di)
    assert d is None, "first diverging cycle %d: %s" % d
    for src in sources:
        ids = dev.get_synapses_from(src)
        assert list(ids) == list(host.get_synapses_from(src))
        for sid in ids:
            assert dev.get_synapse_target(sid) == host.get_synapse_target(sid)
            assert dev.get_synapse_delay(sid) == host.get_synapse_delay(sid)
            assert dev.get_synapse_weight(sid) == host.get_synapse_weight(sid)
            assert dev.get_synapse_plastic(sid) == host.get_synapse_plastic(sid)
    return idev


@pytest.mark.parametrize("kind", ["csr", "stdp", "shuffled", "dense_near", "saturating", "plugin"])
def test_device_construction(nemo_mod, kind):
    """Synapse structures built on the device (atomic bucket cursors, so the
    records of a bucket land in any order) and on the host give the same
    simulation: geometry and mode decisions, rasters, synapse queries and,
    with STDP, weights after applyStdp. Covers networks given as CSR (sources
    in order), synapse by synapse in shuffled order, with buckets of more
    than 15 delay-1 records (no deferred delivery), with saturating sums,
    and on the plugin step (built on the host from the device's mapping)."""
    rng = np.random.default_rng(3)
    stdp = None
    step_mode = None
    if kind in ("csr", "stdp", "dense_near", "plugin"):
        spec = examples.izhikevich_80_20(3000 if kind != "stdp" else 2000, 150, seed=31, stdp=(kind == "stdp"))

        def fill(b):
            b.load_spec(spec)
            if kind == "dense_near":
                for s in range(20):
                    for t in range(40):
                        b.add_synapse(s, t, 1, 0.3, 0)
        n = spec.n
        if kind == "stdp":
            pre, post = examples.stdp_window()
            stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
        if kind == "plugin":
            step_mode = "plugins"
    else:
        spec = examples.izhikevich_80_20(1000, 40, seed=32)
        order = rng.permutation(spec.nsyn)
        src_of = np.searchsorted(spec.rowptr, np.arange(spec.nsyn), side="right") - 1
        scale = 3000.0 if kind == "saturating" else 1.0

        def fill(b):
            for name, idx, args in spec.groups:
                if name not in b.types:
                    b.types[name] = b.net.add_neuron_type(name)
                b.net.add_neurons(b.types[name], idx, args)
            for k in order:
                w = float(spec.weights[k]) * scale
                w = max(-2047.0, min(2047.0, w))
                b.add_synapse(int(src_of[k]), int(spec.targets[k]), int(spec.delays[k]), w, 0)
        n = spec.n
    sims = _construction_pair(nemo_mod, fill, stdp=stdp, step_mode=step_mode)
    info = _same_simulation(sims, n, 300, [0, 1, n // 2, n - 1])
    assert sims[0].info()["device_built"] == (0 if kind == "plugin" else 1)
    assert sims[1].info()["device_built"] == 0
    if kind == "dense_near":
        assert info["deferred"] == 0
    if kind == "saturating":
        assert info["saturating"] == 1
    if kind == "stdp":
        for sim in sims:
            sim.apply_stdp(1.0)
        _same_simulation(sims, n, 100, [0, 5, 999, 1500])


def test_device_construction_invalid(nemo_mod):
    """A synapse to a neuron that does not exist: device construction hands
    the network to the host's checks, which raise NeMo's error."""
    from backends import HipBuilder
    msgs = []
    for dev in (True, False):
        b = HipBuilder(nemo_mod)
        b.load_spec(examples.izhikevich_80_20(500, 20, seed=4))
        b.add_synapse(3, 777, 2, 0.5, 0)
        b.conf.set_device_construction(dev)
        with pytest.raises(nemo_mod.NemoError) as e:
            b.simulation()
        msgs.append(str(e.value))
    assert msgs[0] == msgs[1] and "777" in msgs[0], msgs


@pytest.mark.parametrize("step_mode", [None, "plugins"])
def test_poisson_driven_network(nemo_mod, oracle_mod, step_mode):
    """An Izhikevich 80/20 network driven by 300 PoissonSource neurons and 20
    Input neurons (forced now and then), on the fused kernel (the any-model
    variant) and on the plugin step: rasters and Izhikevich state bit-exact."""
    from backends import HipBuilder, OracleBuilder
    spec = examples.izhikevich_80_20(1000, 100, seed=61)
    rng = np.random.default_rng(62)
    syn = []
    for k in range(320):
        for _ in range(30):
            syn.append((1000 + k, int(rng.integers(0, 1000)), int(rng.integers(1, 21)),
                        float(rng.uniform(0.0, 2.0)), 0))
    builders = [HipBuilder(nemo_mod), OracleBuilder(oracle_mod)]
    for b in builders:
        b.load_spec(spec)
        for k in range(300):
            b.add_neuron("PoissonSource", 1000 + k, 0.02)
        for k in range(20):
            b.add_neuron("Input", 1300 + k)
        for s_, t_, d_, w_, p_ in syn:
            b.add_synapse(s_, t_, d_, w_, p_)
    sim = builders[0].simulation(step_mode=step_mode)
    orc = builders[1].simulation()
    assert sim.info()["fused"] == (0 if step_mode else 1)
    compare_run(sim, orc, 300)
    for t in range(50):
        fs = sorted(rng.choice(20, size=5, replace=False) + 1300) if t % 5 == 0 else []
        fs = [int(x) for x in fs]
        assert sim.step(fs) == orc.step(fs), t
    compare_run(sim, orc, 200)
    compare_state(sim, orc, 1000, 2, sample=range(0, 1000, 3))


def test_synapse_queries(nemo_mod, oracle_mod):
    spec = examples.izhikevich_80_20(300, 20, seed=2)
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    for src in [0, 17, 299]:
        ids = sim.get_synapses_from(src)
        assert len(ids) == 20
        for sid in ids:
            assert sim.get_synapse_target(sid) == orc.get_synapse_target(sid)
            assert sim.get_synapse_delay(sid) == orc.get_synapse_delay(sid)
            assert sim.get_synapse_weight(sid) == orc.get_synapse_weight(sid)
            assert sim.get_synapse_plastic(sid) == orc.get_synapse_plastic(sid)
            assert sim.get_synapse_source(sid) == src


def test_batch_equals_alone(nemo_mod, oracle_mod):
    """Networks of uneven size in one batched persistent launch: each raster
    equals its own oracle run."""
    sizes = [700, 1500, 3100, 1000, 5000]
    specs = [examples.izhikevich_80_20(n, 100, seed=100 + i) for i, n in enumerate(sizes)]
    sims = [hip_sim(nemo_mod, s) for s in specs]
    steps = [400, 300, 400, 250, 400]
    nemo_mod.run_batch(sims, steps)
    for s, sim, k in zip(specs, sims, steps):
        orc = oracle_mod.from_spec(s)
        rc, ri = orc.run(k)
        gc, gi = sim.read_fired()
        d = first_divergence(rc, ri, gc, gi)
        assert d is None, "network %d: first diverging cycle %d: %s" % ((s.n,) + d)


@pytest.mark.parametrize("stdp_on", [False, True])
def test_longest_delays_fused(nemo_mod, oracle_mod, stdp_on):
    """The fused kernel at the longest delays NeMo's CPU backend takes:
    64 (src/nemo/cpu/Simulation.cpp:52-56), and 44 with STDP, whose
    postfire window of 20 must fit beside the delay in the 64-cycle firing
    history. 8000 neurons x 200 synapses, delays uniform up to the limit,
    700 cycles in uneven launches (with applyStdp between them under STDP):
    rasters, state and weights bit-exact."""
    dmax = 44 if stdp_on else 64
    spec = examples.izhikevich_80_20(8000, 200, seed=41, dmax=dmax, stdp=stdp_on)
    stdp = None
    if stdp_on:
        pre, post = examples.stdp_window()
        stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    sim = hip_sim(nemo_mod, spec, stdp=stdp)
    info = sim.info()
    assert info["fused"] == 1 and info["max_delay"] == dmax, info
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for k in [150, 1, 64, 285, 200]:
        compare_run(sim, orc, k)
        if stdp_on:
            sim.apply_stdp(1.0)
            orc.apply_stdp(1.0)
    compare_state(sim, orc, spec.n, 2, sample=range(0, spec.n, 7))
    if stdp_on:
        compare_weights(sim, orc, spec, 4000, seed=9)


def test_batched_kernel_time_per_network(nemo_mod):
    """After a batched run each simulation reports its own step-kernel
    time (its workgroups' first start to last end on the device clock):
    of two like networks in one launch, the one that runs four times the
    cycles takes between 2.5 and 6 times as long, and neither takes longer
    than the launch."""
    specs = [examples.izhikevich_80_20(4000, 100, seed=500 + i) for i in range(2)]
    sims = [hip_sim(nemo_mod, s) for s in specs]
    nemo_mod.run_batch(sims, [1, 1])   # the launch graph is built outside the timed runs
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        nemo_mod.run_batch(sims, [100, 400])
        wall_ms = (time.perf_counter() - t0) * 1e3
        a, b = sims[0].last_kernel_ms(), sims[1].last_kernel_ms()
        assert 0.0 < a < b <= wall_ms, (a, b, wall_ms)
        times.append(b / a)
    ratio = sorted(times)[1]
    assert 2.5 < ratio < 6.0, times


def test_apply_stdp_batch_equals_apply_stdp(nemo_mod, oracle_mod):
    """nemo_hip_apply_stdp_batch gives each simulation what its own
    nemo_apply_stdp gives (the oracle's weights and raster after it), and
    refuses a simulation without STDP."""
    pre, post = examples.stdp_window()
    stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    specs = [examples.izhikevich_80_20(n, 100, seed=300 + i, stdp=True) for i, n in enumerate([1200, 3000, 2000])]
    sims = [hip_sim(nemo_mod, s, stdp=stdp) for s in specs]
    orcs = [oracle_mod.from_spec(s, stdp=stdp) for s in specs]
    rewards = [1.0, 0.5, 0.8]
    for sim in sims:
        sim.run(500)
    nemo_mod.apply_stdp_batch(sims, rewards)
    for sim in sims:
        sim.run(300)
    for s, sim, orc, r in zip(specs, sims, orcs, rewards):
        orc.run(500)
        orc.apply_stdp(r)
        rc, ri = orc.run(300)
        gc, gi = sim.read_fired()
        d = first_divergence(rc, ri, gc, gi)
        assert d is None, "network %d: first diverging cycle %d: %s" % ((s.n,) + d)
        compare_weights(sim, orc, s, 2000, seed=1)
    plain = hip_sim(nemo_mod, examples.izhikevich_80_20(500, 50, seed=1))
    with pytest.raises(nemo_mod.NemoError):
        nemo_mod.apply_stdp_batch([sims[0], plain], [1.0, 1.0])


@pytest.mark.parametrize("dense", [False, True])
def test_deferred_and_immediate_delivery(nemo_mod, oracle_mod, dense):
    """The fused kernel delivers delay-1 records in the cycle and the rest one
    cycle later when every (source, partition) bucket has at most 15 delay-1
    records; a denser bucket switches the network to immediate delivery of
    all delays. Both forms give the oracle's raster and state."""
    from backends import HipBuilder, OracleBuilder
    spec = examples.izhikevich_80_20(1500, 100, seed=77)
    builders = [HipBuilder(nemo_mod), OracleBuilder(oracle_mod)]
    for b in builders:
        b.load_spec(spec)
        if dense:
            for s in range(20):
                for t in range(40):
                    b.add_synapse(s, t, 1, 0.3, 0)
    sim = builders[0].simulation()
    orc = builders[1].simulation()
    info = sim.info()
    assert info["fused"] == 1
    assert info["deferred"] == (0 if dense else 1), info
    compare_run(sim, orc, 400)
    compare_state(sim, orc, spec.n, 2, sample=range(0, spec.n, 7))


@pytest.mark.parametrize("lanes", [1, 2, 16, 64])
def test_delivery_lanes(nemo_mod, oracle_mod, lanes):
    """Every lanes-per-source setting gives the oracle's raster: with 1 or 2
    lanes most buckets are longer than one pass covers (their remainders are
    packed into further passes, more than 63 at a time with 1 lane), with
    64 lanes most lanes of a group are idle."""
    spec = examples.izhikevich_80_20(2500, 300, seed=91)
    net = spec.to_nemo(nemo_mod)
    conf = nemo_mod.Configuration()
    conf.set_partition_size(256)
    conf.set_delivery_lanes(lanes)
    sim = nemo_mod.Simulation(net, conf)
    assert sim.info()["lanes_per_source"] == lanes
    orc = oracle_mod.from_spec(spec)
    compare_run(sim, orc, 200)
    # bursts: 1200 forced sources in one cycle put far more than 63 long
    # buckets into each wave's batch
    rng = np.random.default_rng(5)
    for t in range(40):
        fs = sorted(rng.choice(2500, size=1200, replace=False).tolist()) if t % 10 == 0 else []
        assert sim.step(fs) == orc.step(fs), t
    compare_run(sim, orc, 100)


def test_batch_coresidency_lds_rounding(nemo_mod):
    """A batch whose workgroups fit the device only by the occupancy query's
    count: 12 networks of 100 partitions of 192 neurons (about 32.5 KB of LDS
    each, which the query fits 5 to a CU but the device only 4) must be split
    into launches of workgroups that are all resident, and give each
    network's own raster."""
    specs = [examples.izhikevich_80_20(19200, 50, seed=200 + i) for i in range(12)]
    sims, alone = [], []
    for s in specs:
        for target in (sims, alone):
            conf = nemo_mod.Configuration()
            conf.set_partition_size(192)
            target.append(nemo_mod.Simulation(s.to_nemo(nemo_mod), conf))
    lds = sims[0].info()["lds_bytes"]
    assert sims[0].info()["partitions"] == 100 and 32 * 1024 < lds + 512 <= 33 * 1024, lds
    nemo_mod.run_batch(sims, [200] * len(sims))
    for k, (b, a) in enumerate(zip(sims, alone)):
        a.run(200)
        bc, bi = b.read_fired()
        ac, ai = a.read_fired()
        d = first_divergence(ac, ai, bc, bi)
        assert d is None, "network %d: first diverging cycle %d: %s" % ((k,) + d)


def test_batch_graph_replays(nemo_mod, oracle_mod):
    """The batched launch is an instantiated HIP graph replayed once per
    call: repeated calls on the same pack (same graph), a changed pack (a
    network with 0 cycles drops out, then rejoins) and eager queries between
    replays; after every replay each network's raster and membrane
    potentials equal its own oracle run."""
    sizes = [900, 2100, 1300]
    specs = [examples.izhikevich_80_20(n, 100, seed=300 + i) for i, n in enumerate(sizes)]
    sims = [hip_sim(nemo_mod, s) for s in specs]
    orcs = [oracle_mod.from_spec(s) for s in specs]
    chunks = [[50, 50, 50], [50, 50, 50], [30, 70, 50], [60, 0, 40], [50, 50, 50], [50, 50, 50]]
    for c, ch in enumerate(chunks):
        nemo_mod.run_batch(sims, ch)
        for i, (spec, sim, orc, k) in enumerate(zip(specs, sims, orcs, ch)):
            if k == 0:
                continue
            rc, ri = orc.run(k)
            gc, gi = sim.read_fired()
            d = first_divergence(rc, ri, gc, gi)
            assert d is None, "chunk %d network %d: first diverging cycle %d: %s" % ((c, i) + d)
            for g in (0, spec.n // 2, spec.n - 1):
                assert same_float(sim.get_membrane_potential(g), orc.get_membrane_potential(g)), (c, i, g)


@pytest.mark.parametrize("partition_size", [None, 128])
def test_high_firing(nemo_mod, oracle_mod, partition_size):
    """src/test/test.cpp:697-730: the torus network (one 32 x 32 patch,
    m = 1000, sigma = 16) with every neuron forced to fire every cycle, so
    every partition's fired list is full. Most synapses have delay 1
    (targets within 32 neurons), so buckets hold hundreds of delay-1 records
    and delivery is immediate. 300 forced cycles, then 100 free ones, whose
    rasters and the final state carry the currents still in flight from the
    forced phase: bit-exact against the oracle. Then the network is forced
    for 600 more cycles (NeMo's test runs 1000)."""
    spec = examples.torus(1, 1000, seed=3)
    sim = hip_sim(nemo_mod, spec, partition_size=partition_size)
    info = sim.info()
    assert info["deferred"] == 0, info
    orc = oracle_mod.from_spec(spec)
    everyone = list(range(spec.n))
    for t in range(300):
        got = sim.step(everyone)
        assert got == everyone, t
        assert orc.step(everyone) == everyone, t
    for t in range(100):
        assert sim.step() == orc.step(), t
    compare_state(sim, orc, spec.n, 2)
    for t in range(600):
        assert sim.step(everyone) == everyone, t


@pytest.mark.parametrize("stdp", [False, True])
@pytest.mark.parametrize("network", ["random", "torus"])
def test_mapping_partition_sizes(nemo_mod, oracle_mod, network, stdp):
    """src/test/test.cpp:104-132, 486-512: the random network (1000 x 1000,
    delay 1) and the torus (one patch, m = 1000, sigma 16), with and without
    STDP (applyStdp(1.0) after every 1000 cycles, src/test/utils.cpp:9-46),
    give the same raster at every partition size. Here every partition size
    is compared with the oracle over 2 simulated seconds. With STDP the
    plastic weights may grow to 10, so the sums are bounded only by
    saturation and the saturating variant runs."""
    import scenarios
    if network == "random":
        spec = examples.izhikevich_80_20(1000, 1000, seed=41, dmax=1, stdp=stdp)
    else:
        spec = examples.torus(1, 1000, seed=43, stdp=stdp)
    cfg = scenarios.reference_stdp_config() if stdp else None
    orc = oracle_mod.from_spec(spec, stdp=cfg)
    ref = []
    for s in range(2):
        ref.append(orc.run(1000))
        if stdp:
            orc.apply_stdp(1.0)
    assert sum(int(c.sum()) for c, _ in ref) > 0
    for ps in (None, 512, 256, 128, 64):
        sim = hip_sim(nemo_mod, spec, stdp=cfg, partition_size=ps)
        assert sim.info()["saturating"] == int(stdp), sim.info()
        for s in range(2):
            sim.run(1000)
            gc, gi = sim.read_fired()
            d = first_divergence(ref[s][0], ref[s][1], gc, gi)
            assert d is None, "partition size %s, second %d: first diverging cycle %d: %s" % ((ps, s) + d)
            if stdp:
                sim.apply_stdp(1.0)
        if stdp:
            compare_weights(sim, orc, spec, 2000)


def test_stdp_all_static(nemo_mod, oracle_mod):
    """src/test/test.cpp:661-673: STDP configured but every synapse static
    (1000 neurons x 1000 synapses, delay 1), four epochs of 1000 cycles each
    followed by applyStdp(1.0): rasters bit-exact, weights unchanged."""
    import scenarios
    spec = examples.izhikevich_80_20(1000, 1000, seed=17, dmax=1)
    stdp = scenarios.reference_stdp_config()
    sim = hip_sim(nemo_mod, spec, stdp=stdp)
    assert sim.info()["plastic_synapses"] == 0
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for epoch in range(4):
        compare_run(sim, orc, 1000)
        sim.apply_stdp(1.0)
        orc.apply_stdp(1.0)
    compare_weights(sim, orc, spec, 2000)
    initial = np.array([int(np.trunc(np.float64(w) * (1 << 20))) for w in spec.weights], np.int64)
    assert np.array_equal(orc.weights_fx(spec.n, spec.nsyn), initial)


@pytest.mark.slow
def test_c2_full_network_prefix(nemo_mod, oracle_mod):
    """The 200k x 1000 network (2e8 synapses) for its first 100 cycles: full
    raster bit-exact against the oracle."""
    spec = examples.izhikevich_80_20(200000, 1000, seed=1)
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    compare_run(sim, orc, 100)


@pytest.mark.slow
def test_c3_stdp_epochs(nemo_mod, oracle_mod):
    """The 50k x 1000 STDP network (excitatory synapses plastic, weights in
    [0,1]) for three 1000-cycle epochs with applyStdp(1.0) after each:
    rasters bit-exact every epoch, then a sample of 20000 weights."""
    n = 50000
    spec = examples.izhikevich_80_20(n, 1000, seed=3, stdp=True)
    pre, post = examples.stdp_window()
    stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    sim = hip_sim(nemo_mod, spec, stdp=stdp)
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for epoch in range(3):
        compare_run(sim, orc, 1000)
        sim.apply_stdp(1.0)
        orc.apply_stdp(1.0)
    ref_w = orc.weights_fx(n, spec.nsyn)
    sample = np.random.default_rng(1).integers(0, spec.nsyn, size=20000)
    for k in sample:
        src = int(np.searchsorted(spec.rowptr, k, side="right") - 1)
        sid = (src << 32) | int(k - spec.rowptr[src])
        got = sim.get_synapse_weight(sid)
        exp = np.float32(np.float32(ref_w[k]) / np.float32(1 << 20))
        assert got == exp, (k, got, exp)



@pytest.mark.slow
def test_c2_10000_cycles(nemo_mod, oracle_mod, progress):
    """The 200k x 1000 network for 10000 cycles (1.5e10 synaptic events):
    the full raster bit-exact against the oracle, 1000 cycles at a time,
    then a sample of the neuron state."""
    spec = examples.izhikevich_80_20(200000, 1000, seed=1)
    sim = hip_sim(nemo_mod, spec)
    orc = oracle_mod.from_spec(spec)
    for block in range(10):
        rc, _ = compare_run(sim, orc, 1000)
        progress("C2 cycles %d-%d identical, %d spikes" % (1000 * block, 1000 * block + 999, int(rc.sum())))
    compare_state(sim, orc, spec.n, 2, sample=range(0, spec.n, 97))


@pytest.mark.slow
def test_c3_60_epochs(nemo_mod, oracle_mod, progress):
    """The 50k x 1000 STDP network for 60 simulated seconds with
    applyStdp(1.0) after each: every second's raster bit-exact against the
    oracle, then a sample of 200000 weights."""
    n = 50000
    spec = examples.izhikevich_80_20(n, 1000, seed=3, stdp=True)
    pre, post = examples.stdp_window()
    stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    sim = hip_sim(nemo_mod, spec, stdp=stdp)
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for epoch in range(60):
        rc, _ = compare_run(sim, orc, 1000)
        sim.apply_stdp(1.0)
        orc.apply_stdp(1.0)
        progress("C3 second %d identical, %d spikes" % (epoch, int(rc.sum())))
    compare_weights(sim, orc, spec, 200000, seed=5)


@pytest.mark.slow
@pytest.mark.parametrize("n,m,dmax,stdp_on", [(300000, 100, 20, False), (100000, 100, 64, False),
                                              (150000, 100, 44, True)])
def test_beyond_the_fused_range(nemo_mod, oracle_mod, progress, n, m, dmax, stdp_on):
    """Networks the fused kernel cannot hold -- more partitions than can be
    co-resident (300k neurons at delay 20, 100k at delay 64) or an LDS ring
    too deep for enough partitions (150k at delay 44 with STDP, whose 20-cycle
    postfire window brings the history to the limit of 64) -- run on the
    plugin step in the automatic mode: 200 cycles (with applyStdp after 100)
    against the oracle, full raster and a sample of the state and weights."""
    spec = examples.izhikevich_80_20(n, m, seed=61, dmax=dmax, stdp=stdp_on)
    stdp = None
    if stdp_on:
        pre, post = examples.stdp_window()
        stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    sim = hip_sim(nemo_mod, spec, stdp=stdp)
    info = sim.info()
    assert info["fused"] == 0, info
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    rc, _ = compare_run(sim, orc, 100)
    if stdp_on:
        sim.apply_stdp(1.0)
        orc.apply_stdp(1.0)
    compare_run(sim, orc, 100)
    progress("%d x %d, delays up to %d%s: plugin step, 200 cycles identical (%d spikes in the first 100)"
             % (n, m, dmax, ", STDP" if stdp_on else "", int(rc.sum())))
    compare_state(sim, orc, spec.n, 2, sample=range(0, spec.n, 997))
    if stdp_on:
        compare_weights(sim, orc, spec, 3000, seed=7)
