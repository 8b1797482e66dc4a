"""The batched and serving workloads at their stated scale, every network
checked against its own oracle run.

  * batched: 256 independent Izhikevich 80/20 networks, N log-uniform in
    [1k, 20k], 100 synapses per neuron, delays U{1..20} ms, distinct seeds,
    2000 cycles each, in batched persistent launches with partition sizes and
    delivery lanes from the tuning table -- the networks and seeds of
    `bench.py --workload batch`;
  * serving: a seeded draw of Poisson requests (50 per second; N
    log-uniform in [1k, 50k]; 100-1000 synapses; 100-2000 ms; 30 % STDP with
    applyStdp(1.0) every 1000 ms; 30 % IF_curr_exp_noise) plus the largest
    STDP case of that distribution, 50k x 1000 for 2000 ms, served under an
    HBM budget that holds only part of them at once, arrivals on the wall
    clock: each request's raster equals its own run alone and its oracle run,
    and every STDP request's weights equal the oracle's.
"""
import numpy as np
import pytest

import oracle
from backends import first_divergence
from nemosim_amd import autotune, examples, serving
from parity import compare_weights


def stdp_tuple():
    pre, post = examples.stdp_window()
    return (pre, post, 0.0, 1.0, 0.0, -1.0)


@pytest.mark.gpu
@pytest.mark.slow
def test_batch_256_networks(nemo_mod, oracle_mod, progress):
    count, steps = 256, 2000
    sizes = examples.log_uniform_sizes(count, 1000, 20000, seed=1)
    specs = [examples.izhikevich_80_20(int(n), 100, seed=2 + i) for i, n in enumerate(sizes)]
    table = autotune.load_table()
    sims = []
    for s in specs:
        net = s.to_nemo(nemo_mod)
        conf = nemo_mod.Configuration()
        conf.set_write_only_synapses()
        T = autotune.partition_size(s.n, table)
        if T:
            conf.set_partition_size(T)
        conf.set_delivery_lanes(autotune.delivery_lanes(s.n, table))
        sims.append(nemo_mod.Simulation(net, conf))
    ev = nemo_mod.run_batch(sims, [steps] * count)
    got = [sim.read_fired() for sim in sims]
    assert all(sim.info()["fused"] == 1 for sim in sims)
    del sims
    progress("batched %d networks (%d neurons), %d synaptic events; checking each against the oracle"
             % (count, int(sizes.sum()), int(np.sum(ev))))

    def check(i):
        orc = oracle.from_spec(specs[i])
        rc, ri = orc.run(steps)
        d = first_divergence(rc, ri, got[i][0], got[i][1])
        return None if d is None else (i, specs[i].n) + d

    def report(done, total):
        if done % 32 == 0 or done == total:
            progress("  %d/%d networks checked" % (done, total))

    bad = [r for r in oracle.map_parallel(check, range(count), threads_each=1, on_done=report) if r is not None]
    assert not bad, "networks diverging from their oracle runs (index, neurons, cycle, detail): %s" % bad[:5]


def oracle_replay(r):
    """The request on the oracle, with applyStdp at the server's boundaries."""
    orc = oracle.from_spec(r.spec(), stdp=stdp_tuple() if r.stdp else None)
    counts, ids = [], []
    done = 0
    while done < r.duration_ms:
        k = r.duration_ms - done
        if r.stdp:
            k = min(k, r.stdp_period - (done % r.stdp_period))
        c, i = orc.run(k)
        counts.append(c)
        ids.append(i)
        done += k
        if r.stdp and done % r.stdp_period == 0:
            orc.apply_stdp(r.reward)
    return orc, np.concatenate(counts), np.concatenate(ids)


@pytest.mark.gpu
@pytest.mark.slow
def test_serving_distribution(oracle_mod, progress):
    reqs = serving.poisson_requests(15, rate_hz=50.0, seed=8)
    kinds = [(r.model, r.stdp) for r in reqs]
    assert ("IF_curr_exp_noise", True) in kinds and ("IF_curr_exp_noise", False) in kinds
    assert ("Izhikevich", True) in kinds and ("Izhikevich", False) in kinds
    reqs.append(serving.Request(15, reqs[-1].arrival_s + 0.02, 50000, 1000, 2000, stdp=True, seed=1015))
    total = sum(r.estimated_device_bytes() for r in reqs)
    budget = max(total // 3, max(r.estimated_device_bytes() for r in reqs))
    srv = serving.Server(hbm_budget_bytes=budget, chunk_cycles=100, build_workers=4, keep_finished=True)
    stats = srv.serve(reqs, realtime=True)
    assert stats["requests"] == len(reqs)
    # admission was staggered: some request waited for budget
    assert any(r.admitted_s > r.arrival_s + 0.05 for r in reqs if r.admitted_s is not None)
    progress("served %d requests: p50 %.2f s, p99 %.2f s; each alone, then each on the oracle"
             % (len(reqs), stats["p50_latency_s"], stats["p99_latency_s"]))
    for r in reqs:
        got_c, got_i = serving.raster_of(r)
        alone_c, alone_i = serving.run_alone(r)
        d = first_divergence(alone_c, alone_i, got_c, got_i)
        assert d is None, "request %d (%s): mix vs alone, cycle %d: %s" % ((r.rid, r.model) + d)

    def check(r):
        orc, rc, ri = oracle_replay(r)
        got_c, got_i = serving.raster_of(r)
        d = first_divergence(rc, ri, got_c, got_i)
        if d is not None:
            return "request %d (%s, %d x %d, stdp %s) vs oracle, cycle %d: %s" % (
                (r.rid, r.model, r.neurons, r.synapses, r.stdp) + d)
        if r.stdp:
            try:
                compare_weights(r.sim, orc, r.spec(), 20000, seed=r.rid)
            except AssertionError as e:
                return "request %d weights differ: %s" % (r.rid, e)
        return None

    bad = [x for x in oracle.map_parallel(check, reqs, workers=8, threads_each=2,
                                          on_done=lambda d, n: progress("  %d/%d requests checked" % (d, n)))
           if x is not None]
    assert not bad, bad
