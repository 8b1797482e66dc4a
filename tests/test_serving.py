"""Serving: continuous batching under an HBM budget gives every request the
same raster as running it alone (and the oracle's), with STDP requests
applying their reward on their own schedule."""
import numpy as np
import pytest

from backends import first_divergence
from nemosim_amd import serving, examples


def test_budget_admission_order_cpu():
    """Admission is FIFO by arrival and never exceeds the budget (host
    logic, no GPU)."""
    reqs = serving.poisson_requests(20, rate_hz=50.0, seed=3, size_range=(1000, 5000), syn_range=(50, 100),
                                    dur_range=(50, 100))
    assert [r.rid for r in sorted(reqs, key=lambda r: r.arrival_s)] == list(range(20))
    for r in reqs:
        assert r.estimated_device_bytes() > 8 * r.neurons * r.synapses
        assert r.construction_bytes() > 30 * r.neurons * r.synapses
    too_big = serving.Request(0, 0.0, 50000, 1000, 10)
    srv = serving.Server(hbm_budget_bytes=10 << 20)
    with pytest.raises(ValueError):
        srv.serve([too_big], realtime=False)


def test_balanced_chunk_plan_cpu():
    """Cycles per chunk by measured cost (host logic, no GPU): unmeasured
    requests run a whole chunk; measured ones run about chunk x median
    cost / own cost, clamped to [chunk/10, 4 chunk], never past their
    duration or an STDP request's applyStdp boundary; without balance every
    request runs a chunk."""
    def req(rid, cost, done=0, dur=10000, stdp=False):
        r = serving.Request(rid, 0.0, 1000, 100, dur, stdp=stdp)
        r.cost_ms, r.cycles_done = cost, done
        return r
    srv = serving.Server(hbm_budget_bytes=1 << 30, chunk_cycles=100)
    active = [req(0, None), req(1, 0.01), req(2, 0.02), req(3, 0.04), req(4, 0.5), req(5, 0.0001),
              req(6, 0.01, done=9950), req(7, 0.01, done=1930, stdp=True)]
    # median of the measured costs: 0.01 ms per cycle, so a budget of 1 ms
    assert srv.plan_steps(active) == [100, 100, 50, 25, 10, 400, 50, 70]
    srv.balance = False
    assert srv.plan_steps(active) == [100, 100, 100, 100, 100, 100, 50, 70]


@pytest.mark.gpu
@pytest.mark.parametrize("device_construction,tuned", [(False, False), (True, False), (False, True)])
def test_alone_equals_mix(oracle_mod, device_construction, tuned):
    """With tuned set the server takes partition sizes and delivery lanes
    from the tuning table while the requests run alone with the backend's
    defaults: the rasters must still be identical."""
    from nemosim_amd import autotune
    reqs = [
        serving.Request(0, 0.00, 1500, 100, 350, seed=31),
        serving.Request(1, 0.00, 4000, 80, 250, model="IF_curr_exp_noise", seed=32),
        serving.Request(2, 0.01, 2500, 100, 1300, stdp=True, seed=33, stdp_period=500),
        serving.Request(3, 0.02, 1200, 60, 420, seed=34),
    ]
    # a budget that holds only some of them at once (resident, plus the
    # construction memory of one being built), so admission is staggered
    budget = (reqs[0].estimated_device_bytes() + reqs[1].estimated_device_bytes() + reqs[2].estimated_device_bytes()
              + max(r.construction_bytes() for r in reqs))
    srv = serving.Server(hbm_budget_bytes=budget, chunk_cycles=64, build_workers=2,
                         device_construction=device_construction,
                         tuning=autotune.load_table() if tuned else None)
    stats = srv.serve(reqs, realtime=False)
    assert stats["requests"] == 4
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0
    for r in reqs:
        got_c, got_i = serving.raster_of(r)
        alone_c, alone_i = serving.run_alone(r)
        d = first_divergence(alone_c, alone_i, got_c, got_i)
        assert d is None, "request %d: mix vs alone, cycle %d: %s" % ((r.rid,) + d)
        if r.model == "Izhikevich" and not r.stdp:
            orc = oracle_mod.from_spec(r.spec())
            rc, ri = orc.run(r.duration_ms)
            d = first_divergence(rc, ri, got_c, got_i)
            assert d is None, "request %d vs oracle, cycle %d: %s" % ((r.rid,) + d)
