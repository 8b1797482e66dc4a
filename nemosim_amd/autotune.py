"""Autotuner for the batched step: neurons per partition (T) per network size.

T is the workgroup's neuron tile -- one thread per neuron, an LDS delay
ring of D x T int32 pairs -- and so also fixes how many workgroups a network
needs (ceil(n/T)), how many fit on a CU at once (LDS) and how much of the
device a batch fills. The fired-list bin of a partition holds up to T
neurons, so it cannot overflow and has no separate size to tune.

The delivery lanes G -- how many lanes of a wavefront share one fired
source while its synapses are read, the bin of the delivery queue -- are
tuned second, at the best T.

``tune`` measures, for each size bucket, the aggregate synaptic events per
second of a batch of same-bucket networks at every candidate, on the
current device, and returns {bucket_upper_bound: (T, G)}.
``partition_size`` and ``delivery_lanes`` look a network size up in such a
table. A measured table is kept in
nemosim_amd/tuning/<arch>.json.
"""
import json
import os
import time

import numpy as np

from . import examples
from . import nemo as nemo_api

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "gfx950.json")


def load_table(path=DEFAULT_TABLE):
    """{bucket upper bound: (partition size T, delivery lanes G)} or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    lanes = t.get("delivery_lanes", {})
    return {int(k): (int(v), int(lanes.get(k, 0))) for k, v in t["partition_size"].items()}


def _lookup(n, table):
    if not table:
        return None
    for bound in sorted(table):
        if n <= bound:
            return table[bound]
    return table[max(table)]


def partition_size(n, table):
    """Tuned T for a network of n neurons (None: let the backend choose)."""
    e = _lookup(n, table)
    return None if e is None else e[0]


def delivery_lanes(n, table):
    """Tuned lanes per fired source for a network of n neurons (0: automatic)."""
    e = _lookup(n, table)
    return 0 if e is None else e[1]


def _batch_rate(sizes, m, T, lanes, steps, seed):
    sims = []
    for i, n in enumerate(sizes):
        spec = examples.izhikevich_80_20(int(n), m, seed=seed + i)
        net = spec.to_nemo(nemo_api)
        conf = nemo_api.Configuration()
        conf.set_write_only_synapses()
        conf.set_partition_size(T)
        conf.set_delivery_lanes(lanes)
        conf.set_step_mode("fused")
        sims.append(nemo_api.Simulation(net, conf))
    nemo_api.run_batch(sims, [50] * len(sims))          # warm-up
    t0 = time.perf_counter()
    ev = nemo_api.run_batch(sims, [steps] * len(sims))
    dt = time.perf_counter() - t0
    return float(np.sum(ev)) / dt, dt / steps


def tune(buckets=((2000, 1000), (5000, 2500), (10000, 7000), (20000, 15000), (50000, 35000),
                  (250000, 200000)),
         m=100, candidates=(64, 128, 256, 512, 768, 960), lane_candidates=(1, 2, 4, 8, 16, 32, 64),
         steps=200, batch_neurons=250000, seed=1, log=print):
    """buckets: (upper bound, representative size). Each bucket is measured as
    a batch of copies of the representative network totalling about
    batch_neurons neurons: first every T with automatic lanes, then every
    lane count at the best T."""
    table, detail = {}, {}
    for bound, n in buckets:
        copies = max(1, batch_neurons // n)
        best = None
        for T in candidates:
            if T > 2 * n:
                continue
            try:
                rate, per_cycle = _batch_rate([n] * copies, m, T, 0, steps, seed)
            except nemo_api.NemoError as e:
                log("  n=%d T=%d: %s" % (n, T, e))
                continue
            log("  n=%d x%d T=%d: %.3g events/s, %.1f us/cycle" % (n, copies, T, rate, per_cycle * 1e6))
            detail.setdefault(str(bound), {})["T=%d" % T] = rate
            if best is None or rate > best[2]:
                best = (T, 0, rate)
        if best is None:
            continue
        for G in lane_candidates:
            rate, per_cycle = _batch_rate([n] * copies, m, best[0], G, steps, seed)
            log("  n=%d x%d T=%d lanes=%d: %.3g events/s, %.1f us/cycle" % (n, copies, best[0], G, rate,
                                                                          per_cycle * 1e6))
            detail[str(bound)]["T=%d,lanes=%d" % (best[0], G)] = rate
            if rate > best[2]:
                best = (best[0], G, rate)
        table[bound] = (best[0], best[1])
    return table, detail


def save_table(table, detail, path=DEFAULT_TABLE, arch="gfx950", m=100):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"arch": arch, "synapses_per_neuron": m,
                   "partition_size": {str(k): v[0] for k, v in sorted(table.items())},
                   "delivery_lanes": {str(k): v[1] for k, v in sorted(table.items())},
                   "measured_events_per_s": detail}, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    t, d = tune()
    save_table(t, d)
    print(json.dumps({str(k): v for k, v in t.items()}))
