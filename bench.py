#!/usr/bin/env python3
"""Benchmark of the HIP simulation step (one GPU).

Workload (default): the Izhikevich 80/20 network at 200,000 neurons x 1000
synapses (2e8 synapses), delays U{1..20} ms, STDP off, synthetic seeded
construction, run with the plain single-network persistent step kernel.
A "step" is one 1 ms simulation cycle of the whole network.

Metric: synaptic events per wall second (exact delivered synapses), plus
ms of wall time per simulated second. Timing follows NeMo's benchmark()
(src/examples/common.cpp:13-122): warm-up cycles first, then the timed
cycles; the network and its state are resident on the GPU before timing.

Prints one JSON line with the roofline of the step kernel (algorithmic
bytes per launch / kernel time from HIP events on the kernel's stream) and a
CPU baseline: the oracle (a restatement of NeMo's OpenMP CPU backend) timed
on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def delay_tables(spec):
    """Per source: out-degree and number of distinct delays (rows)."""
    n = spec.n
    deg = np.diff(spec.rowptr).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    key = np.unique(src * 64 + spec.delays.astype(np.int64))
    rows = np.bincount(key // 64, minlength=n).astype(np.int64)
    return deg, rows


def algorithmic_bytes(fired_ids, deg, rows, n_neurons, steps):
    """Bytes the step touches per NeMo's data layout: 8 B per delivered
    synapse record {target, Q11.20 weight}, 8 B per delivered (source, delay)
    row descriptor, 68 B per Izhikevich neuron and cycle (a,b,c,d,sigma,u,v,
    RNG read; u,v,RNG written)."""
    e = int(deg[fired_ids].sum())
    r = int(rows[fired_ids].sum())
    return 8 * e + 8 * r + 68 * n_neurons * steps, e, r


def cpu_baseline(spec, target_s, warm_steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    log("cpu baseline: building the oracle network ...")
    orc = oracle.from_spec(spec)
    orc.run(warm_steps)
    deg, _ = delay_tables(spec)
    # probe the step time, then size the sample to ~target_s of CPU work
    t0 = time.perf_counter()
    c, ids = orc.run(10)
    probe = (time.perf_counter() - t0) / 10
    nsamp = int(max(20, min(20000, target_s / max(probe, 1e-6))))
    log("cpu baseline: %.2f ms/cycle probe; sampling %d cycles" % (probe * 1e3, nsamp))
    t0 = time.perf_counter()
    counts, ids = orc.run(nsamp)
    wall = time.perf_counter() - t0
    events = int(deg[ids].sum())
    return {
        "value": events / wall,
        "unit": "synaptic events/s",
        "cores": int(oracle.load().oracle_num_threads()),
        "kind": "port",
        "sample": "same %d-neuron network, cycles %d..%d after %d warm-up cycles (%.1f s, %d events); "
                  "OpenMP restatement of NeMo's CPU backend" % (spec.n, warm_steps + 10, warm_steps + 10 + nsamp,
                                                              warm_steps, wall, events),
        "ms_per_sim_second": wall * 1e3 / (nsamp / 1000.0),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000, help="timed cycles")
    ap.add_argument("--warmup", type=int, default=1000, help="warm-up cycles")
    ap.add_argument("--launch-steps", type=int, default=1000, help="cycles per persistent launch")
    ap.add_argument("--neurons", type=int, default=200000)
    ap.add_argument("--synapses", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(one network per GPU, see DESIGN.md for how it would shard)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import nemosim_amd
    from nemosim_amd import examples

    n, m = args.neurons, args.synapses
    log("generating Izhikevich 80/20 network: %d neurons x %d synapses ..." % (n, m))
    t0 = time.perf_counter()
    spec = examples.izhikevich_80_20(n, m, seed=args.seed)
    log("  %.1f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    net = spec.to_nemo(nemosim_amd)
    conf = nemosim_amd.Configuration()
    conf.set_cuda_backend(0)
    conf.set_write_only_synapses()
    sim = nemosim_amd.Simulation(net, conf)
    del net
    info = sim.info()
    log("  simulation built in %.1f s: %s" % (time.perf_counter() - t0, info))

    L = max(1, min(args.launch_steps, args.steps))
    deg, rows = delay_tables(spec)

    log("warm-up: %d cycles" % args.warmup)
    done = 0
    while done < args.warmup:
        k = min(L, args.warmup - done)
        sim.run(k)
        done += k

    log("timed: %d cycles in launches of %d" % (args.steps, L))
    events = 0
    kernel_ms = 0.0
    fired = []
    launches = 0
    wall = 0.0
    done = 0
    while done < args.steps:
        k = min(L, args.steps - done)
        t0 = time.perf_counter()
        events += sim.run(k)
        wall += time.perf_counter() - t0
        kernel_ms += sim.info()["last_kernel_ms"]
        launches += 1
        _, ids = sim.read_fired()      # outside the timed interval
        fired.append(ids)
        done += k
    fired = np.concatenate(fired) if fired else np.zeros(0, np.uint32)
    abytes, e_check, r_total = algorithmic_bytes(fired.astype(np.int64), deg, rows, n, args.steps)
    if e_check != events:
        log("warning: device event count %d != raster-derived %d" % (events, e_check))
    kernel_s = kernel_ms / 1e3
    achieved = abytes / launches / (kernel_s / launches) / 1e9

    result = {
        "metric": "synaptic events/s",
        "value": events / wall,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "ms_per_sim_second": wall * 1e3 / (args.steps / 1000.0),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "q11.20-i32/f32/f64",
        "data": "synthetic (seeded Izhikevich 80/20 construction)",
        "config": {
            "workload": "izhikevich_80_20 plain step, STDP off",
            "neurons": n,
            "synapses_per_neuron": m,
            "delays": "U{1..20} ms",
            "launch_cycles": L,
            "partitions": info["partitions"],
            "partition_size": info["partition_size"],
            "firing_rate_hz": float(len(fired)) / n / (args.steps / 1000.0),
            "synaptic_events": events,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "nemo_step_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": abytes / launches,
            "kernel_ms_per_launch": kernel_ms / launches,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(spec, args.cpu_seconds, warm_steps=200)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
