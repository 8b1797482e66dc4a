#!/usr/bin/env python3
"""Benchmark of the HIP simulation step (one GPU).

Default workload ("plain"): the Izhikevich 80/20 network at 200,000 neurons
x 1000 synapses (2e8 synapses), delays U{1..20} ms, STDP off, synthetic
seeded construction, run with the plain single-network persistent step
kernel. A "step" is one 1 ms simulation cycle of the whole network.

Metric: synaptic events per wall second (exact delivered synapses), plus
ms of wall time per simulated second. Timing follows NeMo's benchmark()
(src/examples/common.cpp:13-122): warm-up cycles first, then the timed
cycles; the network and its state are resident on the GPU before timing.

The default prints one JSON line with the roofline of the step kernel
(algorithmic bytes per launch / kernel time from HIP events on the kernel's
stream) and a CPU baseline: the oracle (a restatement of NeMo's OpenMP CPU
backend) timed on a bounded sample of the same workload on this host.

Other workloads (--workload): "stdp" (50k x 1000, STDP, applyStdp every
1000 cycles), "batch" (256 networks of 1k-20k neurons, batched persistent
launch vs one plain run per network), "serve" (Poisson requests under an
HBM budget, p50/p99 latency), "step" (NeMo's per-cycle nemo_step at
1k x 100 and 200k x 1000), "kuramoto" (Kuramoto oscillators through the
plugin; give --neurons and --synapses, e.g. 50000 and 100).
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


def oracle_leg(spec, run, target_s, parity_s, warm_steps=200, stdp=None, baseline=True):
    """The oracle (a restatement of NeMo's CPU backend) on the same network:

    parity: replays the GPU run's cycles from cycle 0 -- in the same launch
    chunks, with applyStdp after the same ones -- for as long as parity_s of
    host time allows, and compares every cycle's fired set with the GPU's;
    cpu_baseline: then times a bounded sample of further cycles (about
    target_s seconds) on this host's cores.

    Returns (parity, cpu_baseline); cpu_baseline is None unless baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    log("oracle: building the network ...")
    orc = oracle.from_spec(spec, stdp=stdp)
    counts, ids, chunks = run["counts"], run["ids"], run["chunks"]
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    t_start = time.perf_counter()
    checked, cycle, diverged = 0, 0, None
    for k, apply in chunks:
        if time.perf_counter() - t_start > parity_s:
            break
        rc, ri = orc.run(k)
        gc = counts[cycle:cycle + k]
        gi = ids[offs[cycle]:offs[cycle + k]]
        d = oracle.first_divergence(rc, ri, gc, gi)
        if d is not None:
            diverged = (cycle + d[0], d[1])
            break
        if apply:
            orc.apply_stdp(1.0)
        cycle += k
        checked = cycle
    timed0, timed1 = run["timed"]
    if diverged is not None:
        status = "DIVERGED at cycle %d (%s)" % diverged
    elif checked >= timed1:
        status = "bit-exact: every cycle 0..%d, all timed cycles %d..%d included" % (checked - 1, timed0, timed1 - 1)
    elif checked > timed0:
        status = "bit-exact: cycles 0..%d (timed cycles %d..%d of %d..%d)" % (checked - 1, timed0, checked - 1,
                                                                              timed0, timed1 - 1)
    else:
        status = "bit-exact: cycles 0..%d (warm-up only; no timed cycle checked)" % (checked - 1)
    parity = {"status": status, "checked_cycles": [0, checked], "timed_cycles": [timed0, timed1],
              "first_divergence": None if diverged is None else diverged[0],
              "oracle_s": time.perf_counter() - t_start}
    log("parity: " + status)
    if not baseline or diverged is not None:
        return parity, None
    if cycle < warm_steps:
        orc.run(warm_steps - cycle)
        cycle = warm_steps
    deg, _ = delay_tables(spec)
    t0 = time.perf_counter()
    orc.run(10)
    probe = (time.perf_counter() - t0) / 10
    nsamp = int(max(20, min(20000, target_s / max(probe, 1e-6))))
    log("cpu baseline: %.2f ms/cycle probe; sampling %d cycles" % (probe * 1e3, nsamp))
    t0 = time.perf_counter()
    _, sids = orc.run(nsamp)
    wall = time.perf_counter() - t0
    events = int(deg[sids].sum())
    first = cycle + 10
    return parity, {
        "value": events / wall,
        "unit": "synaptic events/s",
        "cores": int(oracle.load().oracle_num_threads()),
        "kind": "port",
        "sample": "same %d-neuron network, cycles %d..%d (%.1f s, %d events); "
                  "OpenMP restatement of NeMo's CPU backend" % (spec.n, first, first + nsamp - 1, wall, events),
        "ms_per_sim_second": wall * 1e3 / (nsamp / 1000.0),
    }


def build_sim(nemosim_amd, spec, stdp=None, partition_size=None, mode="auto", lanes=0, exchange_waves=0):
    net = spec.to_nemo(nemosim_amd)
    conf = nemosim_amd.Configuration()
    conf.set_cuda_backend(0)
    conf.set_write_only_synapses()
    conf.set_step_mode(mode)
    if exchange_waves:
        conf.set_exchange_waves(exchange_waves)
    if partition_size:
        conf.set_partition_size(partition_size)
    if lanes:
        conf.set_delivery_lanes(lanes)
    if stdp is not None:
        pre, post, a, b, c, d = stdp
        conf.set_stdp_function(pre, post, min_e=a, max_e=b, min_i=c, max_i=d)
    return nemosim_amd.Simulation(net, conf)


def timed_runs(sim, steps, warmup, L, stdp_period=None):
    """warm-up, then `steps` timed cycles in launches of L. The raster of
    every launch, warm-up included, is read back outside the timed interval.
    Returns wall_s, kernel_ms, events, launches and the run record for the
    parity check: per-cycle fired counts and ids, the launch chunks
    (cycles, applyStdp after it) and the timed cycle range."""
    counts, ids, chunks = [], [], []

    def launch(k, done_after):
        ev = sim.run(k)
        apply = bool(stdp_period) and done_after % stdp_period == 0
        if apply:
            sim.apply_stdp(1.0)
        return ev, apply

    done = 0
    while done < warmup:
        k = min(L, warmup - done)
        _, apply = launch(k, done + k)
        c, i = sim.read_fired()
        counts.append(c)
        ids.append(i)
        chunks.append((k, apply))
        done += k
    events, kernel_ms, wall, launches = 0, 0.0, 0.0, 0
    done = 0
    while done < steps:
        k = min(L, steps - done)
        t0 = time.perf_counter()
        ev, apply = launch(k, warmup + done + k)
        wall += time.perf_counter() - t0
        events += ev
        kernel_ms += sim.info()["last_kernel_ms"]
        launches += 1
        c, i = sim.read_fired()        # outside the timed interval
        counts.append(c)
        ids.append(i)
        chunks.append((k, apply))
        done += k
    run = {
        "counts": np.concatenate(counts).astype(np.uint32) if counts else np.zeros(0, np.uint32),
        "ids": np.concatenate(ids).astype(np.uint32) if ids else np.zeros(0, np.uint32),
        "chunks": chunks,
        "timed": (warmup, warmup + steps),
    }
    return wall, kernel_ms, events, launches, run


def bench_plain(args, stdp=False):
    import nemosim_amd
    from nemosim_amd import examples

    n, m = args.neurons, args.synapses
    if stdp and n == 200000:
        n = 50000
    log("generating Izhikevich 80/20 network: %d neurons x %d synapses%s ..." % (n, m, " (STDP)" if stdp else ""))
    spec = examples.izhikevich_80_20(n, m, seed=args.seed, stdp=stdp)
    stdp_fn = None
    if stdp:
        pre, post = examples.stdp_window()
        stdp_fn = (pre, post, 0.0, 1.0, 0.0, -1.0)
    t0 = time.perf_counter()
    sim = build_sim(nemosim_amd, spec, stdp=stdp_fn, partition_size=args.partition_size or None,
                    exchange_waves=args.exchange_waves)
    info = sim.info()
    log("  simulation built in %.1f s: %s" % (time.perf_counter() - t0, info))
    L = max(1, min(args.launch_steps, args.steps))
    deg, rows = delay_tables(spec)
    log("warm-up %d cycles, timed %d cycles in launches of %d" % (args.warmup, args.steps, L))
    wall, kernel_ms, events, launches, run = timed_runs(sim, args.steps, args.warmup, L,
                                                        stdp_period=1000 if stdp else None)
    t0c = args.warmup
    fired = run["ids"][int(run["counts"][:t0c].sum()):]       # the timed cycles' spikes
    abytes, e_check, r_total = algorithmic_bytes(fired.astype(np.int64), deg, rows, n, args.steps)
    if e_check != events:
        log("warning: device event count %d != raster-derived %d" % (events, e_check))
    if stdp:
        # plastic targets that fired (postLen cycles later) walk ~0.8 m in-synapses:
        # 4 B reverse record + 8 B source history + 8 B accumulator read/write each
        indeg_plastic = 0.8 * m
        abytes += int(len(fired) * indeg_plastic * 20) + 8 * n * args.steps
    achieved = abytes / (kernel_ms / 1e3) / 1e9
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
            "workload": "izhikevich_80_20 plain step, STDP %s" % ("on, applyStdp(1.0) every 1000 cycles" if stdp else "off"),
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
            # rocprof FETCH_SIZE + WRITE_SIZE per launch of this shape (profiles/), if measured
            "traffic": traffic_from_profile(profile_for(stdp, n, m, L)),
            "algorithmic_bytes_per_launch": abytes / launches,
            "kernel_ms_per_launch": kernel_ms / launches,
        },
        "cpu_baseline": None,
        "parity": None,
    }
    if not args.no_parity or not args.no_cpu_baseline:
        parity, base = oracle_leg(spec, run, args.cpu_seconds, args.parity_seconds, stdp=stdp_fn,
                                  baseline=not args.no_cpu_baseline)
        result["parity"] = parity
        result["cpu_baseline"] = base
    return result


def bench_step(args):
    """NeMo's per-cycle API: nemo_step (one 1 ms cycle per call, the fired
    ids returned to the host) at the 1k x 100 and 200k x 1000 sizes, called
    through the C ABI with no Python work in the timed loop beyond copying
    the fired ids into a preallocated buffer."""
    import ctypes
    import nemosim_amd
    from nemosim_amd import examples, _lib
    lib = _lib.load()
    lines = []
    for n, m in [(1000, 100), (200000, 1000)]:
        spec = examples.izhikevich_80_20(n, m, seed=args.seed)
        sim = build_sim(nemosim_amd, spec)
        warm, steps = args.warmup, args.steps
        counts = np.zeros(warm + steps, np.uint32)
        ids = np.zeros((warm + steps) * 64 + n, np.uint32)
        fired = ctypes.POINTER(ctypes.c_uint)()
        nf = ctypes.c_size_t()
        pos = 0
        wall = 0.0
        for t in range(warm + steps):
            t0 = time.perf_counter()
            st = lib.nemo_step(sim._h, None, 0, None, None, 0, ctypes.byref(fired), ctypes.byref(nf))
            k = nf.value
            if pos + k > len(ids):
                ids = np.concatenate([ids, np.zeros(len(ids), np.uint32)])
            if k:
                ctypes.memmove(ids.ctypes.data + 4 * pos, fired, 4 * k)
            if t >= warm:
                wall += time.perf_counter() - t0
            if st != 0:
                raise RuntimeError(lib.nemo_strerror().decode())
            counts[t] = k
            pos += k
        deg, _ = delay_tables(spec)
        timed_ids = ids[int(counts[:warm].sum()):pos]
        events = int(deg[timed_ids.astype(np.int64)].sum())
        run = {"counts": counts, "ids": ids[:pos], "chunks": [(1, False)] * (warm + steps), "timed": (warm, warm + steps)}
        parity, _ = oracle_leg(spec, run, 0.0, args.parity_seconds, baseline=False) if not args.no_parity else (None, None)
        lines.append({"neurons": n, "synapses_per_neuron": m, "steps": steps, "warmup": warm,
                      "ms_per_step": wall * 1e3 / steps, "ms_per_sim_second": wall * 1e3 / (steps / 1000.0),
                      "synaptic_events_per_s": events / wall if wall > 0 else None,
                      "kernel_ms_last_step": sim.info()["last_kernel_ms"], "parity": parity})
        log("nemo_step %d x %d: %.1f us per step" % (n, m, wall * 1e6 / steps))
        del sim
    return {
        "metric": "wall ms per nemo_step (one 1 ms cycle per call)",
        "value": lines[-1]["ms_per_step"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "q11.20-i32/f32/f64",
        "data": "synthetic (seeded Izhikevich 80/20 construction)",
        "config": {"workload": "nemo_step per cycle at 1k x 100 and 200k x 1000, STDP off"},
        "sizes": lines,
    }


def profile_for(stdp, n, m, launch):
    """The committed counter summary of this workload and launch length."""
    if not stdp and n == 200000 and m == 1000:
        name = "nemo_step_kernel_c2_launch%d_counters.txt" % launch
    elif stdp and n == 50000 and m == 1000 and launch == 1000:
        name = "nemo_step_kernel_c3_stdp_counters.txt"
    else:
        return None
    return os.path.join(ROOT, "profiles", name)


def traffic_from_profile(path):
    """HBM bytes per step-kernel launch from the committed rocprof counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None if there is none."""
    if not path or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def bench_batch(args):
    """256 networks, N log-uniform in [1k, 20k], 100 synapses, 2000 cycles each:
    one batched persistent launch set vs one plain run per network."""
    import nemosim_amd
    from nemosim_amd import examples, autotune
    count, steps = args.batch_networks, args.batch_steps
    sizes = examples.log_uniform_sizes(count, 1000, 20000, seed=args.seed)
    table = autotune.load_table()
    log("batch: %d networks, %d neurons total, autotune table %s" % (count, int(sizes.sum()), table))
    specs = [examples.izhikevich_80_20(int(n), 100, seed=args.seed + 1 + i) for i, n in enumerate(sizes)]
    t0 = time.perf_counter()
    plain = [build_sim(nemosim_amd, s) for s in specs]
    batched = [build_sim(nemosim_amd, s, partition_size=autotune.partition_size(s.n, table),
                         lanes=autotune.delivery_lanes(s.n, table)) for s in specs]
    log("  built %d simulations in %.1f s" % (2 * count, time.perf_counter() - t0))
    # plain: one persistent run per network, back to back
    t0 = time.perf_counter()
    ev_plain = 0
    for s in plain:
        ev_plain += s.run(steps)
    wall_plain = time.perf_counter() - t0
    # batched: all networks, packed into as few persistent launches as fit
    t0 = time.perf_counter()
    ev_batch = int(np.sum(nemosim_amd.run_batch(batched, [steps] * count)))
    wall_batch = time.perf_counter() - t0
    exact = 0
    abytes = 0
    for spec, a, b in zip(specs, plain, batched):
        ca, ia = a.read_fired()
        cb, ib = b.read_fired()
        exact += int(np.array_equal(ca, cb) and np.array_equal(ia, ib))
        deg, rows = delay_tables(spec)
        abytes += algorithmic_bytes(ib.astype(np.int64), deg, rows, spec.n, steps)[0]
    achieved = abytes / wall_batch / 1e9
    return {
        "metric": "synaptic events/s (batched vs plain)",
        "value": ev_batch / wall_batch,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "q11.20-i32/f32/f64",
        "data": "synthetic (seeded Izhikevich 80/20 construction)",
        "config": {"workload": "%d izhikevich_80_20 networks, N log-uniform [1k,20k], m=100" % count,
                   "neurons_total": int(sizes.sum())},
        "plain_events_per_s": ev_plain / wall_plain,
        "batched_over_plain": (ev_batch / wall_batch) / (ev_plain / wall_plain),
        "wall_plain_s": wall_plain,
        "wall_batched_s": wall_batch,
        "rasters_identical": "%d/%d" % (exact, count),
        "roofline": {
            "bound": "hbm",
            "kernel": "nemo_step_kernel (batched launches)",
            # the algorithmic bytes of every network (8 B per delivered
            # record, 8 B per delivered row, 68 B per neuron and cycle) over
            # the wall time of the whole batched run: launches of several
            # groups of networks and the host work between them included
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes": abytes,
            "time_basis": "wall of the batched run",
        },
    }


def bench_serve(args):
    import nemosim_amd  # noqa: F401
    from nemosim_amd import serving
    reqs = serving.poisson_requests(args.serve_requests, rate_hz=args.serve_rate, seed=args.seed)
    budget = int(args.serve_budget_gb * (1 << 30))
    from nemosim_amd import autotune
    srv = serving.Server(hbm_budget_bytes=budget, chunk_cycles=args.serve_chunk,
                         build_workers=args.serve_workers, keep_rasters=False,
                         device_construction=not args.serve_host_construction,
                         tuning=autotune.load_table() if args.serve_tuned else None,
                         balance=not args.serve_no_balance)
    log("serve: %d requests, budget %.0f GiB, chunks of %d cycles" % (len(reqs), args.serve_budget_gb, args.serve_chunk))
    stats = srv.serve(reqs, realtime=True)
    # sustained load: latency and its parts by arrival order (tenths of the
    # requests), queue depth over the run, offered against completed rate
    arr = np.array([r.arrival_s for r in reqs])
    adm = np.array([r.admitted_s for r in reqs])
    blt = np.array([r.built_s for r in reqs])
    fin = np.array([r.done_s for r in reqs])
    tenths = np.array_split(np.arange(len(reqs)), 10)
    by_arrival = [{"requests": [int(ix[0]), int(ix[-1])],
                   "latency_s_mean": float(np.mean(fin[ix] - arr[ix])),
                   "latency_s_p99": float(np.percentile(fin[ix] - arr[ix], 99)),
                   "wait_s_mean": float(np.mean(adm[ix] - arr[ix])),
                   "build_s_mean": float(np.mean(blt[ix] - adm[ix])),
                   "run_s_mean": float(np.mean(fin[ix] - blt[ix]))} for ix in tenths if len(ix)]
    d = np.array(srv.depth) if srv.depth else np.zeros((0, 4))
    quarters = np.array_split(np.arange(len(d)), 4)
    depth = {"max_waiting": int(d[:, 1].max()) if len(d) else 0,
             "max_building": int(d[:, 2].max()) if len(d) else 0,
             "max_resident": int(d[:, 3].max()) if len(d) else 0,
             "mean_queued_by_quarter": [float(np.mean(d[q, 1] + d[q, 2])) for q in quarters if len(q)],
             "mean_resident_by_quarter": [float(np.mean(d[q, 3])) for q in quarters if len(q)]}
    offered = len(reqs) / float(arr[-1]) if len(reqs) else None
    log("serve: offered %.1f req/s over %.1f s, completed %d in %.1f s" % (offered, arr[-1], stats["requests"], stats["wall_s"]))
    for b in by_arrival:
        log("  requests %4d-%4d: latency mean %.2f s (wait %.2f, build %.2f, run %.2f), p99 %.2f s"
            % (b["requests"][0], b["requests"][1], b["latency_s_mean"], b["wait_s_mean"], b["build_s_mean"],
               b["run_s_mean"], b["latency_s_p99"]))
    log("  queue depth %s" % depth)
    log("  GPU busy %.1f s of %.1f s in %d batched chunks; applyStdp %.1f s, releasing finished simulations %.1f s"
        % (stats["gpu_busy_s"], stats["wall_s"], stats["chunks"], stats["apply_stdp_s"], stats["release_s"]))
    return {
        "metric": "synaptic events/s (serving)",
        "value": stats["events_per_s"],
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": stats["sim_cycles"],
        "warmup": 0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "q11.20-i32/f32/f64",
        "data": "synthetic (Poisson %g req/s, N log-uniform [1k,50k], m U{100..1000}, 30%% STDP, "
                "30%% IF_curr_exp_noise)" % args.serve_rate,
        "config": {"workload": "serving", "requests": stats["requests"], "budget_gib": args.serve_budget_gb,
                   "chunk_cycles": args.serve_chunk, "build_workers": args.serve_workers,
                   "device_construction": not args.serve_host_construction,
                   "balanced_chunks": not args.serve_no_balance},
        "p50_latency_s": stats["p50_latency_s"],
        "p99_latency_s": stats["p99_latency_s"],
        "requests_per_s": stats["requests_per_s"],
        "offered_requests_per_s": offered,
        "gpu_busy_s": stats["gpu_busy_s"],
        "chunks": stats["chunks"],
        "wall_s": stats["wall_s"],
        "latency_by_arrival": by_arrival,
        "queue_depth": depth,
    }


def bench_kuramoto(args):
    """Kuramoto oscillators through the Kuramoto plugin: N oscillators with m
    incoming couplings each (random sources, delays U{1..31}, weights
    N(0, 0.02), frequencies U[0, 0.2)), run in launches of --launch-steps
    cycles. Metric: coupling terms per second (each cycle evaluates every
    coupling once in each of the four Runge-Kutta stages). Parity: the
    oracle runs the same cycles and every oscillator's final phase must
    match bit for bit; the oracle's own time on the host is the CPU
    baseline."""
    import nemosim_amd
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n, m = args.neurons, args.synapses
    rng = np.random.default_rng(args.seed)
    freq = rng.uniform(0.0, 0.2, n).astype(np.float32)
    phase = rng.uniform(0.0, 2 * np.pi, n).astype(np.float32)
    rowptr = np.arange(n + 1, dtype=np.uint64) * m
    tgt = rng.integers(0, n, n * m).astype(np.uint32)
    dly = rng.integers(1, 32, n * m).astype(np.uint8)
    w = rng.normal(0.0, 0.02, n * m).astype(np.float32)
    args2 = np.stack([freq, phase], axis=1)
    net = nemosim_amd.Network()
    ty = net.add_neuron_type("Kuramoto")
    net.add_neurons(ty, np.arange(n, dtype=np.uint32), args2)
    net.add_synapses_csr(rowptr, tgt, dly, w)
    t0 = time.perf_counter()
    sim = nemosim_amd.Simulation(net, nemosim_amd.Configuration())
    log("kuramoto: %d oscillators x %d couplings built in %.1f s" % (n, m, time.perf_counter() - t0))
    L = max(1, min(args.launch_steps, args.steps))
    done = 0
    while done < args.warmup:
        k = min(L, args.warmup - done)
        sim.run(k)
        done += k
    wall, kernel_ms, launches, done = 0.0, 0.0, 0, 0
    while done < args.steps:
        k = min(L, args.steps - done)
        t0 = time.perf_counter()
        sim.run(k)
        wall += time.perf_counter() - t0
        kernel_ms += sim.info()["last_kernel_ms"]
        launches += 1
        done += k
    terms = 4.0 * n * m * args.steps
    got = np.array([sim.get_neuron_state(i, 0) for i in range(n)], np.float32)
    result = {
        "metric": "Kuramoto coupling terms/s",
        "value": terms / wall,
        "unit": "coupling terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "ms_per_sim_second": wall * 1e3 / (args.steps / 1000.0),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (sinf reduction and polynomial in f64)",
        "data": "synthetic (seeded random oscillator network)",
        "config": {"workload": "kuramoto", "oscillators": n, "couplings_per_oscillator": m,
                   "delays": "U{1..31}", "launch_cycles": L},
        "roofline": {
            "bound": "hbm",
            "kernel": "kuramoto_update_kernel + per-cycle plugin-step kernels",
            # 16 B per coupling per cycle read once (source, delay, weight,
            # source phase) + 12 B per oscillator (frequency, phase in, phase out)
            "achieved": (16.0 * n * m + 12.0 * n) * args.steps / (kernel_ms / 1e3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (16.0 * n * m + 12.0 * n) * args.steps / (kernel_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "kernel_ms_per_launch": kernel_ms / launches,
        },
        "cpu_baseline": None,
        "parity": None,
    }
    if not args.no_parity:
        orc = oracle.Oracle()
        orc.add_neurons("Kuramoto", np.arange(n, dtype=np.uint32), args2)
        orc.add_synapses_csr(rowptr, tgt, dly, w)
        orc.finalize()
        t0 = time.perf_counter()
        for _ in range(args.warmup + args.steps):
            orc.step()
        host = time.perf_counter() - t0
        ref = orc.state_all(0)
        bad = int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
        status = ("bit-exact: every oscillator's phase after cycle %d" % (args.warmup + args.steps - 1)
                  if bad == 0 else "DIVERGED: %d of %d phases differ" % (bad, n))
        log("parity: " + status)
        result["parity"] = {"status": status, "oracle_s": host}
        if not args.no_cpu_baseline:
            cyc = args.warmup + args.steps
            result["cpu_baseline"] = {
                "value": 4.0 * n * m * cyc / host, "unit": "coupling terms/s",
                "cores": int(oracle.load().oracle_num_threads()), "kind": "port",
                "sample": "the same network, cycles 0..%d on the oracle (%.1f s)" % (cyc - 1, host),
                "ms_per_sim_second": host * 1e3 / (cyc / 1000.0)}
    return result


def copy_calibration():
    """A plain device copy of known size (1 GiB read + 1 GiB written per
    launch, 4 launches) for scaling rocprof's FETCH_SIZE / WRITE_SIZE to
    bytes: run under the same counter as the step kernel."""
    import torch
    n = 1 << 28                                   # 2^28 float32 = 1 GiB
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"metric": "copy calibration", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
            "launches": 4, "kernel": "torch copy_ (elementwise copy kernel)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="plain",
                    choices=["plain", "stdp", "batch", "serve", "step", "kuramoto", "copy-calibration"])
    ap.add_argument("--steps", type=int, default=5000, help="timed cycles")
    ap.add_argument("--warmup", type=int, default=1000, help="warm-up cycles")
    ap.add_argument("--launch-steps", type=int, default=1000, help="cycles per persistent launch")
    ap.add_argument("--neurons", type=int, default=200000)
    ap.add_argument("--synapses", type=int, default=1000)
    ap.add_argument("--partition-size", type=int, default=0, help="neurons per workgroup (0: automatic)")
    ap.add_argument("--exchange-waves", type=int, default=0,
                    help="waves per workgroup running the exchange (0: automatic, < 0: none)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true", help="skip the oracle replay of the GPU run's cycles")
    ap.add_argument("--parity-seconds", type=float, default=60.0,
                    help="host time the oracle may spend replaying the GPU run's cycles")
    ap.add_argument("--batch-networks", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-chunk", type=int, default=100)
    ap.add_argument("--serve-workers", type=int, default=16,
                    help="serving: host threads building requests (the GPU host's 16-core share)")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-host-construction", action="store_true",
                    help="serving: build each network's synapse structures on the host instead of the device")
    ap.add_argument("--serve-no-balance", action="store_true",
                    help="serving: the same chunk of cycles for every request (no balancing by measured cost)")
    ap.add_argument("--serve-tuned", action="store_true", help="serving: partition sizes and delivery lanes "
                    "from the batch tuning table instead of the backend's defaults")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU; --gpus %d is not supported "
              "(each network, and each batch of networks, runs on a single GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "plain":
        result = bench_plain(args)
    elif args.workload == "stdp":
        result = bench_plain(args, stdp=True)
    elif args.workload == "batch":
        result = bench_batch(args)
    elif args.workload == "step":
        result = bench_step(args)
    elif args.workload == "kuramoto":
        result = bench_kuramoto(args)
    elif args.workload == "copy-calibration":
        result = copy_calibration()
    else:
        result = bench_serve(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
