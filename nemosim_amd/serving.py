"""Serving: continuous batching of simulation requests on one GPU.

Requests (a network to build and a number of 1 ms cycles to run, optionally
with STDP and periodic applyStdp) arrive over time. The server

  * admits a request when its estimated device footprint (plus, when networks
    are built on the device, the memory the build holds meanwhile) fits the
    HBM budget
    (requests wait in arrival order otherwise);
  * builds admitted networks on host worker threads while the GPU runs;
  * keeps every admitted simulation resident on the device and advances all
    of them together, a chunk of cycles at a time, in one batched persistent
    launch (nemo_hip_run_batch) -- requests that arrive while others run join
    at the next chunk boundary, finished ones leave and free their budget;
  * records each request's raster and its latency (completion minus arrival).

Batching never changes a result: each simulation has its own state, RNG
streams and delivery, so a request's raster is the same alone and in the mix
(tests/test_serving.py checks this).
"""
import concurrent.futures
import heapq
import threading
import time

import numpy as np

from . import autotune
from . import examples
from . import nemo as nemo_api


class Request(object):
    def __init__(self, rid, arrival_s, neurons, synapses, duration_ms, model="Izhikevich",
                 stdp=False, seed=0, stdp_period=1000, reward=1.0):
        self.rid = rid
        self.arrival_s = arrival_s
        self.neurons = int(neurons)
        self.synapses = int(synapses)
        self.duration_ms = int(duration_ms)
        self.model = model
        self.stdp = bool(stdp)
        self.seed = int(seed)
        self.stdp_period = int(stdp_period)
        self.reward = float(reward)
        # filled by the server
        self.admitted_s = None
        self.built_s = None
        self.done_s = None
        self.cycles_done = 0
        self.events = 0
        self.counts = []
        self.ids = []
        self.sim = None
        self.cost_ms = None   # step-kernel ms per cycle in its last batched launch

    @property
    def latency_s(self):
        return None if self.done_s is None else self.done_s - self.arrival_s

    def spec(self):
        if self.model == "Izhikevich":
            return examples.izhikevich_80_20(self.neurons, self.synapses, seed=self.seed, stdp=self.stdp)
        return examples.if_curr_exp_80_20(self.neurons, self.synapses, seed=self.seed, stdp=self.stdp)

    def estimated_device_bytes(self):
        """Footprint of the resident simulation: synapse records 8/synapse,
        bucket offsets 4*(P+1)/neuron (P ~ n/256 partitions), neuron arrays
        ~100 B/neuron, delay ring and exchange buffers, STDP reverse matrix
        12 B/plastic synapse plus 16 B/neuron of history."""
        n, m = self.neurons, self.synapses
        p = max(1, (n + 255) // 256)
        b = 8 * n * m + 4 * n * (p + 2) + 100 * n + 8 * 20 * n + 16 * n
        if self.stdp:
            b += 12 * int(0.8 * n * m) + 16 * n
        return int(b * 1.1) + (1 << 20)

    def construction_bytes(self):
        """Device memory held only while the network is built on the device:
        the uploaded synapses and their mapping (30 B/synapse: source,
        target, weight, local source, local target, Q11.20 weight, delay,
        plastic flag, record index) and three bucket counters per (source,
        partition)."""
        n, m = self.neurons, self.synapses
        p = max(1, (n + 255) // 256)
        return 30 * n * m + 12 * n * (p + 1) + (1 << 20)


def build_simulation(req, device_construction=False, tuning=None, synapse_queries=False):
    """The request's simulation. Built on the host by default: a serving
    worker then needs the device only for the final upload, while a device
    build holds the device lock between chunks (measured on 200 Poisson
    requests: p50 latency 1.8 s host-built, 3.1 s device-built). With a
    tuning table (nemosim_amd.autotune.load_table()) the partition size and
    delivery lanes come from it instead of the backend's defaults; results
    do not depend on either. Synapse queries (getSynapseWeight ...) are off
    unless synapse_queries: their host tables are not kept."""
    spec = req.spec()
    net = spec.to_nemo(nemo_api)
    conf = nemo_api.Configuration()
    if not synapse_queries:
        conf.set_write_only_synapses()
    conf.set_device_construction(device_construction)
    if tuning:
        T = autotune.partition_size(req.neurons, tuning)
        if T:
            conf.set_partition_size(T)
        conf.set_delivery_lanes(autotune.delivery_lanes(req.neurons, tuning))
    if req.stdp:
        pre, post = examples.stdp_window()
        conf.set_stdp_function(pre, post, min_e=0.0, max_e=1.0, min_i=0.0, max_i=-1.0)
    return nemo_api.Simulation(net, conf)


def poisson_requests(count, rate_hz, seed, size_range=(1000, 50000), syn_range=(100, 1000),
                     dur_range=(100, 2000), stdp_frac=0.3, if_frac=0.3):
    """The serving workload: Poisson arrivals, log-uniform sizes, uniform
    synapse counts and durations, a fraction with STDP, a fraction of
    IF_curr_exp_noise networks."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    reqs = []
    for i in range(count):
        t += rng.exponential(1.0 / rate_hz)
        n = int(round(np.exp(rng.uniform(np.log(size_range[0]), np.log(size_range[1])))))
        m = int(rng.integers(syn_range[0], syn_range[1] + 1))
        dur = int(rng.integers(dur_range[0], dur_range[1] + 1))
        stdp = bool(rng.random() < stdp_frac)
        model = "IF_curr_exp_noise" if rng.random() < if_frac else "Izhikevich"
        reqs.append(Request(i, t, n, m, dur, model=model, stdp=stdp, seed=1000 + i))
    return reqs


class Server(object):
    def __init__(self, hbm_budget_bytes, chunk_cycles=100, build_workers=4, keep_rasters=True,
                 device_construction=False, tuning=None, keep_finished=False, balance=True):
        """keep_finished: a finished request keeps its simulation (built
        with synapse queries on) for inspection after serve(); its budget
        is released all the same. balance: cycles per chunk by each
        request's measured cost (see _advance)."""
        self.keep_finished = keep_finished
        self.balance = balance
        self.budget = int(hbm_budget_bytes)
        self.device_construction = device_construction
        self.tuning = tuning
        self.chunk = int(chunk_cycles)
        self.keep_rasters = keep_rasters
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=build_workers)
        # finished simulations are released on this thread: freeing device
        # memory waits for the device, which the serving loop should not
        self.reaper = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        self.used = 0
        self.lock = threading.Lock()
        self.gpu_busy_s = 0.0
        self.chunks = 0       # batched launches of the resident set
        self.apply_s = 0.0    # time in applyStdp calls
        self.release_s = 0.0  # time releasing finished simulations
        self.depth = []       # (time, waiting, building, resident) at each pass of the loop

    def _transient(self, r):
        """Device memory a request holds only while it is built."""
        return r.construction_bytes() if self.device_construction else 0

    def plan_steps(self, active):
        """Cycles each resident request runs in the next chunk. STDP
        requests stop at their applyStdp boundaries. A batched launch lasts
        as long as its slowest network, so with balance each request runs
        about as many cycles as fit in the time the median request needs
        for chunk cycles, by the step-kernel time per cycle each one took
        in its last chunk (between a tenth of a chunk and 4 chunks' worth,
        so that no request is starved; chunk cycles until measured):
        networks that fire fast or are large no longer hold the others' CUs
        idle. Results do not depend on how the cycles are cut."""
        known = [r.cost_ms for r in active if r.cost_ms]
        budget = self.chunk * float(np.median(known)) if (self.balance and known) else None
        steps = []
        for r in active:
            k = self.chunk
            if budget is not None and r.cost_ms:
                k = int(max(max(1, self.chunk // 10), min(4 * self.chunk, round(budget / r.cost_ms))))
            k = min(k, r.duration_ms - r.cycles_done)
            if r.stdp:
                k = min(k, r.stdp_period - (r.cycles_done % r.stdp_period))
            steps.append(k)
        return steps

    def _advance(self, active):
        """One chunk for every resident simulation (plan_steps)."""
        steps = self.plan_steps(active)
        sims = [r.sim for r in active]
        t0 = time.perf_counter()
        ev = nemo_api.run_batch(sims, steps)
        self.gpu_busy_s += time.perf_counter() - t0
        self.chunks += 1
        due = []
        for r, k, e in zip(active, steps, ev):
            r.events += int(e)
            if self.balance and k > 0:
                ms = r.sim.last_kernel_ms()
                if ms > 0.0:
                    r.cost_ms = ms / k
            if self.keep_rasters:
                c, ids = r.sim.read_fired()
                r.counts.append(c.copy())
                r.ids.append(ids.copy())
            r.cycles_done += k
            if r.stdp and r.cycles_done % r.stdp_period == 0:
                due.append(r)
        if due:
            # every applyStdp due at this boundary under one hold of the device lock
            t1 = time.perf_counter()
            nemo_api.apply_stdp_batch([r.sim for r in due], [r.reward for r in due])
            self.apply_s += time.perf_counter() - t1

    def serve(self, requests, realtime=True):
        """Run all requests to completion. With realtime, arrivals follow
        the wall clock from now; otherwise all are present at once.
        Returns a stats dict."""
        pending = sorted(requests, key=lambda r: r.arrival_s)
        t_start = time.perf_counter()
        now = lambda: time.perf_counter() - t_start
        waiting = []          # arrived, not admitted (FIFO)
        building = []         # (future, request)
        active = []
        done = []
        i = 0
        while i < len(pending) or waiting or building or active:
            t = now() if realtime else float("inf")
            while i < len(pending) and pending[i].arrival_s <= t:
                waiting.append(pending[i])
                i += 1
            if not realtime and i < len(pending):
                waiting.extend(pending[i:])
                i = len(pending)
            # admission by budget, in arrival order
            while waiting:
                r = waiting[0]
                need = r.estimated_device_bytes() + self._transient(r)
                if need > self.budget:
                    raise ValueError("request %d needs %d B, more than the whole budget" % (r.rid, need))
                if self.used + need > self.budget:
                    break
                waiting.pop(0)
                self.used += need
                r.admitted_s = now()
                building.append((self.pool.submit(build_simulation, r, self.device_construction, self.tuning,
                                                  self.keep_finished), r))
            # built networks join the resident set
            still = []
            for fut, r in building:
                if fut.done():
                    r.sim = fut.result()
                    r.built_s = now()
                    self.used -= self._transient(r)
                    active.append(r)
                else:
                    still.append((fut, r))
            building = still
            self.depth.append((now(), len(waiting), len(building), len(active)))
            if active:
                self._advance(active)
                finished = [r for r in active if r.cycles_done >= r.duration_ms]
                for r in finished:
                    r.done_s = now()
                    if not self.keep_finished:
                        t1 = time.perf_counter()
                        self.reaper.submit(_release, [r.sim])
                        r.sim = None
                        self.release_s += time.perf_counter() - t1
                    self.used -= r.estimated_device_bytes()
                    done.append(r)
                active = [r for r in active if r.cycles_done < r.duration_ms]
            elif building:
                concurrent.futures.wait([f for f, _ in building], timeout=0.05,
                                        return_when=concurrent.futures.FIRST_COMPLETED)
            elif i < len(pending):
                time.sleep(max(0.0, min(0.01, pending[i].arrival_s - now())))
        self.reaper.submit(lambda: None).result()   # every release done
        wall = now()
        lat = np.array([r.latency_s for r in done])
        events = sum(r.events for r in done)
        return {
            "requests": len(done),
            "wall_s": wall,
            "gpu_busy_s": self.gpu_busy_s,
            "chunks": self.chunks,
            "apply_stdp_s": self.apply_s,
            "release_s": self.release_s,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "synaptic_events": int(events),
            "events_per_s": events / wall if wall > 0 else None,
            "sim_cycles": int(sum(r.duration_ms for r in done)),
            "requests_per_s": len(done) / wall if wall > 0 else None,
        }


def _release(holder):
    """Drop the last reference to a simulation (holder: a one-element
    list), so that it is destroyed on the calling thread."""
    holder.pop()


def raster_of(req):
    """(counts, ids) of a served request over its whole run."""
    c = np.concatenate(req.counts) if req.counts else np.zeros(0, np.uint32)
    i = np.concatenate(req.ids) if req.ids else np.zeros(0, np.uint32)
    return c, i


def run_alone(req, keep_sim=False):
    """The same request on its own (for the alone-vs-mix identity): its
    (counts, ids), and with keep_sim also the simulation (built with
    synapse queries on)."""
    sim = build_simulation(req, synapse_queries=keep_sim)
    counts, ids = [], []
    done = 0
    while done < req.duration_ms:
        k = req.duration_ms - done
        if req.stdp:
            k = min(k, req.stdp_period - (done % req.stdp_period))
        sim.run(k)
        c, i = sim.read_fired()
        counts.append(c.copy())
        ids.append(i.copy())
        done += k
        if req.stdp and done % req.stdp_period == 0:
            sim.apply_stdp(req.reward)
    if keep_sim:
        return np.concatenate(counts), np.concatenate(ids), sim
    return np.concatenate(counts), np.concatenate(ids)
