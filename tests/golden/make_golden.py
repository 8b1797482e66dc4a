"""Writes the regression fixtures of this directory from the oracle (the CPU
restatement of NeMo's CPU backend), for networks fixed by their seeds:

  c1_raster.npz   the 1000-neuron x 100-synapse Izhikevich 80/20 network of
                  nemosim_amd.examples (seed 1), 1000 cycles: fired counts per
                  cycle and fired ids in order
  stdp_weights.npz  the 2000 x 100 network with STDP (seed 13), 5 epochs of
                  100 cycles each followed by applyStdp(1.0): the Q11.20
                  weights of 4000 synapses (indices in CSR order) and the
                  raster of the last epoch

The reference's own regression data files are not shipped with its sources,
so these pin the oracle against itself across changes (test_golden.py), and
the HIP backend against the same numbers on the GPU. Run from the
repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402
from nemosim_amd import examples  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def c1_raster():
    spec = examples.izhikevich_80_20(1000, 100, seed=1)
    o = oracle.from_spec(spec)
    counts, ids = o.run(1000)
    np.savez_compressed(os.path.join(HERE, "c1_raster.npz"), counts=counts.astype(np.uint32),
                        ids=ids.astype(np.uint32))


def stdp_weights():
    spec = examples.izhikevich_80_20(2000, 100, seed=13, stdp=True)
    pre, post = examples.stdp_window()
    o = oracle.from_spec(spec, stdp=(pre, post, 0.0, 1.0, 0.0, -1.0))
    for _ in range(5):
        counts, ids = o.run(100)
        o.apply_stdp(1.0)
    w = o.weights_fx(spec.n, spec.nsyn)
    sample = np.random.default_rng(0).choice(spec.nsyn, size=4000, replace=False).astype(np.uint64)
    np.savez_compressed(os.path.join(HERE, "stdp_weights.npz"), index=sample, weight_fx=w[sample].astype(np.int32),
                        last_counts=counts.astype(np.uint32), last_ids=ids.astype(np.uint32))


if __name__ == "__main__":
    oracle.load()
    c1_raster()
    stdp_weights()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
