"""NeMo's C-API tests (src/test/c_api.cpp) through a C program
(tests/c_api/c_api_check.c) built against include/nemo.h and linked with
libnemo_hip.so: synapse ids, synapse queries from the network and from the
simulation, and a run driven by nemo_step whose raster is compared with the
oracle's on the network the program built."""
import os
import subprocess

import numpy as np
import pytest

from backends import first_divergence
from nemosim_amd import examples

HERE = os.path.dirname(os.path.abspath(__file__))
PROG = os.path.join(HERE, "c_api", "c_api_check")


def _run(*args):
    if not os.path.exists(PROG):
        subprocess.check_call(["make", "-C", os.path.join(HERE, "c_api")])
    r = subprocess.run([PROG] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)


def test_synapse_ids():
    """src/test/c_api.cpp:336-355 (network only, no device)"""
    _run("ids")


@pytest.mark.gpu
@pytest.mark.parametrize("n0", [0, 1000])
def test_get_synapses(n0):
    """src/test/c_api.cpp:524-597"""
    _run("synapses", n0)


def _load_network(path):
    with open(path, "rb") as f:
        n, m = np.frombuffer(f.read(8), np.uint32)
        n, m = int(n), int(m)
        args = np.frombuffer(f.read(4 * 7 * n), np.float32).reshape(n, 7)
        targets = np.frombuffer(f.read(4 * n * m), np.uint32)
        weights = np.frombuffer(f.read(4 * n * m), np.float32)
    groups = [(examples.IZHIKEVICH, np.arange(n, dtype=np.uint32), args.copy())]
    return examples.NetSpec(groups, np.arange(0, n * m + 1, m, dtype=np.uint64), targets.copy(),
                            np.ones(n * m, np.uint8), weights.copy(), np.zeros(n * m, np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("fstim,istim", [(0, 0), (1, 0), (0, 1)])
def test_run_matches_oracle(oracle_mod, tmp_path, fstim, istim):
    """src/test/c_api.cpp:155-315 (there the C API against the C++ API):
    2000 cycles through nemo_step, optional firing stimulus of neuron 100
    and current stimulus into neurons 20, 40, 60 in cycle 0, against the
    oracle on the same network"""
    netfile, rasterfile = tmp_path / "net.bin", tmp_path / "raster.txt"
    _run("run", netfile, rasterfile, fstim, istim)
    got = [list(map(int, line.split()))[1:] for line in open(rasterfile)]
    assert len(got) == 2000
    orc = oracle_mod.from_spec(_load_network(netfile))
    first = orc.step([100] if fstim else [], [(20, 20.0), (40, 20.0), (60, 20.0)] if istim else None)
    rc, ri = orc.run(1999)
    counts = np.concatenate([[len(first)], rc])
    ids = np.concatenate([np.asarray(first, np.int64), np.asarray(ri, np.int64)])
    gc = np.array([len(f) for f in got])
    gi = np.array([x for f in got for x in f], np.int64)
    d = first_divergence(counts, ids, gc, gi)
    assert d is None, "first diverging cycle %d: %s" % d
    assert counts.sum() > 1000
