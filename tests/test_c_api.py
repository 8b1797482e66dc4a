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


# ------------------------------------------------------------ the C++ surface

CPP_PROG = os.path.join(HERE, "c_api", "cpp_api_check")


def _run_cpp(*args):
    if not os.path.exists(CPP_PROG):
        subprocess.check_call(["make", "-C", os.path.join(HERE, "c_api")])
    r = subprocess.run([CPP_PROG] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)


def test_cpp_network_and_configuration():
    """include/nemo.hpp's Network and Configuration (no device): queries,
    weight range, copies, NeMo's exceptions with their status codes"""
    _run_cpp("net")


def _load_cpp_network(path):
    with open(path, "rb") as f:
        n, m = (int(x) for x in np.frombuffer(f.read(8), np.uint32))
        args = np.frombuffer(f.read(4 * 7 * n), np.float32).reshape(n, 7)
        targets = np.frombuffer(f.read(4 * n * m), np.uint32)
        delays = np.frombuffer(f.read(n * m), np.uint8)
        weights = np.frombuffer(f.read(4 * n * m), np.float32)
        plastic = np.frombuffer(f.read(n * m), np.uint8)
        pre = np.frombuffer(f.read(4 * 20), np.float32)
        post = np.frombuffer(f.read(4 * 20), np.float32)
    groups = [(examples.IZHIKEVICH, np.arange(n, dtype=np.uint32), args.copy())]
    spec = examples.NetSpec(groups, np.arange(0, n * m + 1, m, dtype=np.uint64), targets.copy(), delays.copy(),
                            weights.copy(), plastic.copy())
    return spec, (pre.copy(), post.copy(), 0.0, 1.0, 0.0, -1.0)


@pytest.mark.gpu
def test_cpp_run_matches_oracle(oracle_mod, tmp_path):
    """A program written against NeMo's C++ classes (nemo::Network,
    nemo::Configuration with the four-argument setStdpFunction,
    nemo::Simulation::step with firing and current stimulus, applyStdp,
    getSynapseWeight): 2000 cycles with STDP against the oracle, raster and
    every weight bit-exact"""
    netfile, rasterfile, weightfile = tmp_path / "net.bin", tmp_path / "raster.txt", tmp_path / "w.bin"
    _run_cpp("run", netfile, rasterfile, weightfile)
    got = [list(map(int, line.split()))[1:] for line in open(rasterfile)]
    assert len(got) == 2000
    spec, stdp = _load_cpp_network(netfile)
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    first = orc.step([1, 2], [(3, 15.0), (4, 15.0), (5, 15.0)])
    rc1, ri1 = orc.run(999)
    orc.apply_stdp(1.0)
    rc2, ri2 = orc.run(1000)
    orc.apply_stdp(1.0)
    counts = np.concatenate([[len(first)], rc1, rc2])
    ids = np.concatenate([np.asarray(first, np.int64), np.asarray(ri1, np.int64), np.asarray(ri2, np.int64)])
    gc = np.array([len(f) for f in got])
    gi = np.array([x for f in got for x in f], np.int64)
    d = first_divergence(counts, ids, gc, gi)
    assert d is None, "first diverging cycle %d: %s" % d
    w_ref = orc.weights_fx(spec.n, spec.nsyn).astype(np.float32) / np.float32(1 << 20)
    w_got = np.fromfile(weightfile, np.float32)
    assert np.array_equal(w_got.view(np.uint32), w_ref.astype(np.float32).view(np.uint32))
    # STDP moved some weights
    assert np.any(w_ref != np.array([np.float32(int(np.trunc(np.float64(w) * (1 << 20)))) / np.float32(1 << 20)
                                     for w in spec.weights], np.float32))
