"""The neuron-model plugin boundary (include/nemo_hip_plugin.h): descriptors
on the plugin path (NeMo's <Model>.ini format, src/nemo/NeuronType.cpp:92-155),
lib<Model>.so loading, and the plugin step giving the same results as the
fused step and the oracle."""
import os
import shutil

import numpy as np
import pytest

from backends import first_divergence
from nemosim_amd import examples

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGINS = os.path.join(ROOT, "nemosim_amd", "plugins")


def _external_model_dir(tmp_path, name, template="Izhikevich"):
    """A plugin directory holding an external model `name`: a copy of a
    built-in descriptor and its plugin library under the new name."""
    d = tmp_path / "plugins"
    (d / "hip").mkdir(parents=True)
    shutil.copy(os.path.join(PLUGINS, template + ".ini"), str(d / (name + ".ini")))
    shutil.copy(os.path.join(PLUGINS, "hip", "lib%s.so" % template), str(d / "hip" / ("lib%s.so" % name)))
    return str(d)


BUILTIN = ("Izhikevich", "IF_curr_exp", "IF_curr_exp_noise", "PoissonSource", "Input")
SHIPPED = BUILTIN + ("Kuramoto",)   # plugins shipped in nemosim_amd/plugins


def test_plugin_libraries_built():
    for m in SHIPPED:
        assert os.path.exists(os.path.join(PLUGINS, "hip", "lib%s.so" % m))
        assert os.path.exists(os.path.join(PLUGINS, m + ".ini"))


def test_plugin_exports_abi_symbols():
    import ctypes
    for m in SHIPPED:
        lib = ctypes.CDLL(os.path.join(PLUGINS, "hip", "lib%s.so" % m))
        assert hasattr(lib, "hip_update_neurons") and hasattr(lib, "hip_init_neurons")


def test_external_descriptor_on_plugin_path(nemo_mod, tmp_path):
    d = _external_model_dir(tmp_path, "MyIzhikevich")
    with pytest.raises(nemo_mod.NemoError) as e:
        nemo_mod.Network().add_neuron_type("MyIzhikevich")
    assert e.value.status == 12
    nemo_mod.add_plugin_path(d)
    net = nemo_mod.Network()
    t = net.add_neuron_type("MyIzhikevich")
    net.add_neuron(t, 0, 0.02, 0.2, -65.0, 8.0, 5.0, -13.0, -65.0)
    with pytest.raises(nemo_mod.NemoError):
        net.add_neuron(t, 1, 0.02, 0.2, -65.0)        # descriptor: 5 + 2 arguments
    assert net.get_neuron_state(0, 1) == np.float32(-65.0)


def test_user_plugin_directory(nemo_mod, tmp_path, monkeypatch):
    """$HOME/.nemo/plugins is searched after the program's own paths and
    before the system directory (src/nemo/NeuronType.cpp:50-67)."""
    d = _external_model_dir(tmp_path, "HomeIzhikevich")
    home = tmp_path / "home"
    (home / ".nemo").mkdir(parents=True)
    shutil.move(d, str(home / ".nemo" / "plugins"))
    with pytest.raises(nemo_mod.NemoError):
        nemo_mod.Network().add_neuron_type("HomeIzhikevich")
    monkeypatch.setenv("HOME", str(home))
    net = nemo_mod.Network()
    t = net.add_neuron_type("HomeIzhikevich")
    net.add_neuron(t, 0, 0.02, 0.2, -65.0, 8.0, 5.0, -13.0, -65.0)


def test_info_tool(tmp_path):
    """python -m nemosim_amd: the counterpart of NeMo's nemo utility
    (src/app/nemo.cpp:46-98)."""
    import subprocess
    import sys

    def run(*args):
        env = dict(os.environ, HOME=str(tmp_path))
        return subprocess.run([sys.executable, "-m", "nemosim_amd"] + list(args), cwd=ROOT, env=env,
                              capture_output=True, text=True)
    r = run("--version")
    assert r.returncode == 0 and r.stdout.strip().startswith("0.7.2")
    r = run("--system-plugin-path")
    assert r.returncode == 0 and os.path.samefile(r.stdout.strip(), PLUGINS)
    r = run("--user-plugin-path")
    assert r.returncode == 0 and r.stdout.strip() == str(tmp_path / ".nemo" / "plugins")
    r = run("--help")
    assert r.returncode == 0 and "--list-devices" in r.stdout
    r = run()
    assert r.returncode == 1 and "--list-devices" in r.stdout
    r = run("--no-such-option")
    assert r.returncode == 1


@pytest.mark.gpu
def test_info_tool_lists_devices():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-m", "nemosim_amd", "-l"], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0
    assert r.stdout.startswith("0: "), r.stdout


def test_plugin_path_must_exist(nemo_mod, tmp_path):
    with pytest.raises(nemo_mod.NemoError) as e:
        nemo_mod.add_plugin_path(str(tmp_path / "nope"))
    assert e.value.status == 6


def _sim(nemo_mod, spec, mode, stdp=None):
    net = spec.to_nemo(nemo_mod)
    conf = nemo_mod.Configuration()
    conf.set_step_mode(mode)
    if stdp is not None:
        pre, post, a, b, c, d = stdp
        conf.set_stdp_function(pre, post, min_e=a, max_e=b, min_i=c, max_i=d)
    return nemo_mod.Simulation(net, conf)


@pytest.mark.gpu
def test_plugin_step_c1_parity(nemo_mod, oracle_mod):
    spec = examples.izhikevich_80_20(1000, 100, seed=11)
    sim = _sim(nemo_mod, spec, "plugins")
    orc = oracle_mod.from_spec(spec)
    sim.run(600)
    gc, gi = sim.read_fired()
    rc, ri = orc.run(600)
    d = first_divergence(rc, ri, gc, gi)
    assert d is None, "first diverging cycle %d: %s" % d
    v = np.array([sim.get_membrane_potential(i) for i in range(1000)], np.float32)
    assert np.array_equal(v.view(np.uint32), orc.state_all(1).view(np.uint32))


@pytest.mark.gpu
def test_plugin_step_equals_fused_step_with_stimulus(nemo_mod):
    spec = examples.if_curr_exp_80_20(1500, 80, seed=6)
    a = _sim(nemo_mod, spec, "plugins")
    b = _sim(nemo_mod, spec, "fused")
    rng = np.random.default_rng(3)
    for t in range(200):
        fs = sorted(set(rng.integers(0, 1500, size=2).tolist()))
        ist = [(int(i), 30.0) for i in rng.integers(0, 1500, size=5)]
        assert a.step(fs, ist) == b.step(fs, ist), t


@pytest.mark.gpu
def test_plugin_step_stdp_parity(nemo_mod, oracle_mod):
    spec = examples.izhikevich_80_20(1500, 100, seed=17, stdp=True)
    pre, post = examples.stdp_window()
    stdp = (pre, post, 0.0, 1.0, 0.0, -1.0)
    sim = _sim(nemo_mod, spec, "plugins", stdp=stdp)
    orc = oracle_mod.from_spec(spec, stdp=stdp)
    for epoch in range(3):
        sim.run(100)
        orc_c, orc_i = orc.run(100)
        gc, gi = sim.read_fired()
        d = first_divergence(orc_c, orc_i, gc, gi)
        assert d is None, "epoch %d first diverging cycle %d: %s" % ((epoch,) + d)
        sim.apply_stdp(1.0)
        orc.apply_stdp(1.0)
    ref_w = orc.weights_fx(1500, spec.nsyn)
    for k in np.random.default_rng(1).integers(0, spec.nsyn, size=2000):
        src = int(np.searchsorted(spec.rowptr, k, side="right") - 1)
        sid = (src << 32) | int(k - spec.rowptr[src])
        assert sim.get_synapse_weight(sid) == np.float32(np.float32(ref_w[k]) / np.float32(1 << 20))


@pytest.mark.gpu
def test_external_model_runs_through_its_plugin(nemo_mod, oracle_mod, tmp_path):
    """A model unknown to the fused kernel (loaded from the plugin path) runs
    through the plugin step; here it is an Izhikevich copy, so the oracle's
    Izhikevich raster is the expected result."""
    nemo_mod.add_plugin_path(_external_model_dir(tmp_path, "IzhikevichExt"))
    spec = examples.izhikevich_80_20(800, 50, seed=2)
    ext = examples.NetSpec([("IzhikevichExt",) + spec.groups[0][1:]], spec.rowptr, spec.targets,
                           spec.delays, spec.weights, spec.plastic)
    sim = _sim(nemo_mod, ext, "auto")
    assert sim.info()["partitions"] == 1 and sim.info()["partition_size"] == 0   # plugin step
    orc = oracle_mod.from_spec(spec)
    sim.run(300)
    gc, gi = sim.read_fired()
    rc, ri = orc.run(300)
    d = first_divergence(rc, ri, gc, gi)
    assert d is None, "first diverging cycle %d: %s" % d


def test_descriptor_history_and_rcm_fields(nemo_mod, tmp_path):
    """history= of a descriptor is read: history=0 (no cycle of state kept)
    is refused when the type is added."""
    d = tmp_path / "plugins"
    (d / "hip").mkdir(parents=True)
    (d / "ZeroHistory.ini").write_text("parameters=1\nstate-variables=1\nmembrane-potential=0\nhistory=0\n")
    nemo_mod.add_plugin_path(str(d))
    net = nemo_mod.Network()
    with pytest.raises(nemo_mod.NemoError):
        net.add_neuron_type("ZeroHistory")
