"""Pins of the oracle's neuron updates: Izhikevich (T2 mixed precision) and
IF_curr_exp / IF_curr_exp_noise.

The expected values are computed here, independently of the oracle's code,
from the reference's CPU plugins:

  - Izhikevich, src/nemo/cpu/plugins/Izhikevich.cpp:58-89: the right-hand side
    of v and the add to v in IEEE binary64 (numpy float64 arrays: one rounding
    per operation, no contraction), rounded to float once per substep; u in
    float32; four substeps with early exit on firing; forced firing and the
    reset after the loop;
  - IF_curr_exp, src/nemo/cpu/plugins/IF_curr_exp.cpp:81-119, every operation
    in float32 (numpy float32 arrays), plus the build-defined Gaussian term of
    IF_curr_exp_noise (sigma * nrand added once the total current is formed);
  - nrand (src/nemo/RNG.cpp:10-33) rebuilt from the xorshift128 recurrence in
    Python integers, the default-seeded MT19937 (in Python) for the seeds,
    and the host C library's logf/sinf/sqrtf called through ctypes.

Each replay is also run with a plausible mistake (float32 right-hand side,
no early exit, a fused multiply-add in u, noise added before the offset,
`>=` threshold, `<` refractory test, ...) and the test requires every mistake
to change some result on the same inputs, so a restatement carrying that
mistake could not pass.
"""
import ctypes
import ctypes.util
from fractions import Fraction

import numpy as np
import pytest

import netutil

F = np.float32
C_SIN = F(1.4629180792671596810513378043098e-9)
TWO_M32 = F(0.00000000023283064365386962890625)

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
for _name in ("logf", "sinf", "sqrtf"):
    getattr(_libm, _name).restype = ctypes.c_float
    getattr(_libm, _name).argtypes = [ctypes.c_float]


def nrand_py(s):
    """src/nemo/RNG.cpp:26-33 on the state list s (advanced in place)."""
    a = F(F(netutil.xorshift128(s)) * C_SIN)
    b = F(F(netutil.xorshift128(s)) * TWO_M32)
    lg = F(_libm.logf(float(F(F(1) - b))))
    r = F(_libm.sqrtf(float(F(F(-2) * lg))))
    return F(F(_libm.sinf(float(a))) * r)


def seeds(n):
    """Per-neuron RNG state of sim-local neurons 0..n-1 (src/nemo/RNG.cpp:38-63)."""
    d = netutil.mt19937(4 * n)
    return [[d[4 * k + j] >> 1 for j in range(4)] for k in range(n)]


def bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


# ------------------------------------------------------------- Izhikevich

def iz_replay(a, b, c, d, u, v, I, fstim, rhs="f64", early_exit=True, u_fma=False):
    """One cycle of src/nemo/cpu/plugins/Izhikevich.cpp:69-89 for arrays of
    neurons given their total input current I (float32). Returns (u, v, fired)."""
    u = u.copy()
    v = v.copy()
    n = len(v)
    fired = np.zeros(n, bool)
    for _ in range(4):
        act = ~fired if early_exit else np.ones(n, bool)
        if rhs == "f64":
            vd = v.astype(np.float64)
            r = (0.04 * vd + 5.0) * vd + 140.0 - u.astype(np.float64) + I.astype(np.float64)
            vn = (vd + 0.25 * r).astype(np.float32)
        else:
            r = (F(0.04) * v + F(5.0)) * v + F(140.0) - u + I
            vn = v + F(0.25) * r
        v = np.where(act, vn, v)
        if u_fma:
            # b*v is exact in binary64 (24 + 24 significant bits); the
            # difference is rounded once, to float32
            bvu = np.array([F(Fraction(float(bb)) * Fraction(float(vv)) - Fraction(float(uu)))
                            for bb, vv, uu in zip(b, v, u)], np.float32)
        else:
            bvu = b * v - u
        un = u + F(0.25) * (a * bvu)
        u = np.where(act, un, u)
        if early_exit:
            fired = fired | (act & (v.astype(np.float64) >= 30.0))
        else:
            fired = v.astype(np.float64) >= 30.0
    fired = fired | fstim
    v = np.where(fired, c, v)
    u = np.where(fired, u + d, u)
    return u, v, fired


def iz_network(oracle_mod, prm, u0, v0, n_src):
    """Izhikevich neurons 0..n-1 (sim-local = global), then per target one
    excitatory and one inhibitory Input source (ids n + 2k, n + 2k + 1) with
    a delay-1 synapse."""
    n = len(v0)
    o = oracle_mod.Oracle()
    args = np.column_stack([prm["a"], prm["b"], prm["c"], prm["d"], prm["sigma"], u0, v0]).astype(np.float32)
    o.add_neurons("Izhikevich", np.arange(n), args)
    o.add_neurons("Input", np.arange(n, n + 2 * n_src), np.zeros((2 * n_src, 0), np.float32))
    for k in range(n_src):
        o.add_synapse(n + 2 * k, k, 1, float(prm["we"][k]), 0)
        o.add_synapse(n + 2 * k + 1, k, 1, float(prm["wi"][k]), 0)
    o.finalize()
    return o


def fmau_case():
    """A neuron whose b*v - u cancels exactly in float32 after the first
    substep (u == float(b*v)) but not in exact arithmetic: a fused
    multiply-add in the u update changes its u (a = 8 makes the residue
    visible in u's last bit). Found by iterating u <- float(b * v1(u))."""
    a, b = F(8.0), F(0.2)
    v0, I = F(-60.1), F(3.3)
    u = F(-12.0)
    for _ in range(50):
        vd = float(v0)
        v1 = F(vd + 0.25 * ((0.04 * vd + 5.0) * vd + 140.0 - float(u) + float(I)))
        u_next = F(b * v1)
        if u_next == u:
            break
        u = u_next
    return a, b, u, v0, I


@pytest.fixture(scope="module")
def iz_case():
    rng = np.random.default_rng(2024)
    n = 3000
    prm = {
        "a": rng.uniform(0.01, 0.12, n).astype(np.float32),
        "b": rng.uniform(0.15, 0.3, n).astype(np.float32),
        "c": rng.uniform(-70, -50, n).astype(np.float32),
        "d": rng.uniform(1, 9, n).astype(np.float32),
        "sigma": np.where(rng.random(n) < 0.5, rng.uniform(0.5, 6, n), 0.0).astype(np.float32),
        "we": rng.uniform(0.0, 8.0, n).astype(np.float32),
        "wi": rng.uniform(-8.0, 0.0, n).astype(np.float32),
    }
    v0 = rng.uniform(-80, 29.9, n).astype(np.float32)
    u0 = rng.uniform(-20, 10, n).astype(np.float32)
    # neurons 0..9 fire in the first substep of cycle 0
    v0[:10] = F(29.5)
    prm["sigma"][:10] = 0
    # neuron 10: u == float(b*v) after substep 1 (the fused multiply-add case)
    a, b, u, v, _ = fmau_case()
    prm["a"][10], prm["b"][10], u0[10], v0[10], prm["sigma"][10] = a, b, u, v, 0
    return prm, u0, v0


def test_izhikevich_update_pinned(oracle_mod, iz_case):
    prm, u0, v0 = iz_case
    n = len(v0)
    o = iz_network(oracle_mod, prm, u0, v0, n)
    rs = seeds(n)
    assert np.array_equal(o.rng_all()[:n], np.array(rs, np.uint32)), "per-neuron seeds (src/nemo/RNG.cpp:38-63)"
    rng = np.random.default_rng(7)
    u, v = u0.copy(), v0.copy()
    src_prev = np.zeros(2 * n, bool)       # Input sources fired in the previous cycle
    ext_fmaI = fmau_case()[4]
    wrong = {"rhs_f32": 0, "no_early_exit": 0, "noise_first": 0, "u_fma": 0, "no_reset_u": 0}
    fired_in_substep1 = 0
    for step in range(12):
        ext = np.zeros(n, np.float32)
        has = rng.random(n) < 0.6
        ext[has] = rng.uniform(-10, 30, int(has.sum())).astype(np.float32)
        ext[10] = ext_fmaI
        has[10] = True
        forced = rng.random(n) < 0.03
        forced[:11] = False
        src_forced = rng.random(2 * n) < 0.2
        istim = [(int(k), float(ext[k])) for k in np.nonzero(has)[0]]
        fstim = [int(k) for k in np.nonzero(forced)[0]] + [int(n + k) for k in np.nonzero(src_forced)[0]]
        got = o.step(fstim, istim)

        # synaptic input: a delay-1 spike of each source fired last cycle;
        # one Q11.20 weight per plane, converted as wfx_toFloat
        E = np.array([netutil.sat_to_float(netutil.fix20(prm["we"][k])) if src_prev[2 * k] else F(0)
                      for k in range(n)], np.float32)
        Ii = np.array([netutil.sat_to_float(netutil.fix20(prm["wi"][k])) if src_prev[2 * k + 1] else F(0)
                       for k in range(n)], np.float32)
        I = (E + Ii) + ext
        noise = np.zeros(n, np.float32)
        for k in np.nonzero(prm["sigma"] != 0)[0]:
            noise[k] = nrand_py(rs[k])
        I_noisy = np.where(prm["sigma"] != 0, I + prm["sigma"] * noise, I)
        un, vn, fired = iz_replay(prm["a"], prm["b"], prm["c"], prm["d"], u, v, I_noisy, forced)

        # the oracle against the replay: fired ids, u, v bit for bit
        exp_ids = sorted(np.nonzero(fired)[0].tolist() + [n + int(k) for k in np.nonzero(src_forced)[0]])
        assert got == exp_ids, "cycle %d: fired set" % step
        assert same(o.state_all(0)[:n], un), "cycle %d: u" % step
        assert same(o.state_all(1)[:n], vn), "cycle %d: v" % step

        # each plausible mistake changes something on these inputs
        def differs(res):
            return int(np.sum((bits(res[0]) != bits(un)) | (bits(res[1]) != bits(vn)) | (res[2] != fired)))
        wrong["rhs_f32"] += differs(iz_replay(prm["a"], prm["b"], prm["c"], prm["d"], u, v, I_noisy, forced,
                                              rhs="f32"))
        wrong["no_early_exit"] += differs(iz_replay(prm["a"], prm["b"], prm["c"], prm["d"], u, v, I_noisy, forced,
                                                    early_exit=False))
        I_first = np.where(prm["sigma"] != 0, (E + Ii) + (ext + prm["sigma"] * noise), I)
        wrong["noise_first"] += differs(iz_replay(prm["a"], prm["b"], prm["c"], prm["d"], u, v, I_first, forced))
        if step == 0:
            sel = slice(10, 11)
            res = iz_replay(prm["a"][sel], prm["b"][sel], prm["c"][sel], prm["d"][sel], u[sel], v[sel],
                            I_noisy[sel], forced[sel], u_fma=True)
            wrong["u_fma"] += int(not same(res[0], un[sel]))
            fired_in_substep1 = int(np.sum(fired[:10]))
            # forced or not, a fired neuron's u is incremented by d after the loop
            d0 = dict(prm)
            d0["d"] = np.zeros(n, np.float32)
            wrong["no_reset_u"] += differs(iz_replay(d0["a"], d0["b"], d0["c"], d0["d"], u, v, I_noisy, forced))
        u, v = un, vn
        src_prev = src_forced
    # RNG: advanced by two draws per cycle exactly for the neurons with sigma != 0
    assert np.array_equal(o.rng_all()[:n], np.array(rs, np.uint32))
    untouched = seeds(n)
    quiet = prm["sigma"] == 0
    assert np.array_equal(o.rng_all()[:n][quiet], np.array(untouched, np.uint32)[quiet]), \
        "sigma = 0 must leave the RNG untouched"
    assert fired_in_substep1 == 10
    for k, cnt in wrong.items():
        assert cnt > 0, "mistake %s would go unnoticed" % k


def test_izhikevich_first_substep_fire_and_forced_reset(oracle_mod):
    """Closed-form cases: a neuron crossing 30 in the first substep updates u
    once and then u += d; a forced neuron is reset whatever its v."""
    o = oracle_mod.Oracle()
    a, b, c, d = F(0.02), F(0.2), F(-65.0), F(8.0)
    # neuron 0: v0 = 29.5 with I = 20 crosses 30 in substep 1
    o.add_neuron("Izhikevich", 0, a, b, c, d, 0.0, F(-10.0), F(29.5))
    # neuron 1: at rest, forced
    o.add_neuron("Izhikevich", 1, a, b, c, d, 0.0, F(b * F(-65.0)), F(-65.0))
    o.finalize()
    fired = o.step([1], [(0, 20.0)])
    assert fired == [0, 1]
    vd = 29.5
    v1 = F(vd + 0.25 * ((0.04 * vd + 5.0) * vd + 140.0 - (-10.0) + 20.0))
    assert float(v1) >= 30.0
    u1 = F(F(-10.0) + F(0.25) * (a * (b * v1 - F(-10.0))))
    assert o.get_neuron_state(0, 1) == c
    assert same([o.get_neuron_state(0, 0)], [u1 + d])
    # neuron 1: four substeps from rest, then the forced reset
    u, v = F(b * F(-65.0)), F(-65.0)
    for _ in range(4):
        vd = float(v)
        v = F(vd + 0.25 * ((0.04 * vd + 5.0) * vd + 140.0 - float(u) + 0.0))
        u = F(u + F(0.25) * (a * (b * v - u)))
    assert float(v) < 30.0
    assert o.get_neuron_state(1, 1) == c
    assert same([o.get_neuron_state(1, 0)], [u + d])


# ---------------------------------------------------- IF_curr_exp (+ noise)

IF_P = ["v_rest", "v_reset", "c_m", "tau_m", "tau_refrac", "tau_syn_E", "tau_syn_I", "v_thresh", "i_offset"]


def if_replay(p, st, E, Ii, ext, noise, fstim, noise_first=False, ge=False, refr_lt=False, decay_sub=False):
    """One cycle of src/nemo/cpu/plugins/IF_curr_exp.cpp:81-119 (float32),
    with the build-defined noise sigma * nrand added to the total current."""
    v, Ie0, Ii0, lf = st
    one = F(1.0)
    if decay_sub:
        Ie = (Ie0 - Ie0 / p["tau_syn_E"]) + E
        Ii_ = (Ii0 - Ii0 / p["tau_syn_I"]) + Ii
    else:
        Ie = ((one - one / p["tau_syn_E"]) * Ie0) + E
        Ii_ = ((one - one / p["tau_syn_I"]) * Ii0) + Ii
    if noise_first:
        I = ((Ie + Ii_) + ext) + (p["i_offset"] + noise)
    else:
        I = (((Ie + Ii_) + ext) + p["i_offset"]) + noise
    refr = (lf < p["tau_refrac"]) if refr_lt else (lf <= p["tau_refrac"])
    vn = np.where(refr, v, v + (I / p["c_m"] + (p["v_rest"] - v) / p["tau_m"]))
    over = (vn >= p["v_thresh"]) if ge else (vn > p["v_thresh"])
    fired = (~refr & over) | fstim
    lf_n = np.where(fired, one, lf + one)
    v_n = np.where(fired, p["v_reset"], vn)
    return (v_n.astype(np.float32), Ie.astype(np.float32), Ii_.astype(np.float32), lf_n.astype(np.float32)), fired


@pytest.mark.parametrize("noisy", [False, True])
def test_if_curr_exp_update_pinned(oracle_mod, noisy):
    rng = np.random.default_rng(11 if noisy else 12)
    n = 2500
    p = {
        "v_rest": rng.uniform(-70, -60, n), "v_reset": rng.uniform(-70, -60, n), "c_m": rng.uniform(0.5, 2, n),
        "tau_m": rng.uniform(5, 40, n), "tau_refrac": rng.integers(0, 5, n).astype(float),
        "tau_syn_E": rng.uniform(1, 10, n), "tau_syn_I": rng.uniform(1, 10, n),
        "v_thresh": rng.uniform(-56, -46, n), "i_offset": rng.uniform(-1, 2, n),
    }
    p = {k: np.asarray(x, np.float32) for k, x in p.items()}
    sigma = np.where(rng.random(n) < 0.6, rng.uniform(0.5, 5, n), 0.0).astype(np.float32) if noisy \
        else np.zeros(n, np.float32)
    v0 = rng.uniform(-70, -45, n).astype(np.float32)
    Ie0 = rng.uniform(-2, 5, n).astype(np.float32)
    Ii0 = rng.uniform(-5, 0, n).astype(np.float32)
    lf0 = np.where(rng.random(n) < 0.3, rng.integers(0, 6, n), 1000).astype(np.float32)
    # neurons 0..4 land exactly on the threshold (no leak, c_m = 1): strict '>' does not fire
    p["tau_m"][:5] = np.inf
    p["c_m"][:5] = 1
    p["i_offset"][:5] = 0
    p["tau_syn_E"][:5] = p["tau_syn_I"][:5] = 1
    Ie0[:5] = Ii0[:5] = 0
    v0[:5] = -60
    p["v_thresh"][:5] = -50
    lf0[:5] = 1000
    sigma[:5] = 0
    # neurons 5..9 sit exactly at the end of their refractory period
    p["tau_refrac"][5:10] = 2
    lf0[5:10] = 2
    # synapses from Input sources (delay 1) into every target
    we = rng.uniform(0, 4, n).astype(np.float32)
    wi = rng.uniform(-4, 0, n).astype(np.float32)
    model = "IF_curr_exp_noise" if noisy else "IF_curr_exp"
    cols = [p[k] for k in IF_P] + ([sigma] if noisy else []) + [v0, Ie0, Ii0, lf0]
    o = oracle_mod.Oracle()
    o.add_neurons(model, np.arange(n), np.column_stack(cols).astype(np.float32))
    o.add_neurons("Input", np.arange(n, 3 * n), np.zeros((2 * n, 0), np.float32))
    for k in range(n):
        o.add_synapse(n + 2 * k, k, 1, float(we[k]), 0)
        o.add_synapse(n + 2 * k + 1, k, 1, float(wi[k]), 0)
    o.finalize()
    rs = seeds(n)
    st = (v0, Ie0, Ii0, lf0)
    src_prev = np.zeros(2 * n, bool)
    wrong = {"ge": 0, "refr_lt": 0, "decay_sub": 0}
    if noisy:
        wrong["noise_first"] = 0
    for step in range(10):
        ext = np.zeros(n, np.float32)
        has = rng.random(n) < 0.6
        ext[has] = rng.uniform(-3, 12, int(has.sum())).astype(np.float32)
        ext[:5] = 10.0
        has[:5] = True
        forced = rng.random(n) < 0.03
        forced[:10] = False
        src_forced = rng.random(2 * n) < 0.25
        istim = [(int(k), float(ext[k])) for k in np.nonzero(has)[0]]
        fstim = [int(k) for k in np.nonzero(forced)[0]] + [int(n + k) for k in np.nonzero(src_forced)[0]]
        got = o.step(fstim, istim)
        E = np.array([netutil.sat_to_float(netutil.fix20(we[k])) if src_prev[2 * k] else F(0)
                      for k in range(n)], np.float32)
        Ii = np.array([netutil.sat_to_float(netutil.fix20(wi[k])) if src_prev[2 * k + 1] else F(0)
                       for k in range(n)], np.float32)
        noise = np.zeros(n, np.float32)
        for k in np.nonzero(sigma != 0)[0]:
            noise[k] = F(sigma[k] * nrand_py(rs[k]))
        new, fired = if_replay(p, st, E, Ii, ext, noise, forced)
        exp_ids = sorted(np.nonzero(fired)[0].tolist() + [n + int(k) for k in np.nonzero(src_forced)[0]])
        assert got == exp_ids, "cycle %d: fired set" % step
        for var in range(4):
            assert same(o.state_all(var)[:n], new[var]), "cycle %d: state variable %d" % (step, var)
        if step == 0:
            assert not fired[:5].any(), "v == v_thresh does not fire (strict >)"

        def differs(res):
            s2, f2 = res
            return int(np.sum((f2 != fired) | np.any([bits(s2[i]) != bits(new[i]) for i in range(4)], axis=0)))
        wrong["ge"] += differs(if_replay(p, st, E, Ii, ext, noise, forced, ge=True))
        wrong["refr_lt"] += differs(if_replay(p, st, E, Ii, ext, noise, forced, refr_lt=True))
        wrong["decay_sub"] += differs(if_replay(p, st, E, Ii, ext, noise, forced, decay_sub=True))
        if noisy:
            wrong["noise_first"] += differs(if_replay(p, st, E, Ii, ext, noise, forced, noise_first=True))
        st = new
        src_prev = src_forced
    assert np.array_equal(o.rng_all()[:n], np.array(rs, np.uint32)), "RNG advanced by the noise draws only"
    for k, cnt in wrong.items():
        assert cnt > 0, "mistake %s would go unnoticed" % k


def test_if_curr_exp_noise_integrator(oracle_mod):
    """With tau_syn = 1, no leak (tau_m = inf), c_m = 1, no offset and an
    unreachable threshold, v is the float32 running sum of sigma * nrand of
    the neuron's own stream; sigma = 0 neurons stay put and keep their seed."""
    n = 6
    sig = np.array([3.0, 0.0, 1.5, 5.0, 0.0, 0.25], np.float32)
    o = oracle_mod.Oracle()
    for k in range(n):
        o.add_neuron("IF_curr_exp_noise", k, -65.0, -65.0, 1.0, np.inf, 0.0, 1.0, 1.0, np.inf, 0.0, sig[k],
                     -65.0, 0.0, 0.0, 1000.0)
    o.finalize()
    rs = seeds(n)
    v = np.full(n, -65.0, np.float32)
    for _ in range(300):
        assert o.step() == []
        for k in range(n):
            if sig[k] != 0:
                v[k] = F(v[k] + F(sig[k] * nrand_py(rs[k])))
    assert same(o.state_all(0), v)
    assert np.array_equal(o.rng_all(), np.array(rs, np.uint32))
    assert np.array_equal(o.rng_all()[sig == 0], np.array(seeds(n), np.uint32)[sig == 0])
