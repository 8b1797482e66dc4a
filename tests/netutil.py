"""Small-network builders shared by the oracle tests and the GPU parity tests.

Each builder takes a ``sim_like`` factory-free target: any object with
add_neuron(model, idx, *args) and add_synapse(s, t, d, w, plastic) -- the
oracle, or an adapter over nemosim_amd.Network (NetAdapter)."""
import numpy as np

IZ = "Izhikevich"
IF = "IF_curr_exp"


class NetAdapter(object):
    """Gives nemosim_amd.Network the oracle's add_neuron(model, idx, ...)."""

    def __init__(self, nemo_mod):
        self.net = nemo_mod.Network()
        self.types = {}

    def add_neuron(self, model, idx, *args):
        if model not in self.types:
            self.types[model] = self.net.add_neuron_type(model)
        self.net.add_neuron(self.types[model], idx, *args)

    def add_synapse(self, s, t, d, w, plastic=0):
        return self.net.add_synapse(s, t, d, w, plastic)


def iz_regular(target, idx, sigma=0.0):
    """Regular-spiking Izhikevich neuron at rest (a,b,c,d = .02,.2,-65,8)."""
    target.add_neuron(IZ, idx, 0.02, 0.2, -65.0, 8.0, sigma, 0.2 * -65.0, -65.0)


def ring(target, n, delay=1, first=0):
    """n Izhikevich neurons in a ring with strong synapses: once neuron 0 is
    forced to fire, exactly one neuron fires every `delay` cycles
    (src/test/test.cpp:299-325)."""
    for i in range(n):
        iz_regular(target, first + i)
    for i in range(n):
        target.add_synapse(first + i, first + (i + 1) % n, delay, 1000.0, 0)


def if_integrator(target, idx):
    """IF_curr_exp neuron that integrates its input exactly and never fires:
    tau_syn=1 (no carry-over of synaptic current), c_m=1, tau_m=inf (no
    leak), huge threshold. v(t+1) = v(t) + ((E+I)+ext)."""
    inf = np.float32(np.inf)
    target.add_neuron(IF, idx, -65.0, -65.0, 1.0, inf, 0.0, 1.0, 1.0, 1e30, 0.0,
                      -65.0, 0.0, 0.0, 1000.0)


def if_source(target, idx):
    """IF_curr_exp neuron that only fires when forced (threshold unreachable)."""
    target.add_neuron(IF, idx, -65.0, -65.0, 1.0, 20.0, 0.0, 5.0, 5.0, 1e30, 0.0,
                      -65.0, 0.0, 0.0, 1000.0)


def fix20(w):
    """Q11.20 truncation of a float32 weight (src/nemo/fixedpoint.cpp:18-29),
    computed with exact integer arithmetic."""
    w = np.float32(w)
    # float32 * 2^20 is exact; truncation toward zero
    return int(np.trunc(np.float64(w) * (1 << 20)))


def sat_to_float(v):
    """int64 fixed-point sum -> float with saturation (fixedpoint.cpp:41-49):
    clamp to int32 range, then float(v)/2^20 in float32."""
    v = max(min(v, 2**31 - 1), -2**31)
    return np.float32(np.float32(v) / np.float32(1 << 20))


def mt19937(n):
    """Reference MT19937 (Matsumoto & Nishimura), default seed 5489, first n
    32-bit outputs, in Python integers."""
    mt = [0] * 624
    mt[0] = 5489
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xffffffff
    idx = 624
    out = []
    for _ in range(n):
        if idx >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7fffffff)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908b0df if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9d2c5680
        y ^= (y << 15) & 0xefc60000
        y ^= y >> 18
        out.append(y & 0xffffffff)
    return out


def xorshift128(s):
    """One step of Marsaglia's xorshift128 on the state list s (updated in
    place); returns the output."""
    t = (s[0] ^ ((s[0] << 11) & 0xffffffff)) & 0xffffffff
    s[0], s[1], s[2] = s[1], s[2], s[3]
    s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8))
    return s[3]
