"""Seeded example networks (the counterpart of NeMo's src/examples/random.cpp).

``izhikevich_80_20`` builds the 80/20 excitatory/inhibitory Izhikevich network
of the benchmark workloads: per neuron r ~ U[0,1); excitatory (first 80 %):
a=0.02, b=0.2, c=-65+15 r^2, d=8-6 r^2, sigma=5; inhibitory: a=0.02+0.08 r,
b=0.25-0.05 r, c=-65, d=2, sigma=2; v0=-65, u0=b*v0. Each neuron has m
outgoing synapses to uniformly random targets (with replacement), delays
U{1..dmax}, weights U[0,0.5) (excitatory) or (-1,0] (inhibitory) as float32.
Excitatory synapses are marked plastic when ``stdp`` is set.

The generator is numpy (PCG64) seeded, so the oracle and the HIP backend are
fed identical arrays. Networks are returned as a ``NetSpec`` that can be
loaded into either.
"""
import numpy as np

IZHIKEVICH = "Izhikevich"
IF_CURR_EXP = "IF_curr_exp"
IF_CURR_EXP_NOISE = "IF_curr_exp_noise"

MODEL_IDS = {IZHIKEVICH: 0, IF_CURR_EXP: 1, IF_CURR_EXP_NOISE: 2}


class NetSpec(object):
    """Neuron groups plus a CSR over sources 0..n-1 of synapses."""

    def __init__(self, groups, rowptr, targets, delays, weights, plastic):
        self.groups = groups          # list of (model name, idx uint32 (k,), args float32 (k, nargs))
        self.rowptr = rowptr          # uint64 (nsrc+1,)
        self.targets = targets        # uint32
        self.delays = delays          # uint8
        self.weights = weights        # float32
        self.plastic = plastic        # uint8
        self.n = int(sum(len(g[1]) for g in groups))

    @property
    def nsyn(self):
        return int(self.rowptr[-1])

    def to_nemo(self, nemo_mod):
        """Build a nemosim_amd Network."""
        net = nemo_mod.Network()
        for name, idx, args in self.groups:
            t = net.add_neuron_type(name)
            net.add_neurons(t, idx, args)
        net.add_synapses_csr(self.rowptr, self.targets, self.delays, self.weights, self.plastic)
        return net


def _synapses(rng, n, m, dmax, exc_mask, stdp):
    nsyn = n * m
    rowptr = np.arange(0, nsyn + 1, m, dtype=np.uint64)
    targets = rng.integers(0, n, size=nsyn, dtype=np.uint32)
    delays = rng.integers(1, dmax + 1, size=nsyn, dtype=np.uint8)
    u = rng.random(nsyn, dtype=np.float32)
    src_exc = np.repeat(exc_mask, m)
    weights = np.where(src_exc, np.float32(0.5) * u, -u).astype(np.float32)
    plastic = (src_exc & bool(stdp)).astype(np.uint8)
    return rowptr, targets, delays, weights, plastic


def izhikevich_80_20(n, m, seed, dmax=20, stdp=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    ne = (n * 4) // 5
    r = rng.random(n).astype(np.float32)
    exc = np.arange(n) < ne
    a = np.where(exc, np.float32(0.02), np.float32(0.02) + np.float32(0.08) * r).astype(np.float32)
    b = np.where(exc, np.float32(0.2), np.float32(0.25) - np.float32(0.05) * r).astype(np.float32)
    c = np.where(exc, np.float32(-65.0) + np.float32(15.0) * r * r, np.float32(-65.0)).astype(np.float32)
    d = np.where(exc, np.float32(8.0) - np.float32(6.0) * r * r, np.float32(2.0)).astype(np.float32)
    sigma = np.where(exc, np.float32(5.0), np.float32(2.0)).astype(np.float32)
    v = np.full(n, -65.0, np.float32)
    u = (b * v).astype(np.float32)
    args = np.stack([a, b, c, d, sigma, u, v], axis=1).astype(np.float32)
    groups = [(IZHIKEVICH, np.arange(n, dtype=np.uint32), args)]
    rowptr, targets, delays, weights, plastic = _synapses(rng, n, m, dmax, exc, stdp)
    return NetSpec(groups, rowptr, targets, delays, weights, plastic)


def if_curr_exp_80_20(n, m, seed, dmax=20, stdp=False, sigma=3.0, noise=True):
    """80/20 network of IF_curr_exp(_noise) neurons: v_rest=-65, v_reset=-65,
    c_m=1, tau_m=20, tau_refrac=2, tau_syn_E=tau_syn_I=5, v_thresh=-50,
    i_offset=0 (+ sigma); initial v=-65, Ie=Ii=0, lastfired=1000."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ne = (n * 4) // 5
    exc = np.arange(n) < ne
    p = [-65.0, -65.0, 1.0, 20.0, 2.0, 5.0, 5.0, -50.0, 0.0]
    if noise:
        p = p + [sigma]
    s = [-65.0, 0.0, 0.0, 1000.0]
    args = np.tile(np.array(p + s, np.float32), (n, 1))
    name = IF_CURR_EXP_NOISE if noise else IF_CURR_EXP
    groups = [(name, np.arange(n, dtype=np.uint32), args)]
    rowptr, targets, delays, weights, plastic = _synapses(rng, n, m, dmax, exc, stdp)
    return NetSpec(groups, rowptr, targets, delays, weights, plastic)


TORUS_PATCH = 32           # patch width and height (neurons)
TORUS_MAX_DELAY = 20


def torus(pcount, m, seed, stdp=False, sigma=16.0):
    """The toroidal network of src/examples/torus.cpp:216-286, the network of
    NeMo's high-firing test: pcount patches of 32 x 32 neurons side by side
    along x (neuron patch*1024 + y*32 + x), each neuron excitatory with
    probability 0.8. Excitatory: a=0.02, b=0.2, c=-65+15 r1^2, d=8-6 r2^2,
    sigma=5; inhibitory: a=0.02+0.08 r1, b=0.25-0.05 r2, c=-65, d=2, sigma=2;
    v0=-65, u0=b*v0 (:68-96). Each neuron has m synapses to the neuron at
    distance 1+|N(0, s)| (s = sigma excitatory, 16 inhibitory) in a uniform
    direction, offsets rounded half away from zero, wrapping around the torus
    (:110-150); delay 1 + distance/32 from the truncated distance, at most 20
    (:153-163); weight 0.5 U (excitatory, plastic when stdp) or -U
    (inhibitory, static). The draws come from numpy's PCG64, not Boost's
    MT19937: the same shape of network, not the same network."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W = TORUS_PATCH
    n = pcount * W * W
    exc = rng.random(n) < 0.8
    r1 = rng.random(n).astype(np.float32)
    r2 = rng.random(n).astype(np.float32)
    v = np.full(n, -65.0, np.float32)
    a = np.where(exc, np.float32(0.02), np.float32(0.02) + np.float32(0.08) * r1).astype(np.float32)
    b = np.where(exc, np.float32(0.2), np.float32(0.25) - np.float32(0.05) * r2).astype(np.float32)
    c = np.where(exc, v + np.float32(15.0) * r1 * r1, v).astype(np.float32)
    d = np.where(exc, np.float32(8.0) - np.float32(6.0) * r2 * r2, np.float32(2.0)).astype(np.float32)
    sig = np.where(exc, np.float32(5.0), np.float32(2.0)).astype(np.float32)
    u = (b * v).astype(np.float32)
    args = np.stack([a, b, c, d, sig, u, v], axis=1).astype(np.float32)
    groups = [(IZHIKEVICH, np.arange(n, dtype=np.uint32), args)]

    src = np.repeat(np.arange(n, dtype=np.int64), m)
    src_exc = exc[src]
    scale = np.where(src_exc, float(sigma), W / 2.0)
    dist = 1.0 + np.abs(rng.normal(0.0, 1.0, size=n * m) * scale)
    theta = rng.uniform(0.0, 2.0 * np.pi, size=n * m)

    def away(x):
        return np.where(x > 0.0, np.floor(x + 0.5), np.ceil(x - 0.5)).astype(np.int64)

    patch, rem = np.divmod(src, W * W)
    sy, sx = np.divmod(rem, W)
    gx = np.mod(patch * W + sx + away(dist * np.cos(theta)), W * pcount)
    ty = np.mod(sy + away(dist * np.sin(theta)), W)
    targets = ((gx // W) * W * W + ty * W + gx % W).astype(np.uint32)
    idist = dist.astype(np.int64)
    delays = np.where(idist >= TORUS_MAX_DELAY * W, TORUS_MAX_DELAY, 1 + idist // W).astype(np.uint8)
    uw = rng.random(n * m, dtype=np.float32)
    weights = np.where(src_exc, np.float32(0.5) * uw, -uw).astype(np.float32)
    plastic = (src_exc & bool(stdp)).astype(np.uint8)
    rowptr = np.arange(0, n * m + 1, m, dtype=np.uint64)
    return NetSpec(groups, rowptr, targets, delays, weights, plastic)


def stdp_window(pre_amp=0.1, post_amp=-0.12, tau=20.0, length=20):
    """Exponential STDP window sampled at dt = 1..length
    (src/examples/common.cpp:154-165 sampling convention)."""
    pre = [np.float32(pre_amp) * np.float32(np.exp(-(i + 1) / tau)) for i in range(length)]
    post = [np.float32(post_amp) * np.float32(np.exp(-(i + 1) / tau)) for i in range(length)]
    return np.array(pre, np.float32), np.array(post, np.float32)


def log_uniform_sizes(count, lo, hi, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int64)
