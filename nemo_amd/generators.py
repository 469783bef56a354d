"""Synthetic network generators for the benchmark workloads and tests.

Each generator returns a NetworkSpec of plain numpy arrays, which can be
loaded into any object with NeMo's Network interface (add_neuron_type,
add_neuron, add_synapse), so the same network reaches the HIP backend and the
CPU oracle.

random_izhikevich: the 80/20 excitatory/inhibitory Izhikevich network of
NeMo's examples (src/examples/randomNetUtils.cpp:32-56, after Izhikevich
2003): per-neuron r ~ U[0,1); excitatory a=0.02, b=0.2, c=-65+15r^2,
d=8-6r^2; inhibitory a=0.02+0.08r, b=0.25-0.05r, c=-65, d=2; v=-65, u=b*v;
Gaussian input sigma 5 (exc) / 2 (inh); m outgoing synapses per neuron to
uniformly random targets with integer delays U{1..dmax}; weights 0.5*U[0,1)
(exc) and -U[0,1) (inh); excitatory synapses plastic when stdp is set.
Random streams come from numpy's PCG64 with the given seed (the reference
uses Boost streams, which are not reproducible here).
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class NeuronBlock:
    model: str
    idx: np.ndarray          # uint32 user indices
    args: list               # per-argument arrays (parameters then state)


@dataclass
class NetworkSpec:
    blocks: list = field(default_factory=list)
    source: np.ndarray = None
    target: np.ndarray = None
    delay: np.ndarray = None
    weight: np.ndarray = None
    plastic: np.ndarray = None

    @property
    def neuron_count(self):
        return sum(b.idx.size for b in self.blocks)

    @property
    def synapse_count(self):
        return 0 if self.source is None else self.source.size

    def load(self, net, chunk=1 << 24):
        """Add the neurons and synapses to a Network-like object."""
        types = {}
        for b in self.blocks:
            if b.model not in types:
                types[b.model] = net.add_neuron_type(b.model)
            net.add_neuron(types[b.model], b.idx, *b.args)
        if self.source is not None:
            for s in range(0, self.source.size, chunk):
                e = min(self.source.size, s + chunk)
                net.add_synapse(self.source[s:e], self.target[s:e], self.delay[s:e],
                                self.weight[s:e], self.plastic[s:e])
        return net


def izhikevich_params(n, ne, rng, sigma=(5.0, 2.0)):
    """Per-neuron (a, b, c, d, sigma, u, v) for ne excitatory then n-ne inhibitory."""
    r = rng.random(n, dtype=np.float32)
    exc = np.arange(n) < ne
    a = np.where(exc, np.float32(0.02), np.float32(0.02) + np.float32(0.08) * r).astype(np.float32)
    b = np.where(exc, np.float32(0.2), np.float32(0.25) - np.float32(0.05) * r).astype(np.float32)
    c = np.where(exc, np.float32(-65.0) + np.float32(15.0) * r * r, np.float32(-65.0)).astype(np.float32)
    d = np.where(exc, np.float32(8.0) - np.float32(6.0) * r * r, np.float32(2.0)).astype(np.float32)
    sg = np.where(exc, np.float32(sigma[0]), np.float32(sigma[1])).astype(np.float32)
    v = np.full(n, -65.0, np.float32)
    u = (b * v).astype(np.float32)
    return [a, b, c, d, sg, u, v], exc


def random_izhikevich(n, m, dmax=20, seed=1, stdp=False, sigma=(5.0, 2.0), first_index=0,
                      exc_weight=0.5, inh_weight=1.0):
    rng = np.random.default_rng(seed)
    ne = int(round(0.8 * n))
    args, exc = izhikevich_params(n, ne, rng, sigma)
    idx = (np.arange(n, dtype=np.uint32) + np.uint32(first_index))
    spec = NetworkSpec(blocks=[NeuronBlock("Izhikevich", idx, args)])
    if m > 0:
        src_local = np.repeat(np.arange(n, dtype=np.uint32), m)
        spec.source = src_local + np.uint32(first_index)
        spec.target = rng.integers(0, n, size=n * m, dtype=np.uint32) + np.uint32(first_index)
        spec.delay = rng.integers(1, dmax + 1, size=n * m, dtype=np.uint32)
        u = rng.random(n * m, dtype=np.float32)
        src_exc = src_local < ne
        spec.weight = np.where(src_exc, np.float32(exc_weight) * u, -np.float32(inh_weight) * u).astype(np.float32)
        spec.plastic = (src_exc & stdp).astype(np.uint8)
    return spec


def batch_izhikevich(count=64, nmin=1000, nmax=50000, smin=100, smax=1000, dmax=20, seed=4, seed_base=1000,
                     stdp_every=2):
    """The batched workload: count independent 80/20 Izhikevich networks
    (random_izhikevich), network i with N_i log-uniform in [nmin, nmax],
    S_i uniform in {smin..smax}, its own seed seed_base + i, and plastic
    excitatory synapses when i % stdp_every == 0.  Returns a list of
    (spec, stdp) pairs."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(np.int64)
    ss = rng.integers(smin, smax + 1, count)
    out = []
    for i in range(count):
        stdp = stdp_every > 0 and i % stdp_every == 0
        out.append((random_izhikevich(int(ns[i]), int(ss[i]), dmax=dmax, seed=seed_base + i, stdp=stdp), stdp))
    return out


def torus_mixed(n_lif=40000, n_qif=8000, n_hh=2000, n_poisson=1000, m=500, sigma=0.1, seed=3,
                cm=400.0, i_offset=280.0, qif_a=0.1, rate=0.01, grid=50):
    """Mixed-model network on the unit torus: IF_curr_exp (v_rest -65 mV,
    v_reset -65, tau_m 20 ms, refractory 2 ms, tau_syn 5 ms, threshold -50 mV;
    cm 400 nF and i_offset 280 nA, so that the neuron rests 1 mV below
    threshold (v_rest + tau_m i_offset / cm = -51 mV) and one synaptic event
    of weight w moves v by about tau_syn w / cm, a few microvolts: the LIF
    population fires at about 10 Hz, where cm 1 nF drives the whole network
    to its refractory ceiling), QIF (a = 0.1), HH
    (resting gating state) and Poisson sources (rate per ms; 0.01 = 10 Hz),
    added in that order, neuron i at a position U[0,1)^2.  Every neuron has m
    outgoing synapses to non-source neurons with probability proportional to
    exp(-d^2 / (2 sigma^2)) of the torus distance d, and delay ceil(20 d) + 1
    ms.  Targets are drawn by adding a Gaussian offset (sigma) to the source
    position and taking a uniformly chosen neuron of the grid cell (side
    1/grid) holding that point: the cell discretises the kernel at 1/grid, far
    below sigma.  Weights U[0,1) from excitatory and -U[0,1) from inhibitory
    sources; 20% of the non-source neurons, chosen at random, are
    inhibitory.  No synapse is plastic."""
    rng = np.random.default_rng(seed)
    n = n_lif + n_qif + n_hh
    total = n + n_poisson
    pos = rng.random((total, 2))
    blocks = []
    base = 0

    def block(model, count, args):
        nonlocal base
        idx = np.arange(base, base + count, dtype=np.uint32)
        base += count
        blocks.append(NeuronBlock(model, idx, [np.broadcast_to(np.float32(a), (count,)).astype(np.float32)
                                               if np.isscalar(a) else np.asarray(a, np.float32) for a in args]))

    block("IF_curr_exp", n_lif, [-65.0, -65.0, cm, 20.0, 2.0, 5.0, 5.0, -50.0, i_offset,
                                 -65.0, 0.0, 0.0, 100.0])
    block("QIF", n_qif, [qif_a, 0.0])
    block("HH", n_hh, [-65.0, 0.3177, 0.0529, 0.5961, 0.0])
    block("PoissonSource", n_poisson, [rate])
    spec = NetworkSpec(blocks=blocks)

    # grid cells of the candidate targets (the non-source neurons)
    cell = (np.floor(pos[:n, 0] * grid).astype(np.int64) % grid) * grid + \
        (np.floor(pos[:n, 1] * grid).astype(np.int64) % grid)
    order = np.argsort(cell, kind="stable")
    starts = np.searchsorted(cell[order], np.arange(grid * grid))
    ends = np.searchsorted(cell[order], np.arange(grid * grid), side="right")
    if np.any(ends == starts):
        raise ValueError("empty grid cell: use a coarser grid")

    src = np.repeat(np.arange(total, dtype=np.uint32), m)
    pt = (pos[src] + rng.normal(0.0, sigma, size=(src.size, 2))) % 1.0
    c = (np.floor(pt[:, 0] * grid).astype(np.int64) % grid) * grid + (np.floor(pt[:, 1] * grid).astype(np.int64) % grid)
    pick = starts[c] + (rng.random(src.size) * (ends[c] - starts[c])).astype(np.int64)
    tgt = order[pick].astype(np.uint32)
    dxy = np.abs(pos[tgt] - pos[src])
    dxy = np.minimum(dxy, 1.0 - dxy)
    dist = np.sqrt((dxy * dxy).sum(axis=1))
    inh = np.zeros(total, bool)
    inh[:n] = rng.random(n) < 0.2
    u = rng.random(src.size, dtype=np.float32)
    spec.source = src
    spec.target = tgt
    spec.delay = (np.ceil(20.0 * dist).astype(np.uint32) + 1).astype(np.uint32)
    spec.weight = np.where(inh[src], -u, u).astype(np.float32)
    spec.plastic = np.zeros(src.size, np.uint8)
    return spec


def ring(ncount, n0=0, plastic=False, nstep=1, delay=1, sigma=0.0):
    """NeMo's test ring (src/test/utils.cpp:117-140): neuron i excites neuron
    i+1 with weight 1000 at the given delay; RS-like Izhikevich neurons."""
    v, b, r = np.float32(-65.0), np.float32(0.2), np.float32(0.5)
    r2 = r * r
    i = np.arange(ncount, dtype=np.uint32)
    idx = (np.uint32(n0) + (i % ncount) * np.uint32(nstep)).astype(np.uint32)
    tgt = (np.uint32(n0) + ((i + 1) % ncount) * np.uint32(nstep)).astype(np.uint32)
    one = np.ones(ncount, np.float32)
    args = [0.02 * one, b * one, (v + np.float32(15.0) * r2) * one, (np.float32(8.0) - np.float32(6.0) * r2) * one,
            np.float32(sigma) * one, (b * v) * one, v * one]
    spec = NetworkSpec(blocks=[NeuronBlock("Izhikevich", idx, [a.astype(np.float32) for a in args])])
    spec.source = idx
    spec.target = tgt
    spec.delay = np.full(ncount, delay, np.uint32)
    spec.weight = np.full(ncount, 1000.0, np.float32)
    spec.plastic = np.full(ncount, 1 if plastic else 0, np.uint8)
    return spec


def merge(*specs):
    """Concatenate neuron blocks and synapses of several specs."""
    out = NetworkSpec()
    for s in specs:
        out.blocks.extend(s.blocks)
    syn = [s for s in specs if s.source is not None]
    if syn:
        out.source = np.concatenate([s.source for s in syn])
        out.target = np.concatenate([s.target for s in syn])
        out.delay = np.concatenate([s.delay for s in syn])
        out.weight = np.concatenate([s.weight for s in syn])
        out.plastic = np.concatenate([s.plastic for s in syn])
    return out


# ------------------------------------------------------------------ STDP
def stdp_exp_function(a_plus=0.1, a_minus=-0.12, tau=20.0, window=20):
    """pre[i] = a_plus exp(-(i+1)/tau), post[i] = a_minus exp(-(i+1)/tau)
    (sampled at dt = i+1 as in src/test/utils.cpp:89-95)."""
    dt = np.arange(1, window + 1, dtype=np.float32)
    pre = (np.float32(a_plus) * np.exp(-dt / np.float32(tau))).astype(np.float32)
    post = (np.float32(a_minus) * np.exp(-dt / np.float32(tau))).astype(np.float32)
    return pre, post
