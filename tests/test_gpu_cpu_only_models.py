"""NeMo's CPU-only models (src/nemo/cpu/plugins: IzhikevichChaos,
IzhikevichSyn, Izhikevich2007, Izhikevich2007DopSyn, BayesianNeuron,
InhomogeneousPoissonSource) on the HIP backend, through the C ABI, against
the oracle: a random Izhikevich network drives a population of each model
through excitatory and inhibitory synapses (and the models drive the
network back), with current and firing stimuli; rasters, every state
variable and the RNG states bit-exact after every step."""
import numpy as np
import pytest

import nemo_amd
from nemo_amd import generators as gen
import oracle_api as oa

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu(require_gpu):
    pass


def same_f32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    an, bn = np.isnan(a), np.isnan(b)
    return np.array_equal(an, bn) and np.array_equal(a.view(np.uint32)[~an], b.view(np.uint32)[~bn])


MSN = [0.01, -20.0, -55.0, 91.0, 40.0, -80.0, -29.7, 1.0, 15.2]
TAIL = [0.9, 1.0, 0.0, 0.66, 1.0, -80.0]


def model_args(model, rng, n):
    """Per-neuron parameter and state columns of n neurons of `model`."""
    u = rng.random(n)
    if model == "IzhikevichChaos":
        return [0.02 + 0.02 * u, 0.2 + 0 * u, -65 + 10 * u, 8 - 4 * u, 0.5 + u, -13 + 0 * u, -65 + 5 * u]
    if model == "IzhikevichSyn":
        return ([0.02 + 0 * u, 0.2 + 0 * u, -65 + 0 * u, 8 + 0 * u, 30 + 0 * u, 2 * u, 0.9 + 0 * u, 1 + 0 * u,
                 0 * u, 0.66 + 0 * u, 1 + 0 * u, -80 + 0 * u, -13 + 0 * u, -65 + 5 * u, 0.3 * u, 0.2 * u])
    if model == "Izhikevich2007":
        return [x + 0 * u for x in MSN] + [0.3 * u] + [x + 0 * u for x in TAIL] + [0 * u, -80 + 10 * u, 0 * u, 0 * u]
    if model == "Izhikevich2007DopSyn":
        return ([x + 0 * u for x in MSN] + [0.3 * u, 0.4 * u, 0.3 * u] + [x + 0 * u for x in TAIL]
                + [0 * u, -80 + 10 * u, 0 * u, 0 * u])
    if model == "BayesianNeuron":
        return [0.001 + 0 * u, 0.01 + 0 * u, 1.0 + u, -4 + 2 * u, 0 * u]
    if model == "InhomogeneousPoissonSource":
        return []
    raise ValueError(model)


MODELS = ["IzhikevichChaos", "IzhikevichSyn", "Izhikevich2007", "Izhikevich2007DopSyn", "BayesianNeuron",
          "InhomogeneousPoissonSource"]
NSTATE = {"IzhikevichChaos": 2, "IzhikevichSyn": 4, "Izhikevich2007": 4, "Izhikevich2007DopSyn": 4,
          "BayesianNeuron": 2, "InhomogeneousPoissonSource": 0}


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("first", [False, True])
def test_cpu_only_model_matches_oracle(model, first):
    """300 neurons of the model beside a 1000-neuron Izhikevich network (the
    model's type added first or second), 60 synapses from the network to
    each model neuron and 20 back; current stimulus on a third of the model
    neurons every step (rates in Hz for the Poisson source, held between
    stimuli), firing stimulus now and then; 250 steps, bit-exact."""
    rng = np.random.default_rng(hash(model) % 1000 + first)
    nm, ni = 300, 1000
    izh = gen.random_izhikevich(ni, 40, dmax=10, seed=9, first_index=0 if not first else nm)
    base = ni if not first else 0
    cols = model_args(model, rng, nm)
    syn_src, syn_tgt = [], []
    syn_src += list(rng.integers(0, ni, nm * 60) + (nm if first else 0))
    syn_tgt += list(np.repeat(np.arange(nm), 60) + base)
    syn_src += list(np.repeat(np.arange(nm), 20) + base)
    syn_tgt += list(rng.integers(0, ni, nm * 20) + (nm if first else 0))
    order = np.argsort(syn_src, kind="stable")
    syn_src = np.asarray(syn_src)[order]
    syn_tgt = np.asarray(syn_tgt)[order]
    delay = rng.integers(1, 11, syn_src.size)
    weight = (rng.random(syn_src.size) * 2 - 0.6).astype(np.float32)
    sims = []
    for api in (oa, nemo_amd):
        net = api.Network()
        if not first:
            izh.load(net)
        t = net.add_neuron_type(model)
        for i in range(nm):
            net.add_neuron(t, base + i, *[float(c[i]) for c in cols])
        if first:
            izh.load(net)
        for s, g, d, w in zip(syn_src, syn_tgt, delay, weight):
            net.add_synapse(int(s), int(g), int(d), float(w), 0)
        sims.append(api.Simulation(net, api.Configuration()))
    ref, hip = sims
    stim_rng = np.random.default_rng(5)
    model_spikes = 0
    for step in range(250):
        idx = stim_rng.choice(nm, nm // 3, replace=False)
        if model == "InhomogeneousPoissonSource":
            istim = [(int(base + i), float(stim_rng.random() * 200)) for i in idx] if step % 50 == 0 else None
        else:
            istim = [(int(base + i), float(stim_rng.normal(3.0, 3.0))) for i in idx]
        fstim = [int(base + i) for i in stim_rng.choice(nm, 3, replace=False)] if step % 17 == 0 else None
        fr = ref.step(fstim=fstim, istim=istim)
        fh = hip.step(fstim=fstim, istim=istim)
        assert fr == fh, step
        model_spikes += sum(1 for n in fh if base <= n < base + nm)
        for var in range(max(2, NSTATE[model])):
            assert same_f32(ref.neuron_states(var), hip.neuron_states(var)), (step, var)
    assert np.array_equal(ref.rng_state(), hip.rng_state())
    assert model_spikes > 0, "the model population never fired: the test would compare silent neurons"
