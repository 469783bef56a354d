"""Kuramoto oscillators (the all-synapse reverse matrix and a model with a
state history) on the HIP backend, through the C ABI, against the oracle's
restatement of cpu/plugins/Kuramoto.cpp: phases bit-exact every step."""
import math

import numpy as np
import pytest

import nemo_amd
from nemo_amd import generators as gen
import oracle_api as oa
import test_oracle_model_pins as pins

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu(require_gpu):
    pass


def same_f32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    an, bn = np.isnan(a), np.isnan(b)
    return np.array_equal(an, bn) and np.array_equal(a.view(np.uint32)[~an], b.view(np.uint32)[~bn])


def oscillators(n, indegree, dmax, seed, cmean, base=0):
    """n oscillators (indices base..base+n-1), each with `indegree` incoming
    synapses from random oscillators, delays U{1..dmax}, weights U[-1, 1);
    omega U[0, 1), theta U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    freq = rng.random(n, dtype=np.float32)
    phase = (rng.random(n) * 2 * math.pi).astype(np.float32)
    tgt = np.repeat(np.arange(n), indegree)
    src = rng.integers(0, n, tgt.size)
    d = rng.integers(1, dmax + 1, tgt.size)
    w = (rng.random(tgt.size, dtype=np.float32) * 2 - 1).astype(np.float32)
    order = np.argsort(src, kind="stable")  # synapses added source by source, in random target order
    return freq, phase, [(int(base + src[i]), int(base + tgt[i]), int(d[i]), float(w[i])) for i in order], cmean


def build(api, spec, izh=None):
    freq, phase, syn, cmean = spec
    net = api.Network()
    if izh is not None:
        izh.load(net)
    t = net.add_neuron_type("Kuramoto")
    base = 0 if izh is None else int(izh.neuron_count)
    for i, (f, p) in enumerate(zip(freq, phase)):
        net.add_neuron(t, base + i, float(f), cmean, float(p))
    for s, g, d, w in syn:
        net.add_synapse(s, g, d, w, 0)
    return api.Simulation(net, api.Configuration())


@pytest.mark.parametrize("n,indegree,dmax,cmean", [(600, 40, 20, 0.0), (300, 100, 40, 100.0), (257, 3, 1, -1.0),
                                                    (400, 60, 64, 0.0)])
def test_kuramoto_network_matches_oracle(n, indegree, dmax, cmean):
    """Random oscillator networks: phases bit-exact with the oracle after
    every step for 120 steps, then after a 200-step run.  Delays beyond the
    32-step history wrap around it, as in the reference; delays 32 and 64
    read the slot the step writes, which the CPU backend's loop has already
    rewritten for lower-index sources (the update passes)."""
    spec = oscillators(n, indegree, dmax, seed=n + indegree, cmean=cmean)
    ref, hip = build(oa, spec), build(nemo_amd, spec)
    assert same_f32(ref.neuron_states(0), hip.neuron_states(0))
    for t in range(120):
        assert ref.step() == hip.step() == []
        assert same_f32(ref.neuron_states(0), hip.neuron_states(0)), t
    for _ in range(200):
        ref.step()
    hip.run(200)
    assert same_f32(ref.neuron_states(0), hip.neuron_states(0))


def test_kuramoto_after_other_types_and_timer_reset():
    """Oscillators added after an Izhikevich network (their device indices
    start past it); the timer reset at step 50 restarts the cycle that
    selects the history slots (the reference reads slot cycle % 32 even
    though its current state is elsewhere), and a phase set between steps
    lands in the current slot: all bit-exact with the oracle, oscillators
    and Izhikevich neurons alike."""
    izh = gen.random_izhikevich(700, 50, dmax=20, seed=61)
    spec = oscillators(400, 30, 25, seed=62, cmean=30.0, base=700)
    ref, hip = build(oa, spec, izh), build(nemo_amd, spec, izh)
    for t in range(150):
        if t == 50:
            ref.reset_timer()
            hip.reset_timer()
        if t == 80:
            for sim in (ref, hip):
                sim.set_neuron_state(900, 0, 1.25)
        assert ref.step() == hip.step(), t
        for var in (0, 1):
            assert same_f32(ref.neuron_states(var), hip.neuron_states(var)), (t, var)


def test_kuramoto_kats_on_hip():
    """The double-precision Kuramoto pins of the oracle (uncoupled advance by
    h omega, delayed coupling, N-to-1 with Cmean) on the HIP backend."""
    freq, phase = [0.1, 1.0, 7.0, 0.0], [0.0, 3.0, 6.0, 1.0]
    sim = pins.kuramoto_net(nemo_amd, freq, phase)
    got = []
    for _ in range(500):
        sim.step()
        got.append([sim.get_neuron_state(i, 0) for i in range(4)])
    assert np.max(pins.wrapped_diff(got, pins.kuramoto_fp64(freq, phase, [], 500))) < 2e-4
    for delay in (1, 3, 20):
        syn = [(0, 1, delay, 1.0)]
        sim = pins.kuramoto_net(nemo_amd, [0.1, 0.1], [0.0, 1.57], synapses=syn)
        got = []
        for _ in range(200):
            sim.step()
            got.append([sim.get_neuron_state(i, 0) for i in range(2)])
        ref = pins.kuramoto_fp64([0.1, 0.1], [0.0, 1.57], syn, 200)
        assert np.max(pins.wrapped_diff(got, ref)) < 1e-3, delay


def test_kuramoto_sine_range_reported():
    """A coupling so strong that a sine argument leaves (-120, 120), where
    the restated sinf does not follow the C library: the step fails with an
    error instead of returning other phases than the reference."""
    spec = ([0.1, 0.1], [0.0, 3.0], [(0, 1, 1, 2000.0)] * 20 + [(1, 0, 1, 2000.0)] * 20, 0.0)
    sim = build(nemo_amd, spec)
    with pytest.raises(nemo_amd.NemoError):
        for _ in range(5):
            sim.step()


def test_kuramoto_refusals():
    """STDP beside a model that reads the reverse matrix (construction/
    RCM.cpp:83-91), and a non-oscillator source of an oscillator, are
    refused (NEMO_API_UNSUPPORTED; a failed construction returns no status
    through the C API, so the message is checked)."""
    net = nemo_amd.Network()
    t = net.add_neuron_type("Kuramoto")
    net.add_neuron(t, 0, 0.1, 0.0, 0.0)
    net.add_neuron(t, 1, 0.1, 0.0, 0.0)
    net.add_synapse(0, 1, 1, 1.0, 1)
    conf = nemo_amd.Configuration()
    pre, post = gen.stdp_exp_function(0.1, -0.12, 20.0, 20)
    conf.set_stdp_function(pre, post, 0.0, 10.0, 0.0, -10.0)
    with pytest.raises(nemo_amd.NemoError) as e:
        nemo_amd.Simulation(net, conf)
    assert "reverse connectivity" in str(e.value)
    izh = gen.random_izhikevich(10, 0, seed=1)
    net = nemo_amd.Network()
    izh.load(net)
    t = net.add_neuron_type("Kuramoto")
    net.add_neuron(t, 10, 0.1, 0.0, 0.0)
    net.add_synapse(3, 10, 1, 1.0, 0)
    with pytest.raises(nemo_amd.NemoError) as e:
        nemo_amd.Simulation(net, nemo_amd.Configuration())
    assert "sources must be of the same type" in str(e.value)


def test_kuramoto_lag_network_matches_oracle():
    """KuramotoLag (a phase lag alpha per oscillator, cpu/plugins/
    KuramotoLag.cpp): 350 oscillators, 50 incoming synapses each, delays
    1..64, lags U[-1, 1): phases bit-exact with the oracle every step for 100
    steps."""
    rng = np.random.default_rng(77)
    freq, phase, syn, cmean = oscillators(350, 50, 64, seed=77, cmean=50.0)
    alpha = (rng.random(350) * 2 - 1).astype(np.float32)
    sims = []
    for api in (oa, nemo_amd):
        net = api.Network()
        t = net.add_neuron_type("KuramotoLag")
        for i in range(350):
            net.add_neuron(t, i, float(freq[i]), cmean, float(alpha[i]), float(phase[i]))
        for s, g, d, w in syn:
            net.add_synapse(s, g, d, w, 0)
        sims.append(api.Simulation(net, api.Configuration()))
    ref, hip = sims
    for t in range(100):
        ref.step()
        hip.step()
        assert same_f32(ref.neuron_states(0), hip.neuron_states(0)), t
