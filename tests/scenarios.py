"""NeMo's behavioural test procedures, written once for any backend module
exposing Network / Configuration / Simulation (nemo_amd or the oracle).

Each function restates a reference test (src/test/*.cpp, cited per
function) and asserts its known answer.
"""
import numpy as np

from nemo_amd import generators as gen


def make(api, spec, conf=None):
    net = api.Network()
    spec.load(net)
    if conf is None:
        conf = api.Configuration()
    return api.Simulation(net, conf)


# src/test/test.cpp:303-329 runRing
def run_ring(api, ncount, delay):
    duration = ncount * 5 // 2
    sim = make(api, gen.ring(ncount, delay=delay))
    sim.step(fstim=[0])
    for ms in range(1, duration):
        fired = sim.step()
        if delay == 1:
            assert fired == [ms % ncount], (ms, fired)
        elif ms % delay == 0:
            assert fired == [(ms // delay) % ncount], (ms, fired)
        else:
            assert fired == [], (ms, fired)


# src/test/test.cpp:333-367 runDoubleRing
def run_double_ring(api):
    ncount = 512
    duration = ncount * 5 // 2
    spec = gen.merge(gen.ring(ncount, 0, delay=1), gen.ring(ncount, ncount, delay=2))
    sim = make(api, spec)
    sim.step(fstim=[0, ncount])
    for ms in range(1, duration):
        fired = sim.step()
        if ms % 2 == 0:
            assert fired == [ms % ncount, ncount + (ms // 2) % ncount], (ms, fired)
        else:
            assert fired == [ms % ncount], (ms, fired)


# src/test/test.cpp:418-470 testFixedPointSaturation
def fixed_point_saturation(api):
    INPUT0, INPUT1, OUTPUT_SAT, OUTPUT_REF, OUTPUT_WEAK, NCOUNT = range(6)
    net = api.Network()
    iz = net.add_neuron_type("Izhikevich")
    params = [0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0]
    for n in range(NCOUNT):
        net.add_neuron(iz, n, *params)
    for src, tgt, w in [(INPUT0, OUTPUT_REF, 1024.0), (INPUT1, OUTPUT_REF, 1024.0),
                        (INPUT0, OUTPUT_WEAK, 1000.0), (INPUT1, OUTPUT_WEAK, 1000.0),
                        (INPUT0, OUTPUT_SAT, 1100.0), (INPUT1, OUTPUT_SAT, 1100.0)]:
        net.add_synapse(src, tgt, 1, w, 0)
    sim = api.Simulation(net, api.Configuration())
    assert len(sim.step(fstim=[INPUT0, INPUT1])) == 2
    assert len(sim.step()) == 3
    u_ref = sim.get_neuron_state(OUTPUT_REF, 0)
    u_weak = sim.get_neuron_state(OUTPUT_WEAK, 0)
    u_sat = sim.get_neuron_state(OUTPUT_SAT, 0)
    eps = 0.00001
    assert abs(u_ref - u_weak) > eps
    assert abs(u_ref - u_sat) < eps
    return u_ref, u_weak, u_sat


# src/test/test.cpp:209-245 testFiringStimulus
def firing_stimulus(api, ncount=3000, cycles=300, seed=7):
    net = api.Network()
    iz = net.add_neuron_type("Izhikevich")
    r = 0.5
    b = 0.25 - 0.05 * r
    v = -65.0
    idx = np.arange(ncount, dtype=np.uint32)
    net.add_neuron(iz, idx, 0.02 + 0.08 * r, b, v, 2.0, 0.0, b * v, v)  # a, b, c, d, sigma, u, v
    sim = api.Simulation(net, api.Configuration())
    rng = np.random.default_rng(seed)
    p_fire = 10.0 / ncount
    for _ in range(cycles):
        fstim = sorted(np.nonzero(rng.random(ncount) < p_fire)[0].tolist())
        assert sim.step(fstim=fstim) == fstim


# src/test/test.cpp:249-274 testCurrentStimulus
def current_stimulus(api):
    ncount = 1500
    sim = make(api, gen.ring(ncount))
    sim.step(istim=[(5, 0.001), (8, 0.001), (0, 1000.0), (100, 0.001), (1, 0.001)])
    for ms in range(1, ncount * 2):
        assert sim.step() == [ms % ncount]


# src/test/test.cpp:277-291 testInvalidCurrentStimulus
def invalid_current_stimulus(api, error):
    ncount = 1000
    sim = make(api, gen.ring(ncount))
    sim.step()
    sim.step()
    try:
        sim.step(istim=[(ncount + 1, 0.5)])
    except error:
        return
    raise AssertionError("invalid current stimulus accepted")


# src/test/test.cpp:523-560 testNonContigousNeuronIndices
def non_contiguous(api, n0, nstep, seconds=2):
    ncount = 1000
    out = []
    for base in (0, n0):
        sim = make(api, gen.ring(ncount, base, nstep=nstep))
        fired = []
        for s in range(seconds * 1000):
            fired.extend(sim.step(fstim=[base] if s == 0 else None))
        out.append(np.asarray(fired, dtype=np.int64))
    assert out[0].size == seconds * ncount
    assert out[1].size == seconds * ncount
    assert np.array_equal(out[0], out[1] - n0)


# src/test/test.cpp:380-413 testNeuronTypeMixture
def neuron_type_mixture(api, sz_others, iz_first):
    sz_ring = 1024
    poisson_idx = np.arange(sz_ring, sz_ring + sz_others, dtype=np.uint32)
    poisson = gen.NetworkSpec(blocks=[gen.NeuronBlock("PoissonSource", poisson_idx,
                                                      [np.full(sz_others, 0.001, np.float32)])])
    r = gen.ring(sz_ring)
    spec = gen.merge(r, poisson) if iz_first else gen.merge(poisson, r)
    sim = make(api, spec)
    sim.step(fstim=[0])
    for ms in range(1, 1000):
        fired = sorted(sim.step())
        assert len(fired) > 0
        assert fired[0] == ms % sz_ring
        if len(fired) > 1:
            assert fired[1] >= sz_ring


# ---------------------------------------------------------------- STDP
# src/test/stdp.cpp:20-265
GROUP = 2 * 768
INIT_W = 5.0
PRE_FIRE = 10
POST_FIRE_DELAY = 10


def dw_pre(dt):
    return np.float32(1.0) * np.exp(np.float32(dt) / np.float32(20.0), dtype=np.float32)


def dw_post(dt):
    return np.float32(-0.8) * np.exp(np.float32(-dt) / np.float32(20.0), dtype=np.float32)


def stdp_configuration(api):
    conf = api.Configuration()
    pre = [dw_pre(-i) for i in range(20)]
    post = [dw_post(i) for i in range(20)]
    conf.set_stdp_function(pre, post, -0.5, 2 * INIT_W)
    return conf


def stdp_network(noise):
    """Group 0 -> group 1 one-to-one plastic synapses with delay 1 + (local % 20);
    optional all-to-all noise synapses of weight -1e-5, plastic when the
    target's local index is odd (stdp.cpp:120-170)."""
    r = 0.5
    b = 0.25 - 0.05 * r
    v = -65.0
    n = 2 * GROUP
    idx = np.arange(n, dtype=np.uint32)
    one = np.ones(n, np.float32)
    block = gen.NeuronBlock("Izhikevich", idx, [np.float32(0.02 + 0.08 * r) * one, np.float32(b) * one,
                                                np.float32(v) * one, np.float32(2.0) * one, 0 * one,
                                                np.float32(b * v) * one, np.float32(v) * one])
    spec = gen.NetworkSpec(blocks=[block])
    local = np.arange(GROUP, dtype=np.uint32)
    src = [local]
    tgt = [GROUP + local]
    dly = [1 + local % 20]
    w = [np.full(GROUP, INIT_W, np.float32)]
    p = [np.ones(GROUP, np.uint8)]
    if noise:
        ls, lt = np.meshgrid(local, local, indexing="ij")
        keep = ls != lt
        ls, lt = ls[keep], lt[keep]
        src.append(ls)
        tgt.append(GROUP + lt)
        dly.append(1 + (lt + ls) % 20)
        w.append(np.full(ls.size, -0.00001, np.float32))
        p.append((lt & 1).astype(np.uint8))
    # synapses of one source must stay together in insertion order: sort by source, stable
    src = np.concatenate(src)
    order = np.argsort(src, kind="stable")
    spec.source = src[order].astype(np.uint32)
    spec.target = np.concatenate(tgt)[order].astype(np.uint32)
    spec.delay = np.concatenate(dly)[order].astype(np.uint32)
    spec.weight = np.concatenate(w)[order]
    spec.plastic = np.concatenate(p)[order]
    return spec


def stdp_expected(epoch, reward):
    local = np.arange(GROUP)
    delay = 1 + local % 20
    dt = -(POST_FIRE_DELAY - delay)
    exp = np.where(dt > 0, [dw_post(x - 1) for x in dt], [dw_pre(x) for x in dt]).astype(np.float64)
    return INIT_W + epoch * reward * exp


def stdp_closed_form(api, noise, reward, nepochs=4):
    """Returns the one-to-one weights after each epoch, checked against
    w = 5 + epoch * reward * dw(dt) within 0.001 % (stdp.cpp:183-218)."""
    spec = stdp_network(noise)
    sim = make(api, spec, stdp_configuration(api))
    # one-to-one synapse of source i is its first synapse
    first = np.zeros(GROUP, dtype=np.int64)
    counts = np.bincount(spec.source, minlength=2 * GROUP)
    starts = np.concatenate([[0], np.cumsum(counts)])[:GROUP]
    first[:] = starts
    nsyn = spec.synapse_count
    history = []
    for epoch in range(1, nepochs + 1):
        for ms in range(100):
            if ms == PRE_FIRE:
                f = list(range(GROUP))
            elif ms == PRE_FIRE + POST_FIRE_DELAY:
                f = list(range(GROUP, 2 * GROUP))
            else:
                f = None
            sim.step(fstim=f)
        sim.apply_stdp(reward)
        w = np.asarray(sim.fixed_weights(nsyn) if _takes_count(sim) else sim.fixed_weights())[first]
        wf = w.astype(np.float64) / (1 << 20)
        expected = stdp_expected(epoch, reward)
        assert np.allclose(wf, expected, rtol=1e-5, atol=0), (epoch, np.max(np.abs(wf - expected) / expected))
        history.append(w.copy())
    # spot-check the query API path on a few synapses
    for i in (0, 7, GROUP - 1):
        ids = sim.get_synapses_from(i)
        sid = ids[0]
        assert sim.get_synapse_target(sid) == GROUP + i
        assert sim.get_synapse_delay(sid) == 1 + i % 20
        assert sim.get_synapse_plastic(sid)
        assert abs(sim.get_synapse_weight(sid) - stdp_expected(nepochs, reward)[i]) <= 1e-5 * abs(
            stdp_expected(nepochs, reward)[i])
    return history


def _takes_count(sim):
    import inspect
    try:
        return len(inspect.signature(sim.fixed_weights).parameters) == 1
    except (TypeError, ValueError):
        return False


# src/test/test.cpp:775-968 testSetNeuron: neuron setters and getters of the
# network and of a running simulation (Izhikevich: parameters a, b, c, d,
# sigma; state u, v), with the reference's float arithmetic
def set_neuron(api, error):
    import pytest
    f = np.float32
    a, b = f(0.02), f(0.2)
    c = f(-65.0) + f(15.0) * f(0.25)
    d = f(8.0) - f(6.0) * f(0.25)
    v = f(-65.0)
    u = b * v
    sigma = f(5.0)
    e = f(0.1)

    def iz(a_, b_, c_, d_, u_, v_, s_):
        """the deprecated Izhikevich argument order -> parameters, state"""
        return [a_, b_, c_, d_, s_, u_, v_]

    net = api.Network()
    with pytest.raises(error):
        net.set_neuron(0, *iz(a, b, c, d, u, v, sigma))
    t = net.add_neuron_type("Izhikevich")
    net.add_neuron(t, 0, *iz(a, b, c - f(0.1), d, u, v - f(1.0), sigma))
    for bad in (lambda: net.get_neuron_parameter(1, 0), lambda: net.get_neuron_state(1, 0),
                lambda: net.get_neuron_parameter(0, 5), lambda: net.get_neuron_state(0, 2)):
        with pytest.raises(error):
            bad()
    net.set_neuron(0, *iz(a - e, b - e, c - e, d - e, u - e, v - e, sigma - e))
    assert [net.get_neuron_parameter(0, p) for p in range(5)] == [a - e, b - e, c - e, d - e, sigma - e]
    assert [net.get_neuron_state(0, s) for s in range(2)] == [u - e, v - e]
    for p, x in enumerate((a, b, c, d, sigma)):
        net.set_neuron_parameter(0, p, x)
        assert net.get_neuron_parameter(0, p) == x
    for s, x in enumerate((u, v)):
        net.set_neuron_state(0, s, x)
        assert net.get_neuron_state(0, s) == x
    for bad in (lambda: net.set_neuron_parameter(1, 0, 0.0), lambda: net.set_neuron_state(1, 0, 0.0),
                lambda: net.set_neuron_parameter(0, 5, 0.0), lambda: net.set_neuron_state(0, 2, 0.0)):
        with pytest.raises(error):
            bad()

    conf = api.Configuration()
    sim = api.Simulation(net, conf)
    sim.step()
    sim.set_neuron_state(0, 0, u - e)
    assert sim.get_neuron_state(0, 0) == u - e
    sim.set_neuron_state(0, 1, v - e)
    assert sim.get_neuron_state(0, 1) == v - e
    u0, v0 = sim.get_neuron_state(0, 0), sim.get_neuron_state(0, 1)
    for p, x in enumerate((a - e, b - e, c - e, d - e, sigma - e)):
        sim.set_neuron_parameter(0, p, x)
        assert sim.get_neuron_parameter(0, p) == x
    sim.step()
    assert [sim.get_neuron_parameter(0, p) for p in range(5)] == [a - e, b - e, c - e, d - e, sigma - e]
    assert sim.get_neuron_state(0, 0) != u0
    assert sim.get_neuron_state(0, 1) != v0
    for bad in (lambda: sim.set_neuron_parameter(1, 0, 0.0), lambda: sim.set_neuron_state(1, 0, 0.0),
                lambda: sim.get_neuron_parameter(1, 0), lambda: sim.get_neuron_state(1, 0),
                lambda: sim.set_neuron_parameter(0, 5, 0.0), lambda: sim.set_neuron_state(0, 2, 0.0),
                lambda: sim.get_neuron_parameter(0, 5), lambda: sim.get_neuron_state(0, 2)):
        with pytest.raises(error):
            bad()
    del sim

    sim = api.Simulation(net, conf)
    sim.step()
    v_ref = sim.get_membrane_potential(0)
    del sim

    # changing c, which only matters when the neuron fires, leaves the step alone
    sim = api.Simulation(net, conf)
    assert sim.get_neuron_state(0, 0) == u and sim.get_neuron_state(0, 1) == v
    sim.set_neuron(0, *iz(a, b, c + f(1.0), d, u, v, sigma))
    sim.step()
    assert sim.get_membrane_potential(0) == v_ref
    assert [sim.get_neuron_parameter(0, p) for p in range(5)] == [a, b, c + f(1.0), d, sigma]
    del sim

    # setting the state is not applied more than once
    net0 = api.Network()
    t0 = net0.add_neuron_type("Izhikevich")
    net0.add_neuron(t0, 0, *iz(a, b, c, d, u, v, f(0.0)))
    sim0 = api.Simulation(net0, conf)
    sim0.step()
    sim0.step()
    sim1 = api.Simulation(net0, conf)
    sim1.step()
    sim1.set_neuron(0, *iz(a, b, c, d, u, v, f(0.0)))
    sim1.step()
    sim1.step()
    assert sim0.get_membrane_potential(0) == sim1.get_membrane_potential(0)

    # the membrane potential set after construction is the one simulated
    net1 = api.Network()
    t1 = net1.add_neuron_type("Izhikevich")
    net1.add_neuron(t1, 0, *iz(a, b, c, d, u, v - f(1.0), sigma))
    sim = api.Simulation(net1, conf)
    sim.set_neuron(0, *iz(a, b, c, d, u, v, sigma))
    assert sim.get_membrane_potential(0) == v
    sim.step()
    assert sim.get_membrane_potential(0) == v_ref


# src/test/PoissonSource.cpp:7-30 testRate: the mean rate of one Poisson
# source over a long run approaches its rate parameter (crude bound of the
# reference), alone or beside a ring of neurons that never fire
def poisson_rate(api, duration, other_neurons):
    net = api.Network()
    if other_neurons:
        gen.ring(1024, 0, delay=1).load(net)
        idx = 1024
    else:
        idx = 0
    t = net.add_neuron_type("PoissonSource")
    rate = np.float32(0.010)
    net.add_neuron(t, idx, rate)
    sim = api.Simulation(net, api.Configuration())
    nfired = 0
    for _ in range(duration):
        nfired += len(sim.step())
    expected = int(abs(np.float32(duration) * rate))
    deviation = (nfired - expected) % (1 << 32)  # the reference's unsigned difference
    assert nfired > 0
    assert deviation < expected * 2
    return nfired


# src/test/Input.cpp:16-65: input neurons (no parameters, no state) fire
# exactly as stimulated, alone or followed by as many Izhikevich neurons
def input_neurons(api, ncount, izhikevich, seed=9):
    net = api.Network()
    t = net.add_neuron_type("Input")
    net.add_neuron(t, np.arange(ncount, dtype=np.uint32))
    if izhikevich:
        iz = net.add_neuron_type("Izhikevich")
        net.add_neuron(iz, np.arange(ncount, 2 * ncount, dtype=np.uint32),
                       0.02, 0.2, -65.0, 8.0, 0.0, np.float32(0.2) * np.float32(-65.0), -65.0)
    sim = api.Simulation(net, api.Configuration())
    rng = np.random.default_rng(seed)
    for _ in range(1000):
        p = 0.5 * rng.random()
        fstim = [n for n in range(ncount) if rng.random() < p]
        fired = sim.step(fstim=fstim)
        assert list(fired) == fstim
