"""Pins the CPU oracle before it is trusted as the parity reference.

Each check is a known answer the oracle must reproduce: the reference's own
behavioural tests (src/test/test.cpp, src/test/stdp.cpp) restated in
scenarios.py, published known-answer values (mt19937, xorshift128), an
independent implementation (numpy's MT19937) and closed forms.
"""
import math

import numpy as np
import pytest

import oracle_api as oa
import scenarios as sc


# ------------------------------------------------------------- RNG seeding
def test_mt19937_10000th_output():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed mt19937
    assert oa.mt19937_nth(5489, 10000) == 4123659995


def test_seed_words_match_independent_mt19937():
    """RNG.cpp:38-63: four words per neuron, mt() >> 1, stream from seed 5489.
    numpy's MT19937 with legacy seeding is an independent implementation."""
    bg = np.random.MT19937()
    bg._legacy_seeding(5489)
    raw = bg.random_raw(4 * 50).astype(np.uint64)
    expected = (raw >> 1).astype(np.uint32).reshape(-1, 4)
    assert np.array_equal(oa.rng_seeds(0, 49), expected)
    # skipping 4*first words gives the same words for later neurons
    assert np.array_equal(oa.rng_seeds(17, 49), expected[17:])
    # the first eight words of mt19937(5489) shifted right by one
    assert expected.ravel()[:8].tolist() == [1749605806, 290934651, 1945173367, 1793167292,
                                              272702102, 2080627695, 1961459714, 474666992]


def test_xorshift128_known_answer():
    # Marsaglia (2003), "Xorshift RNGs", xor128 with its published seed
    state = np.array([123456789, 362436069, 521288629, 88675123], np.uint32)
    assert oa.urand(state) == 3701687786
    assert oa.urand(state) == 458299110


def test_nrand_is_sine_box_muller():
    """nrand = sin(a) sqrt(-2 ln(1-b)) with a from the first draw and b from the
    second (RNG.cpp:25-33); check against a double-precision evaluation of the
    formula to a few float ulps, and the N(0,1) moments."""
    rng = np.random.default_rng(3)
    vals = []
    for _ in range(2000):
        st = rng.integers(1, 2 ** 31, size=4, dtype=np.uint32)
        s2 = st.copy()
        u1 = oa.urand(s2)
        u2 = oa.urand(s2)
        a = np.float32(np.float32(u1) * np.float32(1.4629180792671596810513378043098e-9))
        b = np.float32(np.float32(u2) * np.float32(2.0 ** -32))
        ref = math.sin(float(a)) * math.sqrt(-2.0 * math.log(float(np.float32(1) - b)))
        got = oa.nrand(st)
        assert np.array_equal(st, s2)  # two draws consumed
        assert abs(got - ref) <= 4e-7 * max(1.0, abs(ref)), (got, ref)
        vals.append(got)
    st = np.array([1, 2, 3, 4], np.uint32)
    big = np.array([oa.nrand(st) for _ in range(100000)])
    assert abs(big.mean()) < 0.02
    assert abs(big.var() - 1.0) < 0.02


# ------------------------------------------------------------- fixed point
def test_fixed_point_conversions():
    one = 1 << 20
    assert oa.fx_tofix(1.0) == one
    assert oa.fx_tofix(-1.0) == -one
    assert oa.fx_tofix(0.5) == one // 2
    assert oa.fx_tofix(0.4 / one) == 0            # truncation toward zero
    assert oa.fx_tofix(-0.9 / one) == 0
    assert oa.fx_tofix(2047.5) == int(2047.5 * one)
    for bad in (2048.0, -2048.0, 4096.0):
        with pytest.raises(oa.OracleError):
            oa.fx_tofix(bad)
    # saturation of wide sums (fixedpoint.cpp:40-49)
    assert oa.wfx_tofloat(1 << 40) == np.float32(2 ** 31 - 1) / np.float32(one)
    assert oa.wfx_tofloat(-(1 << 40)) == np.float32(-2 ** 31) / np.float32(one)
    assert oa.wfx_tofloat(3 * one) == 3.0


# --------------------------------------------------- reference behaviour KATs
@pytest.mark.parametrize("ncount,delay", [(1000, 1), (1024, 1), (2000, 1), (4000, 1), (2000, 20), (2000, 63)])
def test_ring(ncount, delay):
    sc.run_ring(oa, ncount, delay)


def test_double_ring():
    sc.run_double_ring(oa)


def test_fixed_point_saturation():
    sc.fixed_point_saturation(oa)


def test_firing_stimulus():
    sc.firing_stimulus(oa)


def test_current_stimulus():
    sc.current_stimulus(oa)


def test_invalid_current_stimulus():
    sc.invalid_current_stimulus(oa, oa.OracleError)


@pytest.mark.parametrize("n0,nstep", [(1, 1), (1000000, 1), (1, 4), (1000000, 4)])
def test_non_contiguous_indices(n0, nstep):
    sc.non_contiguous(oa, n0, nstep, seconds=1)


@pytest.mark.parametrize("others,iz_first", [(0, True), (1, True), (1000, True), (1000, False), (2000, False)])
def test_neuron_type_mixture(others, iz_first):
    sc.neuron_type_mixture(oa, others, iz_first)


@pytest.mark.parametrize("noise,reward", [(False, 1.0), (False, 0.8), (True, 1.0), (True, 0.9)])
def test_stdp_closed_form(noise, reward):
    sc.stdp_closed_form(oa, noise, reward)


def test_stdp_configuration_validation():
    # stdp.cpp:292-338
    pre = list(np.exp(-np.arange(20) / 20.0))
    post = list(-np.exp(-np.arange(20) / 20.0))
    for args in [(-0.01, 1.0, -0.01, -1.0), (0.01, -1.0, -0.01, -1.0), (-0.01, -1.0, -0.01, -1.0),
                 (0.01, 1.0, 0.01, -1.0), (0.01, 1.0, -0.01, 1.0), (0.01, 1.0, 0.01, 1.0),
                 (1.0, 0.01, -0.01, -1.0), (0.01, 1.0, -1.0, -0.01)]:
        with pytest.raises(oa.OracleError):
            oa.Configuration().set_stdp_function(pre, post, *args)
    long_pre = list(np.exp(-np.arange(35) / 35.0))
    with pytest.raises(oa.OracleError):
        oa.Configuration().set_stdp_function(long_pre, long_pre, 0.01, 1.0, -0.01, -1.0)


@pytest.mark.parametrize("stdp", [False, True])
def test_stdp_dynamic_window_length(stdp):
    # stdp.cpp:323-355: delay 34 + postfire 31 > 64 is rejected only with STDP
    conf = oa.Configuration()
    if stdp:
        fn = list(np.exp(-np.arange(31) / 31.0))
        conf.set_stdp_function(fn, fn, 0.01, 1.0, -0.01, -1.0)
    net = oa.Network()
    iz = net.add_neuron_type("Izhikevich")
    net.add_neuron(iz, 0, *([0.0] * 7))
    net.add_neuron(iz, 1, *([0.0] * 7))
    net.add_synapse(0, 1, 34, 1.0, 1 if stdp else 0)
    if stdp:
        with pytest.raises(oa.OracleError):
            oa.Simulation(net, conf)
    else:
        oa.Simulation(net, conf)


def test_apply_stdp_without_stdp_fails():
    # stdp.cpp:279-286
    sim = sc.make(oa, sc.gen.ring(10, plastic=True))
    sim.step()
    with pytest.raises(oa.OracleError):
        sim.apply_stdp(1.0)


def test_delay_limits():
    net = oa.Network()
    iz = net.add_neuron_type("Izhikevich")
    net.add_neuron(iz, [0, 1], *([0.0] * 7))
    with pytest.raises(oa.OracleError):
        net.add_synapse(0, 1, 0, 1.0, 0)        # delay < 1 (NetworkImpl.cpp:98-101)
    net.add_synapse(0, 1, 65, 1.0, 0)
    with pytest.raises(oa.OracleError):
        oa.Simulation(net, oa.Configuration())  # > 64 ms (cpu/Simulation.cpp:52-56)


def test_invalid_synapse_target():
    net = oa.Network()
    iz = net.add_neuron_type("Izhikevich")
    net.add_neuron(iz, [0, 1], *([0.0] * 7))
    net.add_synapse(0, 7, 1, 1.0, 0)
    with pytest.raises(oa.OracleError):
        oa.Simulation(net, oa.Configuration())


def test_empty_network_rejected():
    with pytest.raises(oa.OracleError):
        oa.Simulation(oa.Network(), oa.Configuration())


# ------------------------------------------------------ Izhikevich numerics
def _izh_fp64_reference(a, b, c, d, u, v, currents):
    """Independent numpy evaluation of NeMo's Izhikevich step: 4 Euler
    substeps of 0.25 ms, v in double precision, u in single, no update after
    firing (cpu/plugins/Izhikevich.cpp:67-94)."""
    a, b, c, d = (np.float32(x) for x in (a, b, c, d))
    u, v = np.float32(u), np.float32(v)
    fired_at = []
    for t, I in enumerate(currents):
        I = np.float32(I)
        fired = False
        for _ in range(4):
            if not fired:
                vd = np.float64(v)
                v = np.float32(vd + np.float64(np.float32(0.25)) * (((0.04 * vd + 5.0) * vd + 140.0) - np.float64(u)
                                                                    + np.float64(I)))
                u = np.float32(u + np.float32(0.25) * (a * (b * v - u)))
                fired = bool(v >= np.float32(30.0))
        if fired:
            fired_at.append(t)
            v = c
            u = np.float32(u + d)
    return fired_at, u, v


def _single_neuron(I):
    """Regular-spiking neuron (a=0.02, b=0.2, c=-65, d=8) under constant
    current: oracle spike times and final state equal the independent
    recomputation exactly; firing rate grows with current and is zero at rest."""
    steps = 1000
    net = oa.Network()
    iz = net.add_neuron_type("Izhikevich")
    net.add_neuron(iz, 0, 0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0)
    sim = oa.Simulation(net, oa.Configuration())
    fired = [t for t in range(steps) if sim.step(istim=[(0, I)])]
    ref, u, v = _izh_fp64_reference(0.02, 0.2, -65.0, 8.0, -13.0, -65.0, [I] * steps)
    assert fired == ref
    assert sim.get_neuron_state(0, 0) == u and sim.get_neuron_state(0, 1) == v
    if I == 0.0:
        assert fired == []
    if I == 10.0:
        assert 5 <= len(fired) <= 40  # tonic spiking, tens of Hz
    return len(fired)


@pytest.mark.parametrize("I", [0.0, 4.0, 10.0, 20.0])
def test_izhikevich_single_neuron(I):
    _single_neuron(I)


def test_izhikevich_rate_monotone_in_current():
    counts = [_single_neuron(I) for I in (4.0, 10.0, 20.0)]
    assert counts[0] <= counts[1] < counts[2]


# ------------------------------------------------ random-network invariants
def test_random_network_events_match_raster():
    """Delivered events equal the out-degree sum implied by the raster: every
    firing of n at step s delivers row (n, d) at step s+d (cpu/Simulation.cpp:259-277)."""
    spec = sc.gen.random_izhikevich(1000, 100, dmax=20, seed=1)
    sim = sc.make(oa, spec)
    steps = 300
    counts, fired, events = sim.run(steps)
    n, D = 1000, 20
    rowlen = np.zeros((n, D + 1), np.int64)
    np.add.at(rowlen, (spec.source, spec.delay), 1)
    step_of = np.repeat(np.arange(steps), counts)
    expected = sum(int(rowlen[f, d]) for f, s in zip(fired, step_of) for d in range(1, D + 1) if s + d < steps)
    assert events == expected
    assert counts.sum() > 0


def test_propagate_brute_force():
    """Brian's propagate (cpu/Simulation.cpp:115-137): call j makes the given
    set F_j fire and returns the input of the step after it, so neuron t
    receives the sum of w over synapses (s, t, d, w) with s in F_{j+1-d},
    excitatory and inhibitory apart.  Checked against that sum computed by
    brute force on a small random network with delays up to 5, and the
    neuron state is left alone."""
    rng = np.random.default_rng(12)
    n, m, D = 40, 6, 5
    net = oa.Network()
    t = net.add_neuron_type("Izhikevich")
    net.add_neuron(t, np.arange(n, dtype=np.uint32), 0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0)
    src = np.repeat(np.arange(n), m)
    tgt = rng.integers(0, n, src.size)
    dly = rng.integers(1, D + 1, src.size)
    w = np.round(rng.uniform(-2, 2, src.size) * 1024) / 1024  # exact in Q11.20
    net.add_synapse(src.astype(np.uint32), tgt.astype(np.uint32), dly.astype(np.uint32), w.astype(np.float32), 0)
    sim = oa.Simulation(net, oa.Configuration())
    v0 = [sim.get_neuron_state(i, 1) for i in range(n)]
    fired = []
    for j in range(30):
        F = sorted(rng.choice(n, rng.integers(0, 8), replace=False).tolist())
        fired.append(set(F))
        E, I = sim.propagate(F)
        eE, eI = np.zeros(n), np.zeros(n)
        for s, tt, d, ww in zip(src, tgt, dly, w):
            k = j + 1 - d
            if k >= 0 and s in fired[k]:
                if ww >= 0:
                    eE[tt] += ww
                else:
                    eI[tt] += ww
        assert np.array_equal(E, eE.astype(np.float32)), j
        assert np.array_equal(I, eI.astype(np.float32)), j
    assert [sim.get_neuron_state(i, 1) for i in range(n)] == v0
    assert sim.elapsed_simulation() == 30
