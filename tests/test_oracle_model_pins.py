"""Pins the oracle's QIF, IF_curr_exp, HH, PoissonSource and Input models.

The oracle restates the reference CPU plugins (src/nemo/cpu/plugins/*.cpp);
these checks do not retype its formulas.  They hold it to what the
mathematics fixes -- closed-form and analytic solutions, an independent
double-precision integration of the textbook Hodgkin-Huxley equations,
fixed points, thresholds, a brute-force spike train from the pinned
xorshift128 stream -- and to the reference's own tests for these models
(src/test/PoissonSource.cpp, Input.cpp, IF_curr_exp.cpp).  Each check names
in its docstring the plausible mistake it catches.
"""
import math

import numpy as np
import pytest

from nemo_amd import generators as gen
import oracle_api as oa
import scenarios as sc

f32 = np.float32


def single_neuron(model, args, steps, istim=0.0, fstim=()):
    """One neuron of `model` (parameters then state in args) stepped `steps`
    times with an external current istim every step and forced firing at
    the steps in fstim; returns its state after every step and the steps it
    fired."""
    net = oa.Network()
    t = net.add_neuron_type(model)
    net.add_neuron(t, 0, *args)
    sim = oa.Simulation(net, oa.Configuration())
    nstate = {"QIF": 1, "IF_curr_exp": 4, "HH": 5}[model]
    states, fired = [], []
    for s in range(steps):
        f = sim.step(fstim=[0] if s in fstim else None, istim=[(0, istim)] if istim else None)
        if f:
            fired.append(s)
        states.append([sim.get_neuron_state(0, v) for v in range(nstate)])
    return np.array(states, np.float64), fired


# ------------------------------------------------------------------- QIF
def qif_analytic(a, I, v0, t):
    """v' = a v (v - 1) + I' with constant I' < a/4 (Riccati equation with two
    real equilibria r1 > r2): (v - r1) / (v - r2) = C exp(a (r1 - r2) t)."""
    disc = math.sqrt(a * a - 4 * a * I)
    r1, r2 = (a + disc) / (2 * a), (a - disc) / (2 * a)
    k = (v0 - r1) / (v0 - r2) * math.exp(a * (r1 - r2) * t)
    return (r1 - k * r2) / (1 - k)


@pytest.mark.parametrize("I", [0.0, 10.0])
def test_qif_follows_the_analytic_solution(I):
    """Below threshold, one QIF step integrates dv/dt = a v (v - 1) + 0.015 I
    over one time unit (ten RK4 substeps of 0.1): v after every step lies
    within RK4 error (1e-5) of the analytic solution.  Catches a wrong input
    scale (e.g. 0.15 instead of 0.015), a wrong substep count (nine substeps
    of 1/dt rounded down would be 2e-2 off), or a dropped RK4 stage."""
    a = 2.0
    states, fired = single_neuron("QIF", [a, 0.5], 30, istim=I)
    assert fired == []
    exact = [qif_analytic(a, 0.015 * float(f32(I)), 0.5, s + 1) for s in range(30)]
    assert np.max(np.abs(states[:, 0] - exact)) < 1e-5


def test_qif_fixed_points_and_reset():
    """With no input, v = 0 is a fixed point of a v (v - 1) (it stays exactly
    0); v = 1 is the other equilibrium but also the threshold (v >= 1), so a
    neuron started there fires in the first step and resets to 0.  Catches a
    strict threshold (v > 1) or a reset other than 0."""
    states, fired = single_neuron("QIF", [2.0, 0.0], 20)
    assert fired == [] and np.all(states[:, 0] == 0.0)
    states, fired = single_neuron("QIF", [2.0, 1.0], 3)
    assert fired == [0] and states[0, 0] == 0.0


# ----------------------------------------------------------- IF_curr_exp
# parameter order of plugins/IF_curr_exp.h:4-18
def lif(v_rest=-65.0, v_reset=-65.0, cm=1.0, tau_m=20.0, tau_refrac=2.0, tau_e=5.0, tau_i=5.0, v_thresh=-50.0,
        i_offset=0.0, v=-65.0, ie=0.0, ii=0.0, lastfired=100.0):
    return [v_rest, v_reset, cm, tau_m, tau_refrac, tau_e, tau_i, v_thresh, i_offset, v, ie, ii, lastfired]


def test_lif_subthreshold_geometric_recurrence():
    """Below threshold with a constant input I (offset current), v obeys
    v' = v + I/cm + (v_rest - v)/tau_m, whose solution is the geometric
    approach v_t = v_inf + (v_0 - v_inf)(1 - 1/tau_m)^t to
    v_inf = v_rest + I tau_m / cm.  cm = 2 and v_0 != v_rest so that a
    dropped division by cm, a dropped leak term or a wrong sign fails."""
    states, fired = single_neuron("IF_curr_exp", lif(cm=2.0, i_offset=0.4, v=-70.0), 60)
    assert fired == []
    v_inf = -65.0 + 0.4 * 20.0 / 2.0
    exact = v_inf + (-70.0 - v_inf) * (1 - 1 / 20.0) ** np.arange(1, 61)
    assert np.max(np.abs(states[:, 0] - exact)) < 1e-4
    assert np.all(states[:, 3] == 100.0 + np.arange(1, 61))  # steps since the last firing


@pytest.mark.parametrize("w", [0.75, -0.5])
def test_lif_synaptic_current_decays_exponentially(w):
    """A single spike of weight w arrives one step after an input neuron is
    forced to fire; it enters Ie (w >= 0) or Ii (w < 0), which then decay as
    (1 - 1/tau_syn)^k with tau_syn 5 (E) and 4 (I).  Catches E/I swapped, a
    decay factor of 1/tau or exp(-1/tau), or currents added after the decay
    of the same step."""
    net = oa.Network()
    tin = net.add_neuron_type("Input")
    net.add_neuron(tin, 0)
    t = net.add_neuron_type("IF_curr_exp")
    net.add_neuron(t, 1, *lif(tau_e=5.0, tau_i=4.0))
    net.add_synapse(0, 1, 1, w, 0)
    sim = oa.Simulation(net, oa.Configuration())
    sim.step(fstim=[0])
    ie, ii = [], []
    for _ in range(20):
        sim.step()
        ie.append(sim.get_neuron_state(1, 1))
        ii.append(sim.get_neuron_state(1, 2))
    k = np.arange(20)
    if w > 0:
        assert np.allclose(ie, w * (1 - 1 / 5.0) ** k, rtol=1e-5, atol=0) and not np.any(ii)
    else:
        assert np.allclose(ii, w * (1 - 1 / 4.0) ** k, rtol=1e-5, atol=0) and not np.any(ie)


@pytest.mark.parametrize("tau_refrac", [0.0, 1.0, 2.0, 2.5, 5.0])
def test_lif_refractory_period(tau_refrac):
    """After a firing (forced at step 0) lastfired = 1, and the neuron is
    refractory -- v held at v_reset despite its input -- while
    lastfired <= tau_refrac: exactly floor(tau_refrac) steps.  Catches a
    strict comparison (one step fewer) or counting from 0."""
    states, fired = single_neuron("IF_curr_exp", lif(i_offset=1.0, tau_refrac=tau_refrac), 12, fstim=[0])
    assert fired == [0]
    held = 0
    while states[1 + held, 0] == -65.0:
        held += 1
    assert held == int(math.floor(tau_refrac))
    assert states[1 + held, 0] == f32(-65.0) + f32(1.0)  # the first free step from v_rest


def test_lif_forced_firing_during_refractory_period():
    """Firing can be forced even while refractory (IF_curr_exp.cpp:106-107)."""
    _, fired = single_neuron("IF_curr_exp", lif(tau_refrac=5.0), 4, fstim=[0, 1])
    assert fired == [0, 1]


def test_lif_threshold_is_strict():
    """v reaching exactly v_thresh does not fire; the step after, above it,
    does: from v = v_rest = -60 with I = 10 and cm = 1 the first step lands
    on -50 = v_thresh exactly, the second on -40.5.  Catches v >= v_thresh."""
    states, fired = single_neuron("IF_curr_exp", lif(v_rest=-60.0, v_reset=-60.0, v=-60.0, i_offset=10.0), 3)
    assert states[0, 0] == -50.0
    assert fired[0] == 1


@pytest.mark.parametrize("multiple", [False, True])
def test_lif_create_and_run(multiple):
    """src/test/IF_curr_exp.cpp:15-49: a network with an IF_curr_exp neuron
    (alone or beside a ring of Izhikevich neurons) is created and run 1000
    steps without error."""
    net = oa.Network()
    if multiple:
        gen.ring(1024, 0, delay=1).load(net)
    t = net.add_neuron_type("IF_curr_exp")
    net.add_neuron(t, 2000, -65.0, 1.0, 20.0, 5.0, 2.0, 5.0, 0.1, -70.0, -51.0, -65.0, 0.0, 0.0, 1000.0)
    sim = oa.Simulation(net, oa.Configuration())
    for _ in range(1000):
        sim.step()


# -------------------------------------------------------------------- HH
HH_REST = [-65.0, 0.3177, 0.0529, 0.5961, 0.0]


def hh_textbook(I, steps, substeps, v0=0.0, n=0.3177, m=0.0529, h=0.5961):
    """Hodgkin & Huxley (1952) in double precision, membrane potential V in
    mV relative to rest (V = v + 65), conductances in mS/cm^2, C = 1 uF/cm^2,
    forward Euler with substeps of 0.001 ms, constant injected current I.
    Returns (v, n, m, h) after every step and the substeps at which a spike
    peaks (V >= 20 mV relative to rest, first decrease after an increase)."""
    out, peaks = [], []
    V, dvp = v0, 0.0
    for s in range(steps):
        for k in range(substeps):
            an = 0.01 * (10 - V) / (math.exp((10 - V) / 10) - 1)
            am = 0.1 * (25 - V) / (math.exp((25 - V) / 10) - 1)
            ah = 0.07 * math.exp(-V / 20)
            bn = 0.125 * math.exp(-V / 80)
            bm = 4 * math.exp(-V / 18)
            bh = 1 / (math.exp((30 - V) / 10) + 1)
            m, n, h = m + 0.001 * (am * (1 - m) - bm * m), n + 0.001 * (an * (1 - n) - bn * n), \
                h + 0.001 * (ah * (1 - h) - bh * h)
            i_ion = 120 * m ** 3 * h * (V - 115) + 36 * n ** 4 * (V + 12) + 0.3 * (V - 10.6)
            Vn = V + 0.001 * (I - i_ion)
            if dvp >= 0 and Vn - V < 0 and Vn >= 20:
                peaks.append(s)
            dvp = Vn - V
            V = Vn
        out.append([V - 65, n, m, h])
    return np.array(out), peaks


def test_hh_rest_state_is_stationary():
    """The resting state (v = -65 mV, gating variables at their steady-state
    values there) does not move without input, and never fires.  Catches a
    wrong reversal potential, conductance or sign in the ionic currents."""
    states, fired = single_neuron("HH", HH_REST, 200)
    assert fired == []
    assert np.max(np.abs(states[:, 0] + 65.0)) < 1e-3
    assert np.max(np.abs(states[:, 1:4] - np.array(HH_REST[1:4]))) < 1e-3


@pytest.mark.parametrize("I", [3.0, 10.0])
def test_hh_matches_textbook_equations_with_999_substeps(I):
    """One HH step is 999 forward-Euler substeps of 0.001 ms (the int of
    1/0.001f, which is 999.99994 in float): v, n, m, h after each of 40 steps
    under a constant current stay within 0.05 mV and 1e-3 of a
    double-precision integration of the textbook equations with 999
    substeps, while the same integration with 1000 substeps is off by more
    than 0.3 mV.  Catches 1000 substeps, any wrong rate function or
    constant, and Euler steps taken in another order."""
    states, fired = single_neuron("HH", HH_REST, 40, istim=I)
    ref999, peaks = hh_textbook(I, 40, 999)
    ref1000, _ = hh_textbook(I, 40, 1000)
    assert np.max(np.abs(states[:, 0] - ref999[:, 0])) < 0.05
    assert np.max(np.abs(states[:, 1:4] - ref999[:, 1:4])) < 1e-3
    assert np.max(np.abs(ref1000[:, 0] - ref999[:, 0])) > 0.3
    # spikes are reported at the step of their peak, never before step 10
    assert fired == [s for s in sorted(set(peaks)) if s >= 10]
    assert any(s < 10 for s in peaks)  # the onset spike happens, unreported


# --------------------------------------------------------------- Poisson
def xorshift128(state):
    """Marsaglia (2003) xor128 on an array of generators (n x 4 uint32),
    advanced in place; returns the n outputs."""
    s = state.astype(np.uint64)
    t = (s[:, 0] ^ (s[:, 0] << 11)) & 0xffffffff
    s0, s1, s2, s3 = s[:, 1], s[:, 2], s[:, 3], s[:, 3]
    w = (s3 ^ (s3 >> 19)) ^ (t ^ (t >> 8))
    state[:, 0], state[:, 1], state[:, 2], state[:, 3] = s0, s1, s2, w
    return w.astype(np.uint32)


def poisson_raster(n, rate, steps):
    net = oa.Network()
    t = net.add_neuron_type("PoissonSource")
    net.add_neuron(t, np.arange(n, dtype=np.uint32), rate)
    sim = oa.Simulation(net, oa.Configuration())
    out = np.zeros((steps, n), bool)
    for s in range(steps):
        out[s, sim.step()] = True
    return out


def test_poisson_spike_train_brute_force():
    """A Poisson source fires when the low 16 bits of its generator's next
    output are below trunc(rate * 65536) (PoissonSource.cpp:17-22).  The
    raster of 500 sources over 2000 steps equals the one predicted from an
    independent xorshift128 over the pinned seeds.  The rate 0.0101 gives
    661.9: a rounded threshold (662), a 32-bit comparison or a high-bit
    draw would change the raster (the sample holds draws equal to 661)."""
    n, steps, rate = 500, 2000, 0.0101
    got = poisson_raster(n, rate, steps)
    state = oa.rng_seeds(0, n - 1).copy()
    p0 = int(f32(rate) * f32(65536))
    assert p0 == 661
    low = np.zeros((steps, n), np.uint32)
    for s in range(steps):
        low[s] = xorshift128(state) & 0xffff
    assert np.array_equal(got, low < p0)
    assert np.count_nonzero(low == p0) > 0


@pytest.mark.parametrize("rate,always", [(0.0, False), (1.0, True), (1.5, True)])
def test_poisson_extreme_rates(rate, always):
    """Rate 0 never fires; a rate of 1 or more fires every step."""
    r = poisson_raster(200, rate, 300)
    assert np.all(r) if always else not np.any(r)


@pytest.mark.parametrize("duration,mix", [(1000, False), (10000, False), (1000, True), (10000, True)])
def test_poisson_rate(duration, mix):
    sc.poisson_rate(oa, duration, mix)


# ----------------------------------------------------------------- Input
@pytest.mark.parametrize("ncount,iz", [(1, False), (1000, False), (1, True), (1000, True)])
def test_input_fires_exactly_as_stimulated(ncount, iz):
    sc.input_neurons(oa, ncount, iz)


# --------------------------------------------------------------- Kuramoto
TWO_PI32 = float(np.float32(2 * math.pi))


def kuramoto_net(api, freq, phase, cmean=0.0, synapses=()):
    net = api.Network()
    t = net.add_neuron_type("Kuramoto")
    for i, (f, p) in enumerate(zip(freq, phase)):
        net.add_neuron(t, i, f, cmean, p)
    for src, tgt, d, w in synapses:
        net.add_synapse(src, tgt, d, w, 0)
    return api.Simulation(net, api.Configuration())


def kuramoto_fp64(freq, phase0, synapses, steps, cmean=0.0, h=0.05, H=32):
    """Independent double-precision model of cpu/plugins/Kuramoto.cpp: the
    history before step 0 runs each oscillator back one frequency per step
    (Kuramoto.cpp:137-159); in step c, oscillator j takes one RK4 step of h
    of theta' = w_j + (1/C) sum_k W_kj sin(theta_k(c - d_kj + 1) - theta),
    the source phases held at their history values, then wraps into
    [0, 2 pi)."""
    n = len(freq)
    hist = {0: np.array(phase0, np.float64)}
    for t in range(1, H):
        hist[-t] = np.mod(hist[-t + 1] - np.array(freq), 2 * math.pi)
    C = cmean if cmean > 0 else 1.0
    out = []
    for c in range(steps):
        cur = hist[c]
        nxt = np.empty(n)
        for j in range(n):
            inc = [(hist[c - (d - 1)][k], w) for k, tgt, d, w in synapses if tgt == j]

            def k_of(theta):
                return freq[j] + sum(w * math.sin(p - theta) for p, w in inc) / C
            k0 = k_of(cur[j])
            k1 = k_of(cur[j] + 0.5 * h * k0)
            k2 = k_of(cur[j] + 0.5 * h * k1)
            k3 = k_of(cur[j] + h * k2)
            nxt[j] = math.fmod(cur[j] + h * (k0 + 2 * k1 + 2 * k2 + k3) / 6, 2 * math.pi) % (2 * math.pi)
        hist[c + 1] = nxt
        out.append(nxt)
    return np.array(out)


def wrapped_diff(a, b):
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return np.minimum(d, 2 * math.pi - d)


def test_kuramoto_uncoupled_advance_by_h_omega():
    """Oscillators without coupling advance by h * omega = 0.05 omega per
    step (the step of cpu/plugins/Kuramoto.cpp).  NeMo's own tests
    (src/test/Kuramoto.cpp:48-82) expect a whole omega per step, which no
    NeMo plugin does; this checks the plugins' behaviour.  Catches a step
    other than h, a dropped RK4 weight, or a wrap other than into
    [0, 2 pi)."""
    freq = [0.1, 1.0, 7.0, 0.0]
    phase = [0.0, 3.0, 6.0, 1.0]
    sim = kuramoto_net(oa, freq, phase)
    assert [sim.get_neuron_state(i, 0) for i in range(4)] == [np.float32(p) for p in phase]
    got = []
    for _ in range(500):
        sim.step()
        got.append([sim.get_neuron_state(i, 0) for i in range(4)])
    ref = kuramoto_fp64(freq, phase, [], 500)
    assert np.max(wrapped_diff(got, ref)) < 2e-4
    assert np.all((np.array(got) >= 0) & (np.array(got) < TWO_PI32))
    assert abs(got[-1][0] - 500 * 0.05 * 0.1) < 1e-4


@pytest.mark.parametrize("delay", [1, 3, 20])
def test_kuramoto_delayed_coupling(delay):
    """Two oscillators, 0 -> 1 with strength 1 (the coupling of
    src/test/Kuramoto.cpp:85-124, at h 0.05): oscillator 1 follows the
    double-precision model within 1e-3 rad over 200 steps, including the
    first steps, where a delay > 1 reads the history filled backwards at
    construction.  Catches sin(target - source), a source phase read from
    the wrong history slot, or a history filled with h omega."""
    syn = [(0, 1, delay, 1.0)]
    sim = kuramoto_net(oa, [0.1, 0.1], [0.0, 1.57], synapses=syn)
    got = []
    for _ in range(200):
        sim.step()
        got.append([sim.get_neuron_state(i, 0) for i in range(2)])
    ref = kuramoto_fp64([0.1, 0.1], [0.0, 1.57], syn, 200)
    assert np.max(wrapped_diff(got, ref)) < 1e-3
    # the coupling pulls oscillator 1 towards oscillator 0
    assert wrapped_diff(got[-1][1], got[-1][0]) < wrapped_diff(1.57, 0.0)


def test_kuramoto_n_to_1_with_mean_coupling():
    """src/test/Kuramoto.cpp:139-185 at h 0.05, with noise connections of
    strength 0 and the coupling divided by Cmean = 2: the target follows the
    double-precision model."""
    n = 20
    freq = [0.1] * (n + 1)
    phase = [0.0] + [1.0] * n
    syn = [(k + 1, 0, 1, 1.0 / n) for k in range(n)]
    syn += [(k + 1, t, 2, 0.0) for k in range(n) for t in range(1, n + 1) if t != k + 1][:100]
    sim = kuramoto_net(oa, freq, phase, cmean=2.0, synapses=syn)
    got = []
    for _ in range(50):
        sim.step()
        got.append([sim.get_neuron_state(i, 0) for i in (0, 1)])
    ref = kuramoto_fp64(freq, phase, syn, 50, cmean=2.0)[:, :2]
    assert np.max(wrapped_diff(got, ref)) < 1e-4


def test_kuramoto_refused_with_stdp():
    """construction/RCM.cpp:83-91"""
    net = oa.Network()
    t = net.add_neuron_type("Kuramoto")
    net.add_neuron(t, 0, 0.1, 0.0, 0.0)
    net.add_neuron(t, 1, 0.1, 0.0, 0.0)
    net.add_synapse(0, 1, 1, 1.0, 1)
    conf = oa.Configuration()
    pre, post = gen.stdp_exp_function(0.1, -0.12, 20.0, 20)
    conf.set_stdp_function(pre, post, 0.0, 10.0, 0.0, -10.0)
    with pytest.raises(oa.OracleError) as e:
        oa.Simulation(net, conf)
    assert e.value.code == 5


# ------------------------------------------------------------ KuramotoLag
def kuramoto_lag_net(api, freq, phase, alpha, cmean=0.0, synapses=()):
    net = api.Network()
    t = net.add_neuron_type("KuramotoLag")
    for i, (f, p, a) in enumerate(zip(freq, phase, alpha)):
        net.add_neuron(t, i, f, cmean, a, p)
    for src, tgt, d, w in synapses:
        net.add_synapse(src, tgt, d, w, 0)
    return api.Simulation(net, api.Configuration())


def kuramoto_lag_fp64(freq, phase0, alpha, synapses, steps, cmean=0.0, h=0.05, H=32):
    """kuramoto_fp64 with each oscillator's lag alpha_j inside its sines:
    theta_j' = w_j + (1/C) sum_k W_kj sin(theta_k - theta_j - alpha_j)
    (cpu/plugins/KuramotoLag.cpp:15-30)."""
    n = len(freq)
    hist = {0: np.array(phase0, np.float64)}
    for t in range(1, H):
        hist[-t] = np.mod(hist[-t + 1] - np.array(freq), 2 * math.pi)
    C = cmean if cmean > 0 else 1.0
    out = []
    for c in range(steps):
        cur = hist[c]
        nxt = np.empty(n)
        for j in range(n):
            inc = [(hist[c - (d - 1)][k], w) for k, tgt, d, w in synapses if tgt == j]

            def k_of(theta):
                return freq[j] + sum(w * math.sin(p - theta - alpha[j]) for p, w in inc) / C
            k0 = k_of(cur[j])
            k1 = k_of(cur[j] + 0.5 * h * k0)
            k2 = k_of(cur[j] + 0.5 * h * k1)
            k3 = k_of(cur[j] + h * k2)
            nxt[j] = math.fmod(cur[j] + h * (k0 + 2 * k1 + 2 * k2 + k3) / 6, 2 * math.pi) % (2 * math.pi)
        hist[c + 1] = nxt
        out.append(nxt)
    return np.array(out)


def test_kuramoto_lag_zero_is_kuramoto():
    """With alpha 0 the lagged model is the Kuramoto model bit for bit
    (KuramotoLag.cpp differs from Kuramoto.cpp only by the lag).  Catches a
    lag added where none is asked for, or a parameter read from the wrong
    plane."""
    rng = np.random.default_rng(3)
    n = 40
    freq = rng.random(n).astype(np.float32)
    phase = (rng.random(n) * 2 * math.pi).astype(np.float32)
    syn = [(int(s), int(t), int(d), float(w)) for s, t, d, w in
           zip(rng.integers(0, n, 300), rng.integers(0, n, 300), rng.integers(1, 33, 300), rng.random(300) - 0.5)]
    syn.sort(key=lambda x: x[0])
    a = kuramoto_net(oa, freq, phase, synapses=syn)
    b = kuramoto_lag_net(oa, freq, phase, [0.0] * n, synapses=syn)
    for _ in range(100):
        a.step()
        b.step()
    pa = np.float32([a.get_neuron_state(i, 0) for i in range(n)])
    pb = np.float32([b.get_neuron_state(i, 0) for i in range(n)])
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))


@pytest.mark.parametrize("alpha", [0.5, -1.2])
def test_kuramoto_lag_coupled_pair(alpha):
    """Two mutually coupled oscillators with lags alpha and alpha / 2 (strength
    1, delays 1 and 4) follow the double-precision lagged model within 1e-3
    rad over 200 steps, and an uncoupled third oscillator advances by h omega
    whatever its lag.  Catches the lag's sign, the lag taken from the source
    instead of the target, or the lag applied outside the sine."""
    freq = [0.2, 0.5, 0.3]
    phase = [0.0, 2.0, 1.0]
    al = [alpha, alpha / 2, 0.9]
    syn = [(0, 1, 1, 1.0), (1, 0, 4, 1.0)]
    sim = kuramoto_lag_net(oa, freq, phase, al, synapses=syn)
    got = []
    for _ in range(200):
        sim.step()
        got.append([sim.get_neuron_state(i, 0) for i in range(3)])
    ref = kuramoto_lag_fp64(freq, phase, al, syn, 200)
    assert np.max(wrapped_diff(got, ref)) < 1e-3
    assert np.max(wrapped_diff(np.array(got)[:, 2], ref[:, 2])) < 2e-4
    # the lagged model differs from the unlagged one
    plain = kuramoto_fp64(freq, phase, syn, 200)
    assert np.max(wrapped_diff(np.array(got)[:, :2], plain[:, :2])) > 1e-2
