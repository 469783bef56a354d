"""Pins of the oracle's restatement of NeMo's CPU-only models (src/nemo/
cpu/plugins: IzhikevichChaos, IzhikevichSyn, Izhikevich2007,
Izhikevich2007DopSyn, BayesianNeuron, InhomogeneousPoissonSource), from the
models' published equations and properties they fix, evaluated here with
numpy independently of the oracle.  Each test names the mistake it would
catch."""
import math

import numpy as np
import pytest

import oracle_api as oa

f32 = np.float32


def single(model, args, steps, istim=None, fstim=()):
    net = oa.Network()
    t = net.add_neuron_type(model)
    net.add_neuron(t, 0, *args)
    sim = oa.Simulation(net, oa.Configuration())
    fired = []
    for s in range(steps):
        stim = [(0, istim[s])] if istim is not None else None
        if sim.step(fstim=[0] if s in fstim else None, istim=stim):
            fired.append(s)
    return sim, fired


def izh_v_step(v, u, I, h):
    """v after one substep of h: the Izhikevich (2003) update with the
    quadratic in double precision and v stored as float."""
    vd = np.float64(v)
    return f32(vd + np.float64(f32(h)) * (((0.04 * vd + 5.0) * vd + 140.0) - np.float64(u) + np.float64(I)))


# ----------------------------------------------------------- IzhikevichChaos
@pytest.mark.parametrize("I", [0.0, 10.0])
def test_chaos_without_noise_is_the_izhikevich_model_at_20_substeps(I):
    """sigma 0: spike times and state equal the Izhikevich equations
    integrated with 20 substeps of 0.05 ms (not 4 of 0.25).  Catches the
    substep count or length of the plain Izhikevich plugin."""
    a, b, c, d = f32(0.02), f32(0.2), f32(-65.0), f32(8.0)
    sim, fired = single("IzhikevichChaos", [0.02, 0.2, -65.0, 8.0, 0.0, -13.0, -65.0], 500, istim=[I] * 500)
    u, v = f32(-13.0), f32(-65.0)
    ref = []
    for t in range(500):
        fr = False
        for _ in range(20):
            if not fr:
                v = izh_v_step(v, u, f32(I), 0.05)
                u = f32(u + f32(0.05) * (a * (b * v - u) + f32(0.0)))
                fr = bool(v >= 30.0)
        if fr:
            ref.append(t)
            v, u = c, f32(u + d)
    assert fired == ref
    assert sim.get_neuron_state(0, 0) == u and sim.get_neuron_state(0, 1) == v
    assert (len(fired) > 0) == (I > 0)


def test_chaos_noise_drives_u_not_the_input():
    """a = b = 0 and no input: v follows the noise-free trajectory exactly
    while u moves by 20 x 0.05 x sigma x nrand per step, nrand drawn once per
    step from the neuron's generator.  Catches noise added to the input,
    scaled per substep by sqrt(dt), or drawn per substep."""
    sigma = 0.7
    sim, _ = single("IzhikevichChaos", [0.0, 0.0, -65.0, 8.0, sigma, -13.0, -70.0], 0)
    rng = sim.rng_state().copy()
    u, v = f32(-13.0), f32(-70.0)
    for t in range(50):
        noise = f32(sigma) * oa.nrand_n(rng)[0]
        sim.step()
        u_new = sim.get_neuron_state(0, 0)
        assert abs(float(u_new) - float(u) - float(noise)) < 1e-5 * max(1.0, abs(float(u)))
        for _ in range(20):
            v = izh_v_step(v, u, f32(0.0), 0.05)
            u = f32(u + f32(0.05) * (f32(0.0) + noise))
        assert sim.get_neuron_state(0, 1) == v
        u = u_new


# ------------------------------------------------------------- IzhikevichSyn
SYN = [0.02, 0.2, -65.0, 8.0, 30.0, 0.0, 0.9, 1.0, 0.0, 0.66, 1.0, -80.0]


def test_syn_conductances_decay_geometrically_and_clamp():
    """Without input ge and gi shrink by their factors every step (float
    products); a conductance above G_exc is clamped only when the step
    computes it (the clamp follows the input).  Catches a decay applied
    after the input or a clamp of the wrong sign."""
    sim, _ = single("IzhikevichSyn", SYN + [-13.0, -65.0, 0.8, -0.5], 0)
    ge, gi = f32(0.8), f32(-0.5)
    for t in range(30):
        sim.step()
        ge, gi = f32(ge * f32(0.9)), f32(gi * f32(0.66))
        assert sim.get_neuron_state(0, 2) == ge and sim.get_neuron_state(0, 3) == gi
    sim2, _ = single("IzhikevichSyn", SYN[:7] + [0.5] + SYN[8:] + [-13.0, -65.0, 3.0, -4.0], 1)
    assert sim2.get_neuron_state(0, 2) == f32(0.5)    # 3.0 * 0.9 clamped to G_exc
    assert sim2.get_neuron_state(0, 3) == f32(-1.0)   # -4.0 * 0.66 clamped to -G_inh


def test_syn_membrane_follows_conductance_input():
    """v and u over 200 steps with ge = 0.3, gi = 0.2 decaying and no
    synapses: the Izhikevich substeps with I = ge (E_exc - v) - gi (E_inh - v)
    and threshold vpeak = 30, recomputed independently.  Catches the input's
    sign convention or reversal potentials swapped."""
    sim, fired = single("IzhikevichSyn", SYN + [-13.0, -65.0, 0.3, 0.2], 200, istim=[8.0] * 200)
    a, b, c, d = f32(0.02), f32(0.2), f32(-65.0), f32(8.0)
    u, v, ge, gi = f32(-13.0), f32(-65.0), f32(0.3), f32(0.2)
    ref = []
    for t in range(200):
        ge, gi = f32(ge * f32(0.9)), f32(gi * f32(0.66))
        I = f32(f32(f32(ge * f32(f32(0.0) - v)) - f32(gi * f32(f32(-80.0) - v))) + f32(8.0))
        fr = False
        for _ in range(4):
            if not fr:
                v = izh_v_step(v, u, I, 0.25)
                u = f32(u + f32(0.25) * (a * (b * v - u)))
                fr = bool(v >= f32(30.0))
        if fr:
            ref.append(t)
            v, u = c, f32(u + d)
    assert fired == ref and len(ref) > 0
    assert sim.get_neuron_state(0, 0) == u and sim.get_neuron_state(0, 1) == v


# ------------------------------------------------- Izhikevich2007 (+ DopSyn)
# striatal medium spiny neuron of Izhikevich (2007), section 8.4.1
MSN = [0.01, -20.0, -55.0, 91.0, 40.0, -80.0, -29.7, 1.0, 15.2, 0.0]
TAIL = [0.9, 1.0, 0.0, 0.66, 1.0, -80.0]


def izh2007_reference(I_ext, steps, v0=-80.0, u0=0.0, dop=None):
    """Izhikevich (2007) eq. 8.5 in float: C v' = k (v - Vr)(v - Vt) - u + I,
    u' = a (b (v - Vr) - u), four substeps of 0.25 ms, reset v = c, u += d
    at vpeak; with dopamine (d1, d2), Vr (1 + K d1), d (1 - L d1) and
    k (1 - alpha d2) in double, rounded to float."""
    a, b, c, d, vpeak, Vr, Vt, k, C = (f32(x) for x in MSN[:9])
    if dop:
        d1, d2 = dop
        Vr = f32(np.float64(Vr) * (1 + 0.0289 * np.float64(f32(d1))))
        d = f32(np.float64(d) * (1 - 0.331 * np.float64(f32(d1))))
        k = f32(np.float64(k) * (1 - 0.032 * np.float64(f32(d2))))
    v, u = f32(v0), f32(u0)
    oneOverC = f32(f32(1.0) / C)
    fired = []
    for t in range(steps):
        I = f32(I_ext)
        fr = False
        for _ in range(4):
            if not fr:
                v = f32(v + f32(f32(f32(0.25) * oneOverC) * f32(f32(f32(f32(k * f32(v - Vr)) * f32(v - Vt)) - u) + I)))
                u = f32(u + f32(0.25) * f32(a * f32(f32(b * f32(v - Vr)) - u)))
                fr = bool(v >= vpeak)
        if fr:
            fired.append(t)
            v, u = c, f32(u + d)
    return fired, u, v


@pytest.mark.parametrize("I", [0.0, 300.0, 600.0])
def test_izhikevich2007_msn(I):
    """The medium spiny neuron (silent at rest, firing under strong drive),
    with zero conductances: spike times and state equal the 2007 equations.
    Catches C applied to u or I only, Vr in place of Vt, or double
    arithmetic where the model is float."""
    sim, fired = single("Izhikevich2007", MSN + TAIL + [0.0, -80.0, 0.0, 0.0], 300, istim=[I] * 300)
    ref, u, v = izh2007_reference(I, 300)
    assert fired == ref
    assert sim.get_neuron_state(0, 0) == u and sim.get_neuron_state(0, 1) == v
    assert (len(fired) > 0) == (I >= 600.0)


def test_izhikevich2007_conductance_set_by_input():
    """Any positive excitatory input sets ge to 1 (not adds), then ge decays
    by its factor: one neuron driving another through a synapse."""
    net = oa.Network()
    t = net.add_neuron_type("Izhikevich2007")
    net.add_neuron(t, 0, *(MSN + TAIL + [0.0, -80.0, 0.0, 0.0]))
    net.add_neuron(t, 1, *(MSN + TAIL + [0.0, -80.0, 0.0, 0.0]))
    net.add_synapse(0, 1, 1, 0.25, 0)
    sim = oa.Simulation(net, oa.Configuration())
    sim.step(fstim=[0])
    sim.step()
    assert sim.get_neuron_state(1, 2) == f32(1.0)
    sim.step()
    assert sim.get_neuron_state(1, 2) == f32(0.9)


def test_dopsyn_without_dopamine_is_izhikevich2007():
    """d1 = d2 = 0, no noise, zero conductances: bit-identical to
    Izhikevich2007 (the two plugins differ only in the dopamine scaling and
    the conductance synapses).  Catches a dopamine term applied at zero
    level."""
    a = single("Izhikevich2007", MSN + TAIL + [0.0, -80.0, 0.0, 0.0], 300, istim=[650.0] * 300)
    b = single("Izhikevich2007DopSyn", MSN + [0.0, 0.0] + TAIL + [0.0, -80.0, 0.0, 0.0], 300, istim=[650.0] * 300)
    assert a[1] == b[1] and len(a[1]) > 0
    for var in range(4):
        assert a[0].get_neuron_state(0, var) == b[0].get_neuron_state(0, var)


def test_dopsyn_dopamine_scales_vr_k_and_d():
    """d1 = 0.3, d2 = 0.2: the spike times and state of the 2007 equations
    with Vr, k and d scaled by the dopamine constants (Vr (1 + 0.0289 d1),
    d (1 - 0.331 d1), k (1 - 0.032 d2)).  Catches the constants' signs or
    the wrong level scaling a parameter."""
    sim, fired = single("Izhikevich2007DopSyn", MSN + [0.3, 0.2] + TAIL + [0.0, -80.0, 0.0, 0.0], 300,
                        istim=[650.0] * 300)
    ref, u, v = izh2007_reference(650.0, 300, dop=(0.3, 0.2))
    assert fired == ref and len(ref) > 0
    assert sim.get_neuron_state(0, 0) == u and sim.get_neuron_state(0, 1) == v


# ------------------------------------------------------------ BayesianNeuron
def test_bayesian_log_odds_dynamics():
    """L and G in double precision (four Euler substeps of 0.25 of
    L' = rOn (1 + e^-L) - rOff (1 + e^L) + I and the same for G without I,
    G += g0 after a spike when L > G + g0/2): spike times equal, state within
    1e-4.  Catches the exponent's sign, the threshold without the half, or
    the adaptation not applied."""
    rOn, rOff, g0 = 0.001, 0.01, 1.5
    I = 0.8
    sim, fired = single("BayesianNeuron", [rOn, rOff, g0, -4.0, 0.0], 400, istim=[I] * 400)
    L, G = -4.0, 0.0
    ref = []
    for t in range(400):
        fr = False
        for _ in range(4):
            if not fr:
                L += 0.25 * (rOn * (1 + math.exp(-L)) - rOff * (1 + math.exp(L)) + I)
                G += 0.25 * (rOn * (1 + math.exp(-G)) - rOff * (1 + math.exp(G)))
                fr = L > G + g0 / 2
        if fr:
            ref.append(t)
            G += g0
    assert fired == ref and len(ref) > 5
    assert abs(sim.get_neuron_state(0, 0) - L) < 1e-4 * max(1, abs(L))
    assert abs(sim.get_neuron_state(0, 1) - G) < 1e-4 * max(1, abs(G))


# ------------------------------------------------ InhomogeneousPoissonSource
def test_inhomogeneous_poisson_rate_is_held_external_current():
    """The rate is the external current in Hz, read every step and not
    cleared: set once, it holds.  A brute-force spike train from the
    neuron's xorshift state (fire when the low 16 bits are below
    0.001 rate 2^16), then rate 0 silences it and rate 1000 Hz fires every
    step.  Catches a rate per ms instead of per s, the current cleared after
    use, or a threshold compared the other way."""
    net = oa.Network()
    t = net.add_neuron_type("InhomogeneousPoissonSource")
    net.add_neuron(t, 0)
    sim = oa.Simulation(net, oa.Configuration())
    rng = sim.rng_state().copy()[0]
    rate = 37.5
    p0 = int(0.001 * rate * 65536.0)
    fired = [bool(sim.step(istim=[(0, rate)] if s == 0 else None)) for s in range(2000)]
    ref = [(oa.urand(rng) & 0xffff) < p0 for _ in range(2000)]
    assert fired == ref and 40 < sum(fired) < 110
    assert not any(sim.step(istim=[(0, 0.0)] if s == 0 else None) for s in range(500))
    assert all(sim.step(istim=[(0, 1000.0)] if s == 0 else None) for s in range(500))
