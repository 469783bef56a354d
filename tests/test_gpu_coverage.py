"""Parity at the benchmark sizes, batches against the oracle, the neuron
setters during a simulation, write-only synapses, and the FP32 state error.

Every call goes through the C ABI of libnemo_hip.so; the oracle is the CPU
restatement of NeMo's CPU backend.  Bar: bit-exact rasters, events, weights,
STDP accumulators, state and RNG state.
"""
import numpy as np
import pytest

import nemo_amd
from nemo_amd import generators as gen
import oracle_api as oa
import scenarios as sc

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu(require_gpu):
    pass


def stdp_fn():
    pre, post = gen.stdp_exp_function(0.1, -0.12, 20.0, 20)
    return (pre, post, 0.0, 10.0, 0.0, -10.0)


def conf_of(api, stdp):
    c = api.Configuration()
    if stdp:
        c.set_stdp_function(*stdp_fn())
    return c


def same_f32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    an, bn = np.isnan(a), np.isnan(b)
    return np.array_equal(an, bn) and np.array_equal(a.view(np.uint32)[~an], b.view(np.uint32)[~bn])


def expected_events(spec, counts, fired, steps):
    """Synaptic events a run of `steps` steps delivers, from its raster, for
    a network of neurons 0..N-1: a firing of n at step s delivers every
    synapse of n whose delay d satisfies s + d <= steps - 1
    (cpu/Simulation.cpp:259-277)."""
    n = int(spec.neuron_count)
    D = int(spec.delay.max())
    key = spec.source.astype(np.int64) * (D + 1) + spec.delay.astype(np.int64)
    rowlen = np.bincount(key, minlength=n * (D + 1)).reshape(n, D + 1)
    cum = np.cumsum(rowlen, axis=1)
    step_of = np.repeat(np.arange(steps), counts)
    horizon = np.minimum(steps - 1 - step_of, D)
    return int(cum[fired.astype(np.int64), horizon].sum())


# ------------------------------------------------------------ C2 at full size
def test_full_size_stdp_network_bit_exact():
    """The 100k-neuron, 1000-synapse Izhikevich network with STDP on its 8e7
    excitatory synapses: 1000 steps with applyStdp(1.0) after step 1000, then
    50 more steps, in the HIP backend (graph runs) and in the oracle stepped
    one step at a time.  Counts, fired lists, events, every weight after
    applyStdp, every STDP accumulator, u and v of every neuron (including the
    three the reference dynamics drive to NaN), firing histories and RNG
    states are bit-exact."""
    spec = gen.random_izhikevich(100000, 1000, dmax=20, seed=2, stdp=True)
    ref = sc.make(oa, spec, conf_of(oa, True))
    hip = sc.make(nemo_amd, spec, conf_of(nemo_amd, True))
    results = []
    for sim in (ref, hip):
        c1, f1, e1 = sim.run(1000, stdp_period=1000, reward=1.0)
        c2, f2, e2 = sim.run(50)
        results.append((np.concatenate([c1, c2]), np.concatenate([f1, f2]), e1 + e2))
    (cr, fr, er), (ch, fh, eh) = results
    assert np.array_equal(cr, ch)
    assert np.array_equal(fr, fh)
    assert er == eh
    assert cr.sum() > 100000  # the network fires throughout
    nsyn = spec.synapse_count
    wr, wh = ref.fixed_weights(nsyn), hip.fixed_weights()
    assert np.array_equal(wr, wh)
    assert not np.array_equal(wh[spec.plastic == 1], (spec.weight[spec.plastic == 1] * 2 ** 20).astype(np.int32))
    ar, ah = ref.stdp_accumulators(nsyn), hip.stdp_accumulators()
    assert np.array_equal(ar, ah)
    assert np.count_nonzero(ah) > 0
    assert np.array_equal(ref.recent_firing(), hip.recent_firing())
    assert np.array_equal(ref.rng_state(), hip.rng_state())
    for var in (0, 1):
        # every neuron; a few (3 of 100k) diverge to NaN under the reference
        # dynamics, and do so here too
        assert same_f32(ref.neuron_states(var), hip.neuron_states(var)), var


# -------------------------------------------------------------- C4 batches
def _poisson_first_spec(seed):
    """Poisson sources added before Izhikevich neurons: the types come in
    another order than in the Izhikevich-only networks of the batch."""
    rng = np.random.default_rng(seed)
    npo, niz = 300, 700
    po = gen.NetworkSpec(blocks=[gen.NeuronBlock("PoissonSource", np.arange(npo, dtype=np.uint32),
                                                 [np.full(npo, 0.02, np.float32)])])
    iz = gen.random_izhikevich(niz, 0, seed=seed, first_index=npo)
    spec = gen.merge(po, iz)
    n = npo + niz
    src = np.repeat(np.arange(n, dtype=np.uint32), 30)
    spec.source = src
    spec.target = rng.integers(npo, n, size=src.size, dtype=np.uint32)
    spec.delay = rng.integers(1, 9, size=src.size, dtype=np.uint32)
    spec.weight = (rng.random(src.size, dtype=np.float32) * np.float32(3.0)).astype(np.float32)
    spec.plastic = np.zeros(src.size, np.uint8)
    return spec


def test_batch_networks_match_oracle():
    """Seven networks (sizes 200..3000, 20..200 synapses per neuron, STDP on
    every other one, one with Poisson sources added first) advanced as one
    batch for 250 steps with applyStdp(1.0) every 100: each network's raster
    equals the oracle's run of that network alone, and so do the weights and
    STDP accumulators of every synapse, and the state of its neurons."""
    specs = gen.batch_izhikevich(count=6, nmin=200, nmax=3000, smin=20, smax=200, seed=13)
    specs.insert(3, (_poisson_first_spec(5), False))
    nets = []
    for spec, _ in specs:
        net = nemo_amd.Network()
        spec.load(net)
        nets.append(net)
    batch = nemo_amd.Batch(nets, conf_of(nemo_amd, True))
    steps, period = 250, 100
    counts, fired, events = batch.run(steps, stdp_period=period, reward=1.0)
    per = nemo_amd.Batch.split(counts, fired)
    ref_events = 0
    wref, aref = [], []
    for k, (spec, stdp) in enumerate(specs):
        ref = sc.make(oa, spec, conf_of(oa, stdp))
        c, f, e = ref.run(steps, stdp_period=period if stdp else 0, reward=1.0)
        assert np.array_equal(per[k][0], c), ("counts", k)
        assert np.array_equal(per[k][1], f), ("fired", k)
        assert c.sum() > 0
        ref_events += e
        wref.append(ref.fixed_weights(spec.synapse_count))
        aref.append(ref.stdp_accumulators(spec.synapse_count))
        iz = spec.blocks[-1].idx
        for i in iz[:: max(1, iz.size // 9)]:
            for var in (0, 1):
                assert same_f32(batch.get_neuron_state(int(i), var, net=k), ref.get_neuron_state(int(i), var)), (k, i)
    assert events == ref_events
    assert np.array_equal(batch.fixed_weights(), np.concatenate(wref))
    acc = batch.stdp_accumulators()
    assert np.array_equal(acc, np.concatenate(aref))
    assert np.count_nonzero(acc) > 0


def test_batch_of_64_networks_full_size():
    """The 64-network batch at full size (neuron counts log-uniform in
    [1k, 50k], U{100..1000} synapses per neuron, STDP on every other
    network): 120 steps with applyStdp(1.0) after step 100.  Every network's
    delivered events equal what its own raster implies, and four networks
    (the smallest, the largest, and one with and one without STDP) have
    rasters and weights identical to their own solo simulations."""
    specs = gen.batch_izhikevich(count=64)
    nets = []
    for spec, _ in specs:
        net = nemo_amd.Network()
        spec.load(net)
        nets.append(net)
    batch = nemo_amd.Batch(nets, conf_of(nemo_amd, True))
    assert batch.size == 64
    steps, period = 120, 100
    counts, fired, events = batch.run(steps, stdp_period=period, reward=1.0)
    per = nemo_amd.Batch.split(counts, fired)
    expected = 0
    for k, (spec, _) in enumerate(specs):
        expected += expected_events(spec, per[k][0], per[k][1], steps)
        assert per[k][0].sum() > 0, k
    assert events == expected
    wbatch = batch.fixed_weights()
    offsets = np.concatenate([[0], np.cumsum([s.synapse_count for s, _ in specs])])
    sizes = [s.neuron_count for s, _ in specs]
    sample = {int(np.argmin(sizes)), int(np.argmax(sizes)), 10, 21}
    for k in sorted(sample):
        spec, stdp = specs[k]
        solo = nemo_amd.Simulation(nets[k], conf_of(nemo_amd, stdp))
        c, f, _ = solo.run(steps, stdp_period=period if stdp else 0, reward=1.0)
        assert np.array_equal(per[k][0], c), ("counts", k)
        assert np.array_equal(per[k][1], f), ("fired", k)
        assert np.array_equal(wbatch[offsets[k]:offsets[k + 1]], solo.fixed_weights()), ("weights", k)
        del solo


# ------------------------------------------------------------- run variants
def test_run_without_raster_matches_run_with_raster():
    """run(n, raster=False) takes the same steps as run(n): same state,
    weights, accumulators and events over a run longer than the pinned
    scalar ring (300 chunks of 64 steps would reuse each of its 256 slots)."""
    spec = gen.random_izhikevich(2000, 100, dmax=20, seed=23, stdp=True)
    a = sc.make(nemo_amd, spec, conf_of(nemo_amd, True))
    b = sc.make(nemo_amd, spec, conf_of(nemo_amd, True))
    _, _, ea = a.run(20000, stdp_period=1000, raster=False)
    _, _, eb = b.run(20000, stdp_period=1000)
    assert ea == eb
    for var in (0, 1):
        assert same_f32(a.neuron_states(var), b.neuron_states(var))
    assert np.array_equal(a.rng_state(), b.rng_state())
    assert np.array_equal(a.fixed_weights(), b.fixed_weights())
    assert np.array_equal(a.stdp_accumulators(), b.stdp_accumulators())
    ca, fa, _ = a.run(100)
    cb, fb, _ = b.run(100)
    assert np.array_equal(ca, cb) and np.array_equal(fa, fb)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_autotune_from_every_stdp_mode_keeps_the_state(mode):
    """Autotuning from a simulation whose STDP accumulation runs in mode 0..4
    changes nothing the simulation computes: after it, rasters, weights and
    accumulators equal an untuned twin's."""
    spec = gen.random_izhikevich(5000, 200, dmax=20, seed=8, stdp=True)
    a = sc.make(nemo_amd, spec, conf_of(nemo_amd, True))
    b = sc.make(nemo_amd, spec, conf_of(nemo_amd, True))
    shape = list(a.launch_shape)
    shape[3] = mode
    shape[4] = 8 if mode == 4 else 2
    a.launch_shape = tuple(shape)
    ca1, fa1, _ = a.run(150, stdp_period=100)
    a.autotune(32)
    ca2, fa2, _ = a.run(150, stdp_period=100)
    cb1, fb1, _ = b.run(150, stdp_period=100)
    cb2, fb2, _ = b.run(150, stdp_period=100)
    assert np.array_equal(np.concatenate([ca1, ca2]), np.concatenate([cb1, cb2]))
    assert np.array_equal(np.concatenate([fa1, fa2]), np.concatenate([fb1, fb2]))
    assert np.array_equal(a.fixed_weights(), b.fixed_weights())
    assert np.array_equal(a.stdp_accumulators(), b.stdp_accumulators())


# ------------------------------------------------------ neuron access (F1)
def test_vprobe():
    """src/test/test.cpp:744-757: the membrane potential of a new simulation
    is the network's."""
    from nemo_amd._lib import check, lib
    import ctypes
    v0 = -65.0
    net = nemo_amd.Network()
    check(lib().nemo_add_neuron_iz(net._h, 0, 0.02, 0.2, -65.0 + 15.0 * 0.25, 8.0 - 6.0 * 0.25, 0.2 * -65.0, v0, 5.0))
    sim = nemo_amd.Simulation(net, nemo_amd.Configuration())
    assert sim.get_membrane_potential(0) == v0


def test_set_neuron():
    """src/test/test.cpp:775-968 (testSetNeuron), on the HIP backend."""
    sc.set_neuron(nemo_amd, nemo_amd.NemoError)


def test_set_neuron_during_a_run_matches_oracle():
    """Parameters and state changed between steps of a running network
    (setNeuron, setNeuronState, setNeuronParameter on a sample of neurons
    every 50 steps) take effect exactly as in the oracle: rasters and state
    bit-exact over 300 steps."""
    spec = gen.random_izhikevich(3000, 100, dmax=20, seed=31, stdp=True)
    ref = sc.make(oa, spec, conf_of(oa, True))
    hip = sc.make(nemo_amd, spec, conf_of(nemo_amd, True))
    rng = np.random.default_rng(31)
    for t in range(300):
        if t % 50 == 25:
            for i in rng.choice(3000, 40, replace=False):
                i = int(i)
                kind = rng.integers(3)
                if kind == 0:
                    v = float(rng.uniform(-70, -50))
                    for sim in (ref, hip):
                        sim.set_neuron_state(i, 1, v)
                elif kind == 1:
                    d = float(rng.uniform(2, 8))
                    for sim in (ref, hip):
                        sim.set_neuron_parameter(i, 3, d)
                else:
                    args = [0.02, 0.2, -60.0, 6.0, 3.0, -12.0, -62.0]
                    ref.set_neuron_state(i, 0, args[5])
                    ref.set_neuron_state(i, 1, args[6])
                    for p in range(5):
                        ref.set_neuron_parameter(i, p, args[p])
                    hip.set_neuron(i, *args)
        fr, fh = ref.step(), hip.step()
        assert fr == fh, ("raster", t)
        if t % 50 == 49:
            for var in (0, 1):
                r = np.array([ref.get_neuron_state(i, var) for i in range(3000)], np.float32)
                assert same_f32(r, hip.neuron_states(var)), (t, var)
    assert np.array_equal(ref.stdp_accumulators(spec.synapse_count), hip.stdp_accumulators())


def test_write_only_synapses():
    """Configuration.setWriteOnlySynapses (Configuration.hpp:132-140): the
    simulation runs exactly as with readable synapses, and synapse queries
    fail with an error instead of answering."""
    spec = gen.random_izhikevich(2000, 100, dmax=20, seed=41)
    a = sc.make(nemo_amd, spec)
    conf = nemo_amd.Configuration()
    conf.set_write_only_synapses()
    b = sc.make(nemo_amd, spec, conf)
    ca, fa, ea = a.run(300)
    cb, fb, eb = b.run(300)
    assert np.array_equal(ca, cb) and np.array_equal(fa, fb) and ea == eb
    for q in (lambda: b.get_synapses_from(0), lambda: b.get_synapse_weight(0), lambda: b.get_synapse_target(0)):
        with pytest.raises(nemo_amd.NemoError):
            q()
    assert len(a.get_synapses_from(0)) == 100


def test_deprecated_izhikevich_setters():
    """nemo_set_neuron_iz_n / nemo_set_neuron_iz_s (src/nemo.h:640-658): the
    arguments are (a, b, c, d, u, v, sigma); parameters are stored as
    a, b, c, d, sigma and state as u, v."""
    from nemo_amd._lib import check, lib
    net = nemo_amd.Network()
    check(lib().nemo_add_neuron_iz(net._h, 0, 0.02, 0.2, -65.0, 8.0, -13.0, -65.0, 5.0))
    check(lib().nemo_set_neuron_iz_n(net._h, 0, 0.03, 0.25, -60.0, 6.0, -14.0, -64.0, 4.0))
    assert [net.get_neuron_parameter(0, p) for p in range(5)] == [np.float32(x) for x in (0.03, 0.25, -60.0, 6.0, 4.0)]
    assert [net.get_neuron_state(0, s) for s in range(2)] == [np.float32(-14.0), np.float32(-64.0)]
    sim = nemo_amd.Simulation(net, nemo_amd.Configuration())
    check(lib().nemo_set_neuron_iz_s(sim._h, 0, 0.02, 0.2, -65.0, 8.0, -13.0, -70.0, 0.0))
    assert [sim.get_neuron_parameter(0, p) for p in range(5)] == [np.float32(x) for x in (0.02, 0.2, -65.0, 8.0, 0.0)]
    assert sim.get_membrane_potential(0) == np.float32(-70.0)
    assert lib().nemo_set_neuron_iz_s(sim._h, 1, 0, 0, 0, 0, 0, 0, 0) != 0


# ------------------------------------------------------- FP32 state error
def izhikevich_fp64_step(v, u, a, b, I):
    """One Izhikevich step (four 0.25 ms substeps, skipped after the
    threshold) of cpu/plugins/Izhikevich.cpp:67-94 with every quantity in
    FP64.  Returns (v, u, fired) before the reset."""
    fired = np.zeros(v.shape, bool)
    for _ in range(4):
        live = ~fired
        vn = v + 0.25 * ((0.04 * v + 5.0) * v + 140.0 - u + I)
        un = u + 0.25 * (a * (b * vn - u))
        v = np.where(live, vn, v)
        u = np.where(live, un, u)
        fired = fired | (live & (v >= 30.0))
    return v, u, fired


def test_fp32_state_error_against_fp64_recomputation(capsys):
    """The small benchmark network (1000 Izhikevich neurons x 100 synapses,
    sigma 5/2, 1000 steps): every step, the state the HIP backend computes in
    the reference's precision (v updated in FP64 and stored as FP32, u in
    FP32) is recomputed on the host in FP64 from the previous step's stored
    state and the same input current (the delivered fixed-point currents
    as floats plus sigma times the same normal deviate).  The one-step error
    of v and u is stated (printed) and bounded; so is the drift of a full
    FP64 trajectory driven by the same inputs, over the neurons whose spike
    trains stay identical."""
    n, steps = 1000, 1000
    spec = gen.random_izhikevich(n, 100, dmax=20, seed=11)
    sim = sc.make(nemo_amd, spec)
    sim.set_debug(True)
    a = np.asarray(spec.blocks[0].args[0], np.float64)
    b = np.asarray(spec.blocks[0].args[1], np.float64)
    c = np.asarray(spec.blocks[0].args[2], np.float64)
    d = np.asarray(spec.blocks[0].args[3], np.float64)
    sigma = np.asarray(spec.blocks[0].args[4], np.float32)
    u32 = sim.neuron_states(0)
    v32 = sim.neuron_states(1)
    vd, ud = v32.astype(np.float64), u32.astype(np.float64)
    diverged = np.zeros(n, bool)
    err_v = err_u = drift_v = drift_u = 0.0
    spikes = 0
    for t in range(steps):
        rng = sim.rng_state().copy()
        noise = oa.nrand_n(rng)
        fired = np.zeros(n, bool)
        fired[sim.step()] = True
        spikes += int(fired.sum())
        E, Ii = sim.currents()
        I = ((E + Ii) + np.float32(0.0)).astype(np.float32)
        I = np.where(sigma != 0, (I + sigma * noise).astype(np.float32), I).astype(np.float64)
        u_new, v_new = sim.neuron_states(0), sim.neuron_states(1)
        # one step in FP64 from the stored FP32 state
        v1, u1, f1 = izhikevich_fp64_step(v32.astype(np.float64), u32.astype(np.float64), a, b, I)
        v1 = np.where(f1, c, v1)
        u1 = np.where(f1, u1 + d, u1)
        same = f1 == fired
        err_v = max(err_v, float(np.max(np.abs(v_new[same] - v1[same]) / np.maximum(np.abs(v1[same]), 1.0))))
        err_u = max(err_u, float(np.max(np.abs(u_new[same] - u1[same]) / np.maximum(np.abs(u1[same]), 1.0))))
        # the FP64 trajectory from the initial state, same inputs
        vd, ud, fd = izhikevich_fp64_step(vd, ud, a, b, I)
        vd = np.where(fd, c, vd)
        ud = np.where(fd, ud + d, ud)
        diverged |= fd != fired
        ok = ~diverged
        drift_v = max(drift_v, float(np.max(np.abs(v_new[ok] - vd[ok]) / np.maximum(np.abs(vd[ok]), 1.0))))
        drift_u = max(drift_u, float(np.max(np.abs(u_new[ok] - ud[ok]) / np.maximum(np.abs(ud[ok]), 1.0))))
        u32, v32 = u_new, v_new
    with capsys.disabled():
        print("\nFP32 state error, 1000 x 100 Izhikevich, %d steps, %d spikes: one-step max relative error "
              "v %.3g, u %.3g; FP64 trajectory drift over %d of %d neurons with identical spike trains: "
              "v %.3g, u %.3g" % (steps, spikes, err_v, err_u, int((~diverged).sum()), n, drift_v, drift_u))
    assert spikes > 1000
    # Measured on an MI355X: v 1.31e-05, u 7.67e-07, and identical spike trains
    # for all 1000 neurons.  The one-step error of v is largest just below the
    # threshold, where each 0.25 ms substep multiplies a rounding of the stored
    # FP32 v by 1 + 0.25 (0.08 v + 5), about 2.9 at v = 30.
    assert err_v < 1e-4 and err_u < 1e-5
    assert (~diverged).sum() >= 990


# ------------------------------------------- reference model KATs on HIP
@pytest.mark.parametrize("duration,mix", [(1000, False), (10000, False), (1000, True), (10000, True)])
def test_poisson_rate(duration, mix):
    """src/test/PoissonSource.cpp:7-41 on the HIP backend."""
    sc.poisson_rate(nemo_amd, duration, mix)


@pytest.mark.parametrize("ncount,iz", [(1, False), (1000, False), (1, True), (1000, True)])
def test_input_fires_exactly_as_stimulated(ncount, iz):
    """src/test/Input.cpp:42-79 on the HIP backend."""
    sc.input_neurons(nemo_amd, ncount, iz)


@pytest.mark.parametrize("multiple", [False, True])
def test_lif_create_and_run(multiple):
    """src/test/IF_curr_exp.cpp:15-49 on the HIP backend, state compared with
    the oracle after 1000 steps."""
    sims = []
    for api in (oa, nemo_amd):
        net = api.Network()
        if multiple:
            gen.ring(1024, 0, delay=1).load(net)
        t = net.add_neuron_type("IF_curr_exp")
        net.add_neuron(t, 2000, -65.0, 1.0, 20.0, 5.0, 2.0, 5.0, 0.1, -70.0, -51.0, -65.0, 0.0, 0.0, 1000.0)
        sim = api.Simulation(net, api.Configuration())
        for _ in range(1000):
            sim.step()
        sims.append(sim)
    for var in range(4):
        assert same_f32(sims[0].get_neuron_state(2000, var), sims[1].get_neuron_state(2000, var))


# ------------------------------------------------ Brian's propagate (F4)
def _hip_runtime():
    """hipMemcpy of the HIP runtime the process already uses (torch's, which
    libnemo_hip.so binds to by soname): loading the library by soname
    reuses it.  A second runtime in the process would not know the first's
    device pointers, so the test refuses to go on if two are mapped."""
    import ctypes
    from nemo_amd._lib import lib
    lib()
    rt = ctypes.CDLL("libamdhip64.so.7")
    runtimes = {l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l}
    assert len(runtimes) == 1, runtimes
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return rt


def test_propagate_matches_oracle():
    """Brian's propagate (Simulation.hpp:93-113) on a 3000-neuron network
    after 50 ordinary steps: 100 calls with random fired sets return the
    oracle's E and I bit for bit (host form, then the device form with a
    device flag array), leave the neuron state alone, and ordinary steps
    afterwards stay identical to the oracle's."""
    import ctypes
    from nemo_amd._lib import check, lib
    spec = gen.random_izhikevich(3000, 100, dmax=20, seed=51)
    ref = sc.make(oa, spec)
    hip = sc.make(nemo_amd, spec)
    for t in range(50):
        assert ref.step() == hip.step(), t
    v_before = hip.neuron_states(1)
    rng = np.random.default_rng(51)
    for j in range(60):
        F = np.sort(rng.choice(3000, int(rng.integers(0, 300)), replace=False)).astype(np.uint32)
        er, ir = ref.propagate(F)
        eh, ih = hip.propagate(F)
        assert same_f32(er, eh) and same_f32(ir, ih), j
    assert same_f32(v_before, hip.neuron_states(1))
    import torch
    rt = _hip_runtime()
    n = hip.neuron_count
    l2g = hip.local_to_global()
    for j in range(40):
        F = np.sort(rng.choice(3000, int(rng.integers(0, 300)), replace=False)).astype(np.uint32)
        flags = torch.tensor(np.isin(l2g, F).astype(np.int32), device="cuda:0")
        torch.cuda.synchronize()
        pe, pi = ctypes.POINTER(ctypes.c_float)(), ctypes.POINTER(ctypes.c_float)()
        check(lib().nemo_hip_propagate_device(hip._h, ctypes.c_void_p(flags.data_ptr()), ctypes.byref(pe),
                                              ctypes.byref(pi)))
        eh, ih = np.zeros(n, np.float32), np.zeros(n, np.float32)
        assert rt.hipMemcpy(eh.ctypes.data, ctypes.cast(pe, ctypes.c_void_p), 4 * n, 2) == 0  # device to host
        assert rt.hipMemcpy(ih.ctypes.data, ctypes.cast(pi, ctypes.c_void_p), 4 * n, 2) == 0
        er, ir = ref.propagate(F)
        assert same_f32(er, eh) and same_f32(ir, ih), ("device", j)
    assert ref.elapsed_simulation() == hip.elapsed_simulation()
    for t in range(100):
        assert ref.step() == hip.step(), ("after", t)
    for var in (0, 1):
        r = np.array([ref.get_neuron_state(i, var) for i in range(3000)], np.float32)
        assert same_f32(r, hip.neuron_states(var))


def test_propagate_refused_with_stdp():
    """The CUDA backend asserts that STDP is off (cuda/Simulation.cpp:332)."""
    spec = gen.random_izhikevich(500, 20, dmax=10, seed=52, stdp=True)
    hip = sc.make(nemo_amd, spec, conf_of(nemo_amd, True))
    with pytest.raises(nemo_amd.NemoError):
        hip.propagate([1, 2, 3])
