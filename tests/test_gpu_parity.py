"""Parity of the HIP backend with the CPU oracle on the same seeded networks.

Bar: bit-exact rasters, fixed-point currents (as the floats the models see),
membrane state, RNG state, firing history, STDP accumulators and weights.
Every call goes through the C ABI of libnemo_hip.so.
"""
import ctypes
import os
import subprocess
import tempfile

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


def both(spec, stdp=None):
    """The same network in the oracle and in the HIP backend."""
    sims = []
    for api in (oa, nemo_amd):
        conf = api.Configuration()
        if stdp is not None:
            conf.set_stdp_function(*stdp)
        sims.append(sc.make(api, spec, conf))
    return sims


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f":
        # NaN payloads differ between x86 and the GPU; any NaN equals any NaN
        an, bn = np.isnan(a), np.isnan(b)
        return np.array_equal(an, bn) and np.array_equal(a.view(np.uint32)[~an], b.view(np.uint32)[~bn])
    return np.array_equal(a, b)


def state_planes(sim, idx, nstate):
    return [np.array([sim.get_neuron_state(int(i), v) for i in idx], np.float32) for v in range(nstate)]


# ------------------------------------------------------------ device math
@pytest.fixture(scope="module")
def host_libm():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "ref.c")
    so = os.path.join(d, "libref.so")
    open(src, "w").write(
        "#include <math.h>\n#include <stddef.h>\n"
        "void eval(int fn, const float* in, float* out, size_t n) {\n"
        "  for (size_t i = 0; i < n; ++i) out[i] = fn == 0 ? logf(in[i]) : fn == 1 ? sinf(in[i]) : expf(in[i]);\n}\n")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", src, "-o", so, "-lm"], check=True)
    l = ctypes.CDLL(so)
    l.eval.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]
    return l


def _eval(fn, x, dev, host_libm):
    out = np.empty_like(x)
    if dev:
        from nemo_amd._lib import check, lib
        check(lib().nemo_hip_eval_math(fn, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), x.size))
    else:
        host_libm.eval(fn, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), x.size)
    return out


def float_u32_domain():
    parts = [np.arange(1 << 24, dtype=np.uint64)]
    for k in range(24, 32):
        parts.append(np.arange(1 << k, 1 << (k + 1), 1 << (k - 23), dtype=np.uint64))
    return np.concatenate(parts).astype(np.float32)


def test_device_logf_sinf_match_host_over_box_muller_domain(host_libm):
    f = float_u32_domain()
    b = (f * np.float32(2.0 ** -32)).astype(np.float32)
    x = np.unique((np.float32(1) - b).astype(np.float32))
    x = x[x > 0]
    assert same_bits(_eval(0, x, True, host_libm), _eval(0, x, False, host_libm))
    a = (f * np.float32(1.4629180792671596810513378043098e-9)).astype(np.float32)
    assert same_bits(_eval(1, a, True, host_libm), _eval(1, a, False, host_libm))


def test_device_expf_matches_host_over_hh_range(host_libm):
    u = np.arange(0, 0x41800000, 7, dtype=np.uint64).astype(np.uint32)
    x = u.view(np.float32)
    x = np.concatenate([x, -x])
    assert same_bits(_eval(2, x, True, host_libm), _eval(2, x, False, host_libm))


# -------------------------------------------------------- reference KATs
@pytest.mark.parametrize("ncount,delay", [(1000, 1), (1024, 1), (2000, 1), (4000, 1), (2000, 20), (2000, 63)])
def test_ring(ncount, delay):
    sc.run_ring(nemo_amd, ncount, delay)


def test_double_ring():
    sc.run_double_ring(nemo_amd)


def test_fixed_point_saturation():
    hip = sc.fixed_point_saturation(nemo_amd)
    ref = sc.fixed_point_saturation(oa)
    assert hip == ref


def test_firing_stimulus():
    sc.firing_stimulus(nemo_amd)


def test_current_stimulus():
    sc.current_stimulus(nemo_amd)


def test_invalid_current_stimulus():
    sc.invalid_current_stimulus(nemo_amd, nemo_amd.NemoError)


@pytest.mark.parametrize("n0,nstep", [(1, 1), (1000000, 4)])
def test_non_contiguous_indices(n0, nstep):
    sc.non_contiguous(nemo_amd, n0, nstep, seconds=1)


@pytest.mark.parametrize("others,iz_first", [(0, True), (1000, True), (2000, False)])
def test_neuron_type_mixture(others, iz_first):
    sc.neuron_type_mixture(nemo_amd, others, iz_first)


@pytest.mark.parametrize("noise,reward", [(False, 1.0), (True, 1.0), (True, 0.9)])
def test_stdp_closed_form(noise, reward):
    hip = sc.stdp_closed_form(nemo_amd, noise, reward)
    ref = sc.stdp_closed_form(oa, noise, reward)
    for h, r in zip(hip, ref):
        assert np.array_equal(h, r)


# ------------------------------------------------ random-network parity
def test_random_network_step_parity():
    """1000 Izhikevich neurons x 100 synapses, delays 1..20, Gaussian input
    sigma 5/2, 1000 steps (the small benchmark network): raster, currents,
    state, RNG and history bit-exact, step by step."""
    spec = gen.random_izhikevich(1000, 100, dmax=20, seed=11)
    ref, hip = both(spec)
    hip.set_debug(True)
    assert np.array_equal(ref.local_to_global(), hip.local_to_global())
    for t in range(1000):
        fr = ref.step()
        fh = hip.step()
        assert fr == fh, ("raster diverged", t)
        if t % 50 == 0 or t < 5:
            er, ir = ref.currents()
            eh, ih = hip.currents()
            assert same_bits(er, eh) and same_bits(ir, ih), ("currents diverged", t)
    idx = np.arange(1000)
    for pr, ph in zip(state_planes(ref, idx, 2), state_planes(hip, idx, 2)):
        assert same_bits(pr, ph)
    assert np.array_equal(ref.rng_state(), hip.rng_state())
    assert np.array_equal(ref.recent_firing(), hip.recent_firing())


def stdp_fn():
    pre, post = gen.stdp_exp_function(0.1, -0.12, 20.0, 20)
    return (pre, post, 0.0, 10.0, 0.0, -10.0)


@pytest.mark.parametrize("n,m,dmax,steps,period", [(2000, 200, 20, 1200, 300), (5000, 100, 20, 600, 200),
                                                    (1500, 1000, 20, 400, 150), (1000, 1500, 12, 300, 100)])
def test_stdp_network_run_parity(n, m, dmax, steps, period):
    """Izhikevich network with plastic excitatory synapses (the STDP benchmark
    rule, reduced size): the run() graph path against the oracle stepped one
    step at a time, applyStdp(1.0) every `period` steps.  Raster, weights and
    accumulators bit-exact.  Rows of about 10, 5, 50 and 125 synapses per
    (source, delay): short rows, and rows a delivery wave takes whole, over
    one or several passes of 64 events."""
    spec = gen.random_izhikevich(n, m, dmax=dmax, seed=5, stdp=True)
    ref, hip = both(spec, stdp_fn())
    cr, fr, er = ref.run(steps, stdp_period=period, reward=1.0)
    ch, fh, eh = hip.run(steps, stdp_period=period, reward=1.0)
    assert np.array_equal(cr, ch)
    assert np.array_equal(fr, fh)
    assert er == eh
    nsyn = spec.synapse_count
    assert np.array_equal(ref.fixed_weights(nsyn), hip.fixed_weights())
    assert np.array_equal(ref.stdp_accumulators(nsyn), hip.stdp_accumulators())
    assert cr.sum() > 0


def test_dense_input_tile_parity():
    """600 neurons whose 800 synapses each all target the first tile (delays
    1..3): tiles take far more events per step than the per-event index of
    the gather holds, and delivery blocks more than their bins, so the
    searched-bucket paths and the overflow accumulators carry the input.
    Rasters and currents bit-exact with the oracle over 150 steps."""
    n, m = 600, 800
    spec = gen.random_izhikevich(n, 0, seed=17)
    rng = np.random.default_rng(17)
    spec.source = np.repeat(np.arange(n, dtype=np.uint32), m)
    spec.target = rng.integers(0, 256, n * m).astype(np.uint32)
    spec.delay = rng.integers(1, 4, n * m).astype(np.uint32)
    spec.weight = rng.uniform(-2.0, 1.0, n * m).astype(np.float32)
    spec.plastic = np.zeros(n * m, np.uint8)
    ref, hip = both(spec)
    hip.set_debug(True)
    fired = 0
    for t in range(150):
        fr = ref.step()
        fh = hip.step()
        assert fr == fh, ("raster diverged", t)
        fired += len(fh)
        if t % 25 == 24:
            er, ir = ref.currents()
            eh, ih = hip.currents()
            assert same_bits(er, eh) and same_bits(ir, ih), ("currents diverged", t)
    assert fired > 150


def test_run_matches_step():
    spec = gen.random_izhikevich(3000, 150, dmax=15, seed=9, stdp=True)
    a = sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn()))
    b = sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn()))
    ca, fa, _ = a.run(300, stdp_period=100)
    fb = []
    cb = []
    for t in range(300):
        f = b.step()
        cb.append(len(f))
        fb.extend(f)
        if (t + 1) % 100 == 0:
            b.apply_stdp(1.0)
    assert np.array_equal(ca, cb)
    assert np.array_equal(fa, fb)
    assert np.array_equal(a.fixed_weights(), b.fixed_weights())


def test_interleaved_runs_steps_and_apply_match_oracle():
    """Runs of several graphs, single steps, applyStdp between them, a raster-
    less run and a state change: the step base that consecutive graphs move
    on themselves, the host's copy of it (used by the flush after each run)
    and the STDP batches must stay right whatever comes between the runs.
    Raster, weights and accumulators bit-exact with the oracle doing the same."""
    spec = gen.random_izhikevich(3000, 120, dmax=12, seed=21, stdp=True)
    ref, hip = both(spec, stdp_fn())
    nsyn = spec.synapse_count

    def same():
        assert np.array_equal(ref.fixed_weights(nsyn), hip.fixed_weights())
        assert np.array_equal(ref.stdp_accumulators(nsyn), hip.stdp_accumulators())

    def run(n, period=0, raster=True):
        cr, fr, _ = ref.run(n, stdp_period=period, reward=1.0, raster=raster)
        ch, fh, _ = hip.run(n, stdp_period=period, reward=1.0, raster=raster)
        if raster:
            assert np.array_equal(cr, ch) and np.array_equal(fr, fh)

    def steps(n):
        for _ in range(n):
            assert np.array_equal(ref.step(), hip.step())

    run(150)
    steps(3)
    run(70, period=35)
    same()
    for s_ in (ref, hip):
        s_.apply_stdp(0.5)
    run(1)
    run(130, raster=False)
    steps(2)
    for s_ in (ref, hip):
        s_.set_neuron_state(5, 1, -70.0)  # v of neuron 5
    run(64)
    run(65)
    same()
    steps(1)


def test_run_readout_longer_than_expected_matches_step():
    """A run's fired list is compacted on the device and copied with the
    counts up to an expected length; a network that fires far more (4000
    Poisson sources at 0.5 per step, added in a shuffled index order, so the
    index map is not linear) takes the rest-of-list copy and grows the host
    buffer, and still equals the lists of single steps."""
    idx = np.random.default_rng(3).permutation(4000).astype(np.uint32)

    def make():
        net = nemo_amd.Network()
        t = net.add_neuron_type("PoissonSource")
        net.add_neuron(t, idx, np.full(idx.size, 0.5, np.float32))
        return nemo_amd.Simulation(net, nemo_amd.Configuration())

    a, b = make(), make()
    ca, fa, _ = a.run(130)
    ca2, fa2, _ = a.run(70)
    fb, cb = [], []
    for _ in range(200):
        f = b.step()
        cb.append(len(f))
        fb.extend(f)
    assert np.array_equal(np.concatenate([ca, ca2]), cb)
    assert np.array_equal(np.concatenate([fa, fa2]), fb)
    assert ca.sum() > 64 * 1024  # more than the host buffer first held for a run of 64 steps


def _conf(api, stdp):
    c = api.Configuration()
    c.set_stdp_function(*stdp)
    return c


def test_mixed_models_parity():
    """QIF first (so the reference's QIF indexing is in bounds), then
    IF_curr_exp, HH, Poisson sources and Input neurons, randomly connected:
    raster and all state variables bit-exact over 300 steps."""
    rng = np.random.default_rng(4)
    blocks = []
    base = 0

    def block(model, n, args):
        nonlocal base
        idx = np.arange(base, base + n, dtype=np.uint32)
        base += n
        blocks.append(gen.NeuronBlock(model, idx, [np.asarray(a, np.float32) * np.ones(n, np.float32) for a in args]))
        return idx

    q = block("QIF", 300, [2.0, 0.0])
    lif = block("IF_curr_exp", 400, [-65.0, -65.0, 1.0, 20.0, 2.0, 5.0, 5.0, -50.0, 0.9, -65.0, 0.0, 0.0, 100.0])
    hh = block("HH", 50, [-65.0, 0.3177, 0.0529, 0.5961, 0.0])
    po = block("PoissonSource", 100, [0.01])
    inp = block("Input", 20, [])
    n = base
    m = 40
    src = np.repeat(np.arange(n, dtype=np.uint32), m)
    tgt = rng.integers(0, n - 120, size=src.size, dtype=np.uint32)  # no synapses onto sources
    dly = rng.integers(1, 11, size=src.size, dtype=np.uint32)
    w = rng.random(src.size, dtype=np.float32) * np.where(rng.random(src.size) < 0.8, 1.0, -1.0).astype(np.float32)
    spec = gen.NetworkSpec(blocks=blocks, source=src, target=tgt, delay=dly, weight=w.astype(np.float32),
                           plastic=np.zeros(src.size, np.uint8))
    ref, hip = both(spec)
    hip.set_debug(True)
    for t in range(300):
        fs = list(inp[rng.random(inp.size) < 0.1])
        istim = [(int(i), 2.0) for i in lif[rng.random(lif.size) < 0.05]]
        fr = ref.step(fstim=fs, istim=istim)
        fh = hip.step(fstim=fs, istim=istim)
        assert fr == fh, ("raster diverged", t)
        er, ir = ref.currents()
        eh, ih = hip.currents()
        assert same_bits(er, eh) and same_bits(ir, ih), ("currents diverged", t)
    for idx, ns in ((q, 1), (lif, 4), (hh, 5)):
        for pr, ph in zip(state_planes(ref, idx, ns), state_planes(hip, idx, ns)):
            assert same_bits(pr, ph)
    assert np.array_equal(ref.rng_state(), hip.rng_state())


def test_events_match_raster_full_size():
    """At the full 100k x 1000 STDP size, where the oracle is too slow to step
    through a test, check a size-independent property: events delivered equal
    the out-degrees implied by the raster (every firing of n at step s delivers
    row (n, d) at step s + d), and firing stays in a physiological range."""
    n, m, D = 100000, 1000, 20
    spec = gen.random_izhikevich(n, m, dmax=D, seed=2, stdp=True)
    sim = sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn()))
    steps = 200
    counts, fired, events = sim.run(steps, stdp_period=100)
    rowlen = np.zeros((n, D + 1), np.int64)
    np.add.at(rowlen, (spec.source.astype(np.int64), spec.delay.astype(np.int64)), 1)
    cum = np.cumsum(rowlen, axis=1)  # cum[n, d] = synapses of n with delay <= d
    step_of = np.repeat(np.arange(steps), counts)
    horizon = np.minimum(steps - 1 - step_of, D)
    expected = int(cum[fired.astype(np.int64), horizon].sum())
    assert events == expected
    rate_hz = counts.sum() / n / (steps / 1000.0)
    assert 0.5 < rate_hz < 200


# ------------------------------------------------------------ batches
def _mixed_spec(seed):
    """Poisson sources (added first) driving Izhikevich neurons: a network
    whose types come in a different order from the Izhikevich-only ones."""
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


def test_batch_matches_solo_runs():
    """Networks of different sizes, types and STDP settings advanced as one
    batch: each network's raster is bit-exact with its own simulation, and the
    weights after applyStdp equal the solo runs' (networks back to back)."""
    specs = gen.batch_izhikevich(count=6, nmin=200, nmax=3000, smin=20, smax=200, seed=11)
    specs.insert(2, (_mixed_spec(3), False))
    conf = _conf(nemo_amd, stdp_fn())
    nets = []
    for spec, _ in specs:
        net = nemo_amd.Network()
        spec.load(net)
        nets.append(net)
    batch = nemo_amd.Batch(nets, conf)
    assert batch.size == len(specs)
    steps, period = 300, 100
    counts, fired, events = batch.run(steps, stdp_period=period, reward=1.0)
    per = nemo_amd.Batch.split(counts, fired)
    solo_events = 0
    weights = []
    for k, ((spec, stdp), net) in enumerate(zip(specs, nets)):
        solo = nemo_amd.Simulation(net, _conf(nemo_amd, stdp_fn()) if stdp else nemo_amd.Configuration())
        c, f, e = solo.run(steps, stdp_period=period if stdp else 0, reward=1.0)
        assert np.array_equal(per[k][0], c), ("counts", k)
        assert np.array_equal(per[k][1], f), ("fired", k)
        assert c.sum() > 0
        solo_events += e
        weights.append(solo.fixed_weights())
        iz_idx = spec.blocks[-1].idx
        for i in iz_idx[:: max(1, iz_idx.size // 7)]:
            assert same_bits(np.float32(batch.get_membrane_potential(int(i), net=k)),
                             np.float32(solo.get_membrane_potential(int(i))))
    assert events == solo_events
    assert np.array_equal(batch.fixed_weights(), np.concatenate(weights))


def test_run_into_caller_buffer_and_spill():
    """nemo_hip_run_into writes fired neurons into the caller's buffer and the
    rest into the simulation's list: with a buffer of 7 entries (almost all
    spill), with none, and through Simulation.run (sized from the previous
    run), the fired lists and counts equal nemo_hip_run's."""
    from nemo_amd._lib import check, lib
    spec = gen.random_izhikevich(3000, 100, dmax=20, seed=17, stdp=True)
    sims = [sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn())) for _ in range(3)]
    c_uint_p = ctypes.POINTER(ctypes.c_uint)
    for nsteps in (100, 150, 70):
        ref_counts, ref_fired, ref_ev = sims[0].run(nsteps, stdp_period=100)
        for sim, cap in ((sims[1], 7), (sims[2], 0)):
            counts = np.zeros(nsteps, np.uint32)
            buf = np.zeros(max(cap, 1), np.uint32)
            used, nspill, ev = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint64()
            spill = c_uint_p()
            check(lib().nemo_hip_run_into(sim._h, nsteps, 100, 1.0, counts.ctypes.data_as(c_uint_p),
                                          buf.ctypes.data_as(c_uint_p), cap, ctypes.byref(used),
                                          ctypes.byref(spill), ctypes.byref(nspill), ctypes.byref(ev)))
            rest = np.ctypeslib.as_array(spill, shape=(nspill.value,)) if nspill.value else np.zeros(0, np.uint32)
            fired = np.concatenate([buf[:used.value], rest])
            assert used.value == min(cap, fired.size)
            assert np.array_equal(counts, ref_counts), (nsteps, cap)
            assert np.array_equal(fired, ref_fired), (nsteps, cap)
            assert ev.value == ref_ev
        assert ref_counts.sum() > 0


def test_single_type_batch_matches_solo_runs():
    """A batch of one neuron type lists each network's neurons in turn on the
    device, and its fired lists are decoded in one pass: each network's raster
    must still equal its own simulation's."""
    specs = gen.batch_izhikevich(count=5, nmin=200, nmax=2000, smin=20, smax=100, seed=12)
    nets = []
    for spec, _ in specs:
        net = nemo_amd.Network()
        spec.load(net)
        nets.append(net)
    batch = nemo_amd.Batch(nets, _conf(nemo_amd, stdp_fn()))
    steps = 200
    counts, fired, events = batch.run(steps, stdp_period=100, reward=1.0)
    per = nemo_amd.Batch.split(counts, fired)
    solo_events = 0
    for k, ((spec, stdp), net) in enumerate(zip(specs, nets)):
        solo = nemo_amd.Simulation(net, _conf(nemo_amd, stdp_fn()) if stdp else nemo_amd.Configuration())
        c, f, e = solo.run(steps, stdp_period=100 if stdp else 0, reward=1.0)
        assert np.array_equal(per[k][0], c), ("counts", k)
        assert np.array_equal(per[k][1], f), ("fired", k)
        assert c.sum() > 0
        solo_events += e
    assert events == solo_events


def test_launch_shapes_give_identical_results():
    """Delivery blocks of both sets, bin capacity (down to 64 events, so that
    most events take the overflow path), STDP placement and block width, and
    the overlap of the next step's delivery with the update (including the
    update's own push of its delay-1 events) change only speed: rasters,
    weights and accumulators are identical."""
    spec = gen.random_izhikevich(4000, 150, dmax=15, seed=21, stdp=True)
    shapes = [None, (1, 1, 64, 0, 1, 1), (7, 3, 256, 1, 4, 0), (200, 0, 1024, 2, 2, 0),
              (900, 124, 4096, 2, 1, 1), (33, 5, 2048, 1, 3, 1), (64, 0, 512, 0, 2, 0),
              (64, 0, 512, 3, 2, 0), (33, 5, 2048, 3, 1, 1), (33, 5, 2048, 2, 3, 2), (64, 16, 1024, 1, 2, 2),
              (200, 40, 4096, 0, 4, 2), (7, 3, 256, 3, 1, 2), (120, 30, 512, 2, 4, 2), (64, 0, 512, 4, 2, 0),
              (33, 5, 2048, 4, 3, 2), (200, 0, 1024, 4, 8, 1), (120, 0, 1024, 4, 32, 1), (90, 0, 2048, 4, 64, 0),
              (33, 5, 2048, 4, 3, 3), (7, 3, 256, 3, 1, 3), (1, 1, 64, 2, 2, 3), (64, 16, 1024, 1, 2, 3),
              (120, 30, 512, 0, 4, 3)]
    results = []
    for shape in shapes:
        sim = sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn()))
        if shape is not None:
            sim.launch_shape = shape
            assert sim.launch_shape == shape
        c, f, e = sim.run(250, stdp_period=100)
        c2, f2, e2 = sim.run(67)  # chunks of 64 and 3 steps
        single = [sim.step() for _ in range(5)]
        results.append((np.concatenate([c, c2]), np.concatenate([f, f2]), e + e2, sim.fixed_weights(),
                        sim.stdp_accumulators(), single))
    for shape, r in zip(shapes[1:], results[1:]):
        for name, x, y in zip(("counts", "fired", "events", "weights", "accumulators"), results[0][:5], r[:5]):
            assert np.array_equal(x, y), (shape, name)
        assert results[0][5] == r[5], (shape, "single steps")
    assert results[0][0].sum() > 0


@pytest.mark.parametrize("which", ["random", "torus"])
def test_ranged_table_reads_give_identical_results(which, monkeypatch):
    """The update reads either a tile's whole table row or only the columns of
    the delivery blocks whose windows hold the tile (NEMO_HIP_BIN_RANGED): the
    same events either way, over several launch shapes (both delivery sets,
    small bins with overflow, co-launched delivery, pushed delay-1 events)."""
    if which == "random":
        spec = gen.random_izhikevich(4000, 150, dmax=15, seed=21, stdp=True)
    else:
        spec = gen.torus_mixed(n_lif=4000, n_qif=800, n_hh=200, n_poisson=100, m=50, grid=15)
    stdp = which == "random"
    shapes = [None, (1, 1, 64, 0, 1, 1), (33, 5, 2048, 4, 3, 2), (7, 3, 256, 3, 1, 2), (200, 0, 1024, 4, 8, 1),
              (33, 5, 2048, 4, 3, 3), (7, 3, 256, 3, 1, 3)]
    results = []
    for ranged in ("0", "1"):
        monkeypatch.setenv("NEMO_HIP_BIN_RANGED", ranged)
        for shape in shapes:
            sim = sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn()) if stdp else None)
            if shape is not None:
                sim.launch_shape = shape
            c, f, e = sim.run(150, stdp_period=100 if stdp else 0)
            results.append(((ranged, shape), c, f, e))
    for key, c, f, e in results[1:]:
        assert np.array_equal(results[0][1], c), (key, "counts")
        assert np.array_equal(results[0][2], f), (key, "fired")
        assert results[0][3] == e, (key, "events")
    assert results[0][1].sum() > 0


def test_autotune_keeps_the_simulation_state():
    spec = gen.random_izhikevich(5000, 200, dmax=20, seed=8, stdp=True)
    a = sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn()))
    b = sc.make(nemo_amd, spec, _conf(nemo_amd, stdp_fn()))
    ca1, fa1, ea1 = a.run(150, stdp_period=100)
    shape = a.autotune(32)
    assert a.launch_shape == shape
    ca2, fa2, ea2 = a.run(150, stdp_period=100)
    cb1, fb1, eb1 = b.run(150, stdp_period=100)
    cb2, fb2, eb2 = b.run(150, stdp_period=100)
    assert np.array_equal(np.concatenate([ca1, ca2]), np.concatenate([cb1, cb2]))
    assert np.array_equal(np.concatenate([fa1, fa2]), np.concatenate([fb1, fb2]))
    assert (ea1, ea2) == (eb1, eb2)
    assert np.array_equal(a.fixed_weights(), b.fixed_weights())
    assert np.array_equal(a.stdp_accumulators(), b.stdp_accumulators())


# ------------------------------------------------------- mixed-model torus
def test_torus_mixed_parity():
    """The mixed-model torus network (IF_curr_exp, QIF, HH, Poisson; distance
    dependent delays) at a tenth of its size: rasters, currents and every
    state variable bit-exact with the oracle over 300 steps."""
    spec = gen.torus_mixed(n_lif=4000, n_qif=800, n_hh=200, n_poisson=100, m=50, grid=15)
    ref, hip = both(spec)
    hip.set_debug(True)
    for t in range(300):
        fr = ref.step()
        fh = hip.step()
        assert fr == fh, ("raster diverged", t)
        if t % 50 == 49:
            er, ir = ref.currents()
            eh, ih = hip.currents()
            assert same_bits(er, eh) and same_bits(ir, ih), ("currents diverged", t)
    for blk, ns in zip(spec.blocks[:3], (4, 1, 5)):
        for pr, ph in zip(state_planes(ref, blk.idx[::7], ns), state_planes(hip, blk.idx[::7], ns)):
            assert same_bits(pr, ph)
    assert np.array_equal(ref.rng_state(), hip.rng_state())


def test_torus_mixed_full_size_events():
    """The full 51k-neuron, 500-synapse torus (2.55e7 synapses) for 200 steps:
    events delivered equal the out-degrees implied by the raster."""
    spec = gen.torus_mixed()
    sim = sc.make(nemo_amd, spec)
    steps = 200
    D = int(spec.delay.max())
    n = spec.neuron_count
    counts, fired, events = sim.run(steps)
    rowlen = np.zeros((n, D + 1), np.int64)
    np.add.at(rowlen, (spec.source.astype(np.int64), spec.delay.astype(np.int64)), 1)
    cum = np.cumsum(rowlen, axis=1)
    step_of = np.repeat(np.arange(steps), counts)
    horizon = np.minimum(steps - 1 - step_of, D)
    assert events == int(cum[fired.astype(np.int64), horizon].sum())
    assert counts.sum() > 0
