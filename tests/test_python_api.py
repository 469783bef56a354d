"""NeMo's Python API tests (src/api/python/test.py:50-384) on nemo_amd.

The reference module is imported as `nemo`; this module is a drop-in for
it.  Network-only cases run without a GPU; the ones that build a Simulation
are marked gpu.  Failures surface as RuntimeError, as in the reference.
"""
import random

import numpy as np
import pytest

import nemo_amd as nemo


class IzNetwork(nemo.Network):
    """test.py:7-15: Izhikevich neurons added through the named-parameter
    dictionaries, positionally."""

    def __init__(self):
        nemo.Network.__init__(self)
        self._type = self.add_neuron_type('Izhikevich')

    def add_neuron(self, nidx, a, b, c, d, sigma, u, v):
        nemo.Network.add_neuron(self, self._type, nidx,
                                {'a': a, 'b': b, 'c': c, 'd': d, 'sigma': sigma}, {'u': u, 'v': v})


def random_source():
    return random.randint(0, 999)


def random_delay():
    return random.randint(1, 20)


def random_weight():
    return random.uniform(-1.0, 1.0)


def random_plastic():
    return random.choice([True, False])


def arg(vlen, gen):
    """a fixed-length list or a scalar, with values drawn from gen"""
    if random.choice([True, False]):
        return [gen() for _ in range(vlen)]
    return gen()


@pytest.fixture(autouse=True)
def _seed():
    random.seed(1)


def check_neuron_function(fun, ncount):
    vlen = random.randint(2, ncount)
    a, b, c, d, u, v, s = (arg(vlen, random.random) for _ in range(7))
    if any(isinstance(x, list) for x in (a, b, c, d, u, v, s)):
        fun(range(vlen), a, b, c, d, s, u, v)
    else:
        fun(random.randint(0, 1000), a, b, c, d, s, u, v)


def check_set_neuron_vector(obj, pop, tests=1000):
    for _ in range(tests):
        vlen = random.randint(2, 100)
        vector = random.choice([True, False])
        if vector:
            neuron = random.sample(pop, vlen)
            value = [random.random() for _ in neuron]
        else:
            neuron = random.choice(pop)
            value = random.random()

        def same(value, ret):
            if vector:
                assert len(ret) == vlen
                assert all(abs(x - y) < 1e-5 for x, y in zip(value, ret))
            else:
                assert abs(value - ret) < 1e-5

        param = random.randint(0, 4)
        obj.set_neuron_parameter(neuron, param, value)
        same(value, obj.get_neuron_parameter(neuron, param))
        var = random.randint(0, 1)
        obj.set_neuron_state(neuron, var, value)
        same(value, obj.get_neuron_state(neuron, var))


# ------------------------------------------------------------ network only
def test_network_set_neuron():
    """test.py:52-122"""
    a, b = 0.02, 0.2
    c, d = -65.0 + 15.0 * 0.25, 8.0 - 6.0 * 0.25
    v = -65.0
    u = b * v
    sigma = 5.0
    net = IzNetwork()
    with pytest.raises(RuntimeError):
        net.set_neuron(0, a, b, c, d, sigma, u, v)
    net.add_neuron(0, a, b, c - 0.1, d, sigma, u, v - 1.0)
    for bad in ((net.get_neuron_parameter, 1, 0), (net.get_neuron_state, 1, 0),
                (net.get_neuron_parameter, 0, 5), (net.get_neuron_state, 0, 2)):
        with pytest.raises(RuntimeError):
            bad[0](*bad[1:])
    e = 0.1
    net.set_neuron(0, a - e, b - e, c - e, d - e, sigma - e, u - e, v - e)
    places = 5
    for p, x in enumerate((a - e, b - e, c - e, d - e, sigma - e)):
        assert round(net.get_neuron_parameter(0, p) - x, places) == 0
    for k, x in enumerate((u - e, v - e)):
        assert round(net.get_neuron_state(0, k) - x, places) == 0
    for p, x in enumerate((a, b, c, d, sigma)):
        net.set_neuron_parameter(0, p, x)
        assert round(net.get_neuron_parameter(0, p) - x, places) == 0
    for k, x in enumerate((u, v)):
        net.set_neuron_state(0, k, x)
        assert round(net.get_neuron_state(0, k) - x, places) == 0
    for bad in ((net.set_neuron_parameter, 1, 0, 0.0), (net.set_neuron_state, 1, 0, 0.0),
                (net.set_neuron_parameter, 0, 5, 0.0), (net.set_neuron_state, 0, 2, 0.0)):
        with pytest.raises(RuntimeError):
            bad[0](*bad[1:])


def test_add_neuron():
    """test.py:138-146: scalar and vector forms of add_neuron"""
    for _ in range(1000):
        net = IzNetwork()
        check_neuron_function(net.add_neuron, ncount=1000)


def test_network_set_neuron_forms():
    """test.py:148-161, network half"""
    net = IzNetwork()
    net.add_neuron(range(1000), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    for _ in range(1000):
        check_neuron_function(net.set_neuron, ncount=1000)


def test_network_set_neuron_vector():
    """test.py:200-212"""
    net = IzNetwork()
    pop = list(range(1000))
    for n in pop:
        net.add_neuron(n, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    check_set_neuron_vector(net, pop)


def test_add_synapse():
    """test.py:291-311: scalar and vector forms of add_synapse; a list of ids
    for vector input, one id for scalar input"""
    net = IzNetwork()
    for _ in range(1000):
        vlen = random.randint(2, 500)
        source = arg(vlen, random_source)
        target = arg(vlen, random_source)
        delay = arg(vlen, random_delay)
        weight = arg(vlen, random_weight)
        plastic = arg(vlen, random_plastic)
        ids = net.add_synapse(source, target, delay, weight, plastic)
        if any(isinstance(n, list) for n in (source, target, delay, weight, plastic)):
            assert isinstance(ids, list) and len(ids) == vlen
        else:
            assert not isinstance(ids, list)


def test_named_parameters_of_other_models():
    """wrappers.py:72-155: IF_curr_exp and PoissonSource take the reference's
    key names; QIF and HH (no entry in the reference's table) take the key
    order of their .ini file; Input takes none."""
    net = nemo.Network()
    lif = net.add_neuron_type("IF_curr_exp")
    params = dict(v_rest=-65.0, v_reset=-70.0, c_m=1.0, tau_m=20.0, tau_refrac=5.0, tau_synE=2.0, tau_synI=5.0,
                  v_thresh=-51.0, I_offset=0.1)
    net.add_neuron(lif, 0, params, dict(v=-65.0, ie=0.0, ii=0.0, lastfired=1000.0))
    assert [net.get_neuron_parameter(0, p) for p in range(9)] == \
        [np.float32(x) for x in (-65.0, -70.0, 1.0, 20.0, 5.0, 2.0, 5.0, -51.0, 0.1)]
    assert net.get_neuron_state(0, 3) == 1000.0
    po = net.add_neuron_type("PoissonSource")
    net.add_neuron(po, [1, 2], {"lambda": [0.01, 0.02]}, {})
    assert net.get_neuron_parameter(2, 0) == np.float32(0.02)
    qif = net.add_neuron_type("QIF")
    net.add_neuron(qif, 3, {"a": 2.0}, {"v": 0.25})
    assert (net.get_neuron_parameter(3, 0), net.get_neuron_state(3, 0)) == (2.0, 0.25)
    hh = net.add_neuron_type("HH")
    net.add_neuron(hh, 4, {}, {"v": -65.0, "n": 0.3, "m": 0.05, "h": 0.6, "dir": 0.0})
    assert [net.get_neuron_state(4, k) for k in range(5)] == [np.float32(x) for x in (-65.0, 0.3, 0.05, 0.6, 0.0)]
    inp = net.add_neuron_type("Input")
    net.add_neuron(inp, 5, {}, {})
    with pytest.raises(KeyError):
        net.add_neuron(lif, 6, {"v_rest": -65.0}, {})
    with pytest.raises(RuntimeError):
        net.add_neuron(17, 7, {}, {})
    assert net.neuron_count() == 6


# ----------------------------------------------------------- simulations
def simple_network():
    net = IzNetwork()
    net.add_neuron(0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    net.add_neuron(1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    net.add_synapse(0, 1, 1, 5.0, False)
    net.add_synapse(1, 0, 1, 5.0, False)
    return net, nemo.Simulation(net, nemo.Configuration())


@pytest.mark.gpu
def test_simulation_set_neuron_forms(require_gpu):
    """test.py:148-161, simulation half"""
    net = IzNetwork()
    net.add_neuron(range(1000), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    sim = nemo.Simulation(net, nemo.Configuration())
    for _ in range(1000):
        check_neuron_function(sim.set_neuron, ncount=1000)


@pytest.mark.gpu
def test_sim_set_neuron_vector(require_gpu):
    """test.py:214-228"""
    net = IzNetwork()
    pop = list(range(1000))
    for n in pop:
        net.add_neuron(n, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    sim = nemo.Simulation(net, nemo.Configuration())
    check_set_neuron_vector(sim, pop)


@pytest.mark.gpu
def test_get_neuron_scalar(require_gpu):
    """test.py:238-252"""
    def check(x):
        x.get_neuron_state([0], 0)
        x.get_neuron_state(0, 0)
        x.get_neuron_parameter([0], 0)
        x.get_neuron_parameter(0, 0)
    net, sim = simple_network()
    check(net)
    check(sim)


@pytest.mark.gpu
def test_set_neuron_scalar(require_gpu):
    """test.py:254-268"""
    def check(x):
        x.set_neuron_state([0], 0, [0])
        x.set_neuron_state(0, 0, 0)
        x.set_neuron_parameter([0], 0, [0])
        x.set_neuron_parameter(0, 0, 0)
    net, sim = simple_network()
    check(net)
    check(sim)


@pytest.mark.gpu
def test_get_synapse_scalar(require_gpu):
    """test.py:270-289"""
    def check(x):
        for q in (x.get_synapse_source, x.get_synapse_target, x.get_synapse_delay, x.get_synapse_weight,
                  x.get_synapse_plastic):
            q(0)
            q([0])
    net, sim = simple_network()
    check(net)
    check(sim)


@pytest.mark.gpu
def test_get_synapses_from_unconnected(require_gpu):
    """test.py:314-319"""
    net = IzNetwork()
    net.add_neuron(0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    assert len(net.get_synapses_from(0)) == 0
    sim = nemo.Simulation(net, nemo.Configuration())
    assert len(sim.get_synapses_from(0)) == 0


@pytest.mark.gpu
def test_get_synapse(require_gpu):
    """test.py:321-380: synapse properties are functions of source and
    target; read them back, scalar and vector, in random order"""
    def delay(source, target):
        return 1 + ((source + target) % 20)

    def plastic(source, target):
        return (source + target) % 1 == 0

    def weight(source, target):
        return float(source) + float(target)

    ncount = 100
    net = IzNetwork()
    for src in range(ncount):
        net.add_neuron(src, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
        for tgt in range(src + 1):
            net.add_synapse(src, tgt, delay(src, tgt), weight(src, tgt), plastic(src, tgt))
    sim = nemo.Simulation(net, nemo.Configuration())

    def check_scalar(x, known_source, sid, source, target):
        assert known_source == source
        assert x.get_synapse_delay(sid) == delay(source, target)
        assert x.get_synapse_weight(sid) == weight(source, target)
        assert x.get_synapse_plastic(sid) == plastic(source, target)

    def check(x):
        for src in range(ncount):
            all_synapses = x.get_synapses_from(src)
            n_queried = random.randint(1, len(all_synapses))
            queried = random.sample(all_synapses, n_queried)
            if len(queried) == 1:
                queried = queried[0]
            sources = x.get_synapse_source(queried)
            targets = x.get_synapse_target(queried)
            if n_queried == 1:
                check_scalar(x, src, queried, sources, targets)
            else:
                for sid, qsrc, tgt in zip(queried, sources, targets):
                    check_scalar(x, src, sid, qsrc, tgt)

    check(net)
    check(sim)
    for x in (net, sim):
        for src in range(ncount):
            x.get_synapse_source(x.get_synapses_from(src))
