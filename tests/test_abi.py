"""The C-ABI library loads and exports every entry point include/*.h declares;
the host-side network API works without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import nemo_amd
from nemo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    skip = {"typedef", "defined", "sizeof", "if", "return"}
    return sorted({n for n in names if n not in skip and (n.startswith("nemo_") or n.startswith("hip_"))})


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions("nemo_hip.h")
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_the_header():
    names = set(declared_functions("nemo_hip.h"))
    protos = {p[0] for p in _lib._PROTOTYPES}
    assert names == protos, (names - protos, protos - names)


@pytest.mark.parametrize("model", ["Izhikevich", "QIF", "IF_curr_exp", "HH", "PoissonSource", "Input"])
def test_plugins_export_the_plugin_abi(model):
    path = os.path.join(nemo_amd.plugin_dir(), "hip", "lib%s.so" % model)
    lib = ctypes.CDLL(path)
    assert hasattr(lib, "hip_update_neurons") and hasattr(lib, "hip_init_neurons")
    assert os.path.exists(os.path.join(nemo_amd.plugin_dir(), model + ".ini"))


def test_network_construction_and_queries():
    net = nemo_amd.Network()
    iz = net.add_neuron_type("Izhikevich")
    assert net.add_neuron_type("Izhikevich") == iz
    net.add_neuron(iz, [0, 1, 2], 0.02, 0.2, -65.0, 8.0, 5.0, -13.0, -65.0)
    assert net.neuron_count() == 3
    sid = net.add_synapse(0, 1, 3, 0.5, 1)
    ids = net.add_synapse(0, [2, 1], [1, 2], [-0.25, 0.75], 0)
    assert sid == 0 and list(ids) == [1, 2]
    assert net.get_synapses_from(0) == [0, 1, 2]
    assert isinstance(ids, list)  # list input gives a list of ids, as NeMo's module
    assert net.get_synapse_target(ids) == [2, 1]
    assert net.get_synapse_delay(sid) == 3
    assert net.get_synapse_weight(1) == -0.25
    assert net.get_synapse_plastic(sid) == 1
    assert net.get_neuron_parameter(1, 4) == 5.0
    net.set_neuron_state(2, 1, -70.0)
    assert net.get_neuron_state(2, 1) == -70.0


def test_network_named_parameters():
    net = nemo_amd.Network()
    t = net.add_neuron_type("IF_curr_exp")
    # the key names of NeMo's Python module (wrappers.py:117-121)
    params = dict(v_rest=-65.0, v_reset=-70.0, c_m=1.0, tau_m=20.0, tau_refrac=2.0, tau_synE=5.0,
                  tau_synI=5.0, v_thresh=-50.0, I_offset=0.1)
    net.add_neuron(t, 7, paramDict=params, stateDict=dict(v=-65.0, ie=0.0, ii=0.0, lastfired=100.0))
    assert net.get_neuron_parameter(7, 1) == -70.0
    assert net.get_neuron_parameter(7, 8) == np.float32(0.1)
    assert net.get_neuron_state(7, 3) == 100.0


@pytest.mark.parametrize("bad", [
    lambda n: n.add_neuron_type("NoSuchModel"),
    lambda n: n.add_neuron(n.add_neuron_type("Izhikevich"), 0, 1.0, 2.0),       # wrong arity
    lambda n: n.add_synapse(0, 1, 0, 1.0, 0),                                    # delay < 1
])
def test_network_errors(bad):
    with pytest.raises(nemo_amd.NemoError):
        bad(nemo_amd.Network())


def test_duplicate_neuron_rejected():
    net = nemo_amd.Network()
    iz = net.add_neuron_type("Izhikevich")
    net.add_neuron(iz, 4, *([0.0] * 7))
    with pytest.raises(nemo_amd.NemoError):
        net.add_neuron(iz, 4, *([0.0] * 7))


def test_configuration():
    conf = nemo_amd.Configuration()
    assert "HIP" in conf.backend_description()
    pre = np.exp(-np.arange(1, 21) / 20.0) * 0.1
    conf.set_stdp_function(pre, -pre, -1.0, 10.0)
    with pytest.raises(nemo_amd.NemoError):
        conf.set_stdp_function(pre, -pre, 0.01, 1.0, 0.01, -1.0)  # positive inhibitory limit
    conf.set_cpu_backend()
    assert "CPU" in conf.backend_description()


def test_cpu_backend_is_not_provided():
    net = nemo_amd.Network()
    iz = net.add_neuron_type("Izhikevich")
    net.add_neuron(iz, 0, *([0.0] * 7))
    conf = nemo_amd.Configuration()
    conf.set_cpu_backend()
    with pytest.raises(nemo_amd.NemoError) as e:
        nemo_amd.Simulation(net, conf)
    assert "HIP backend only" in str(e.value)
