"""Neuron-model descriptions (<Model>.ini) in NeMo's format.

NeMo reads parameters, state-variables, membrane-potential and rng.normal
(required), history, the [rcm] flags and the backend flags, and ignores the
[params:...] and [state_names] sections, whose values are default
expressions for its tools (src/nemo/NeuronType.cpp:93-158).  The HIP
backend reads the same keys plus backends.hip.  No GPU needed: only the
network is built.
"""
import glob
import os

import pytest

import nemo_amd

REFERENCE_PLUGINS = "/root/reference/src/nemo/plugins"


def write_model(tmp_path, name, text):
    d = tmp_path / "plugins"
    d.mkdir(exist_ok=True)
    (d / (name + ".ini")).write_text(text)
    nemo_amd.add_plugin_path(str(d))


def test_reference_dialect_loads(tmp_path):
    """Default expressions, random generators and booleans in the default
    sections do not disturb the parse."""
    write_model(tmp_path, "DialectIzhikevich", """\
# an Izhikevich description in NeMo's own dialect
parameters=5
state-variables=2
membrane-potential=1
history=1

[rcm]
sources=false
delays=false
forward=false
weights=false

[rng]
normal=true

[backends]
cpu=true
cuda=true
hip=true

[params:Excitatory]
a=0.02
c=-65+15*RANDF()**2

[params:Inhibitory]
a=0.02+0.08*RANDF()

[state_names]
v=-65
u=b*v
w=true
""")
    net = nemo_amd.Network()
    t = net.add_neuron_type("DialectIzhikevich")
    net.add_neuron(t, 0, 0.02, 0.2, -65.0, 8.0, 5.0, -13.0, -65.0)
    with pytest.raises(nemo_amd.NemoError):
        net.add_neuron(t, 1, 0.02, 0.2, -65.0, 8.0, 5.0, -13.0)  # one argument short
    assert net.get_neuron_state(0, 1) == -65.0


@pytest.mark.parametrize("drop", ["parameters", "state-variables", "membrane-potential", "normal"])
def test_required_keys(tmp_path, drop):
    text = "parameters=1\nstate-variables=1\nmembrane-potential=0\n[rng]\nnormal=false\n[backends]\nhip=true\n"
    text = "\n".join(l for l in text.splitlines() if not l.startswith(drop + "="))
    name = "Missing_" + drop.replace("-", "_")
    write_model(tmp_path, name, text)
    net = nemo_amd.Network()
    with pytest.raises(nemo_amd.NemoError) as e:
        net.add_neuron_type(name)
    assert e.value.code == 6  # NEMO_INVALID_INPUT
    assert "Missing parameter" in str(e.value)


def test_unparsable_value(tmp_path):
    write_model(tmp_path, "BadCount", "parameters=two\nstate-variables=1\nmembrane-potential=0\n[rng]\nnormal=no\n"
                "[backends]\nhip=yes\n")
    with pytest.raises(nemo_amd.NemoError) as e:
        nemo_amd.Network().add_neuron_type("BadCount")
    assert e.value.code == 6


def test_model_without_hip_backend(tmp_path):
    write_model(tmp_path, "CpuOnly", "parameters=1\nstate-variables=1\nmembrane-potential=0\n[rng]\nnormal=false\n"
                "[backends]\ncpu=true\ncuda=true\n")
    with pytest.raises(nemo_amd.NemoError) as e:
        nemo_amd.Network().add_neuron_type("CpuOnly")
    assert e.value.code == 5  # NEMO_API_UNSUPPORTED


def test_packaged_descriptions_are_in_nemo_format():
    """The six packaged models load, and each .ini holds the keys NeMo's own
    parser requires."""
    for path in glob.glob(os.path.join(nemo_amd.plugin_dir(), "*.ini")):
        text = open(path).read()
        for key in ("parameters=", "state-variables=", "membrane-potential=", "normal="):
            assert key in text, (path, key)
        net = nemo_amd.Network()
        net.add_neuron_type(os.path.basename(path)[:-4])


@pytest.mark.skipif(not os.path.isdir(REFERENCE_PLUGINS), reason="NeMo's source tree is not present")
@pytest.mark.parametrize("model,nargs", [("Izhikevich", 7), ("QIF", 2), ("IF_curr_exp", 13), ("HH", 5),
                                         ("PoissonSource", 1), ("Input", 0)])
def test_nemo_descriptions_load(tmp_path, model, nargs):
    """NeMo's own description of each model, with backends.hip added, loads
    and gives the parameter and state counts NeMo's parser would."""
    text = open(os.path.join(REFERENCE_PLUGINS, model + ".ini")).read()
    write_model(tmp_path, "NeMo" + model, text + "\n[backends]\nhip=true\n")
    net = nemo_amd.Network()
    t = net.add_neuron_type("NeMo" + model)
    net.add_neuron(t, 0, *([0.0] * nargs))
    with pytest.raises(nemo_amd.NemoError):
        net.add_neuron(t, 1, *([0.0] * (nargs + 1)))
