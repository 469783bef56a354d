"""nemo_amd -- Python interface of the MI355X HIP backend for NeMo.

Mirrors NeMo's Python module (src/api/python/wrappers.py:27-331 over
nemo_py.cpp): Configuration, Network and Simulation with the same method
names and argument meaning, plus set_hip_backend (the device selection the
HIP backend adds) and Simulation.run (many steps per host round trip).
Calls go through the C ABI of libnemo_hip.so (include/nemo_hip.h).
"""
import ctypes
import os

import numpy as np

from ._lib import NemoError, check, lib

__all__ = ["Configuration", "Network", "Simulation", "Batch", "NemoError", "version", "hip_device_count",
           "hip_device_description", "add_plugin_path", "plugin_dir"]

_HERE = os.path.dirname(os.path.abspath(__file__))


def plugin_dir():
    """Directory holding <Model>.ini and hip/lib<Model>.so."""
    return os.path.join(_HERE, "plugins")


def version():
    return lib().nemo_version().decode()


def add_plugin_path(path):
    check(lib().nemo_add_plugin_path(path.encode()))
    _extra_paths.append(path)


def hip_device_count():
    n = ctypes.c_uint(0)
    check(lib().nemo_hip_device_count(ctypes.byref(n)))
    return n.value


def hip_device_description(dev):
    s = ctypes.c_char_p()
    check(lib().nemo_hip_device_description(dev, ctypes.byref(s)))
    return s.value.decode()


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


# Named parameters and state variables of each model, in index order: the
# names NeMo's Python module maps (src/api/python/wrappers.py:72-150).
_NAMED_ARGS = {
    "Izhikevich": (("a", "b", "c", "d", "sigma"), ("u", "v")),
    "IF_curr_exp": (("v_rest", "v_reset", "c_m", "tau_m", "tau_refrac", "tau_synE", "tau_synI", "v_thresh",
                     "I_offset"), ("v", "ie", "ii", "lastfired")),
    "PoissonSource": (("lambda",), ()),
    "Input": ((), ()),
    "Kuramoto": (("omega", "Cmean"), ("theta",)),
    "KuramotoLag": (("omega", "Cmean", "alpha"), ("theta",)),
    "IzhikevichChaos": (("a", "b", "c", "d", "sigma"), ("u", "v")),
    # NeMo's module passes these three models' state as v, u, ge, gi although
    # their plugins read index 0 as u and 1 as v (plugins/Izhikevich2007.h,
    # IzhikevichSyn.h); a drop-in keeps the module's mapping
    "IzhikevichSyn": (("a", "b", "c", "d", "vpeak", "sigma", "tMdtOt_exc", "G_exc", "E_exc", "tMdtOt_inh", "G_inh",
                       "E_inh"), ("v", "u", "ge", "gi")),
    "Izhikevich2007": (("a", "b", "c", "d", "vpeak", "Vr", "Vt", "k", "Cm", "sigma", "tMdtOt_exc", "G_exc", "E_exc",
                        "tMdtOt_inh", "G_inh", "E_inh"), ("v", "u", "ge", "gi")),
    "Izhikevich2007DopSyn": (("a", "b", "c", "d", "vpeak", "Vr", "Vt", "k", "Cm", "sigma", "d1", "d2", "tMdtOt_exc",
                              "G_exc", "E_exc", "tMdtOt_inh", "G_inh", "E_inh"), ("v", "u", "ge", "gi")),
    "BayesianNeuron": (("rOn", "rOff", "g0"), ("L", "G")),
    "InhomogeneousPoissonSource": ((), ()),
}


def _ini_names(model):
    """Parameter and state names of a model: the table above, or else the key
    order of the first [params:...] section and of the [state_names] section
    of its .ini file (which NeMo's .ini files list in index order for the
    other models)."""
    if model in _NAMED_ARGS:
        return _NAMED_ARGS[model]
    path = next((os.path.join(d, model + ".ini") for d in _search_paths()
                 if os.path.exists(os.path.join(d, model + ".ini"))), None)
    if path is None:
        return (), ()
    params, state, section, params_section = [], [], "", None
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            if line.startswith("[") and line.endswith("]"):
                section = line[1:-1]
                if section.startswith("params:") and params_section is None:
                    params_section = section
                continue
            if "=" not in line:
                continue
            key = line.split("=", 1)[0].strip()
            if section == params_section:
                params.append(key)
            elif section == "state_names":
                state.append(key)
    return tuple(params), tuple(state)


_extra_paths = []


def _search_paths():
    """Directories searched for <Model>.ini, in the library's order: added
    paths, NEMO_PLUGIN_PATH, then the package's plugins directory."""
    env = [p for p in os.environ.get("NEMO_PLUGIN_PATH", "").split(":") if p]
    return _extra_paths + env + [plugin_dir()]


def _is_seq(x):
    return isinstance(x, (list, tuple, range, np.ndarray))


class Configuration:
    """Simulation configuration (src/nemo/Configuration.hpp:70-195)."""

    def __init__(self):
        self._conf = lib().nemo_new_configuration()

    def __del__(self):
        if getattr(self, "_conf", None):
            lib().nemo_delete_configuration(self._conf)
            self._conf = None

    def set_stdp_function(self, prefire, postfire, minWeight, maxWeight, minInhibitory=None, maxInhibitory=None):
        """STDP function sampled per ms.  With four weights the arguments are
        (minE, maxE, minI, maxI); with two they are (minWeight, maxWeight) as in
        Configuration::setStdpFunction (Configuration.cpp:148-151): excitatory
        [0, maxWeight], inhibitory [minWeight, 0]."""
        if minInhibitory is None:
            minE, maxE, minI, maxI = 0.0, maxWeight, 0.0, minWeight
        else:
            minE, maxE, minI, maxI = minWeight, maxWeight, minInhibitory, maxInhibitory
        pre, post = _f32(prefire), _f32(postfire)
        check(lib().nemo_set_stdp_function(self._conf, _ptr(pre, ctypes.c_float), pre.size,
                                           _ptr(post, ctypes.c_float), post.size, minE, maxE, minI, maxI))

    def set_hip_backend(self, device=-1):
        check(lib().nemo_set_hip_backend(self._conf, device))

    def set_cuda_backend(self, device=-1):
        """The GPU backend choice of the reference; selects the HIP device."""
        self.set_hip_backend(device)

    def set_cpu_backend(self):
        check(lib().nemo_set_cpu_backend(self._conf))

    def backend_description(self):
        s = ctypes.c_char_p()
        check(lib().nemo_backend_description(self._conf, ctypes.byref(s)))
        return s.value.decode()

    def set_write_only_synapses(self):
        check(lib().nemo_set_write_only_synapses(self._conf))

    def log_stdout(self):
        check(lib().nemo_log_stdout(self._conf))


class _SynapseQueries:
    """Synapse getters shared by Network (_n) and Simulation (_s)."""

    def get_synapses_from(self, source):
        ptr = ctypes.POINTER(ctypes.c_uint64)()
        n = ctypes.c_size_t(0)
        check(getattr(lib(), "nemo_get_synapses_from_" + self._sfx)(self._h, source, ctypes.byref(ptr), ctypes.byref(n)))
        return [ptr[i] for i in range(n.value)]

    def _get(self, what, ctype, ids):
        f = getattr(lib(), "nemo_get_synapse_%s_%s" % (what, self._sfx))
        scalar = np.isscalar(ids)
        out = []
        for i in ([ids] if scalar else ids):
            v = ctype()
            check(f(self._h, int(i), ctypes.byref(v)))
            out.append(v.value)
        return out[0] if scalar else out

    def get_synapse_source(self, ids):
        return self._get("source", ctypes.c_uint, ids)

    def get_synapse_target(self, ids):
        return self._get("target", ctypes.c_uint, ids)

    def get_synapse_delay(self, ids):
        return self._get("delay", ctypes.c_uint, ids)

    def get_synapse_weight(self, ids):
        return self._get("weight", ctypes.c_float, ids)

    def get_synapse_plastic(self, ids):
        return self._get("plastic", ctypes.c_ubyte, ids)


class Network(_SynapseQueries):
    """Network under construction (src/nemo/Network.hpp:45-231)."""

    _sfx = "n"

    def __init__(self):
        self._h = lib().nemo_new_network()
        self._types = {}

    def __del__(self):
        if getattr(self, "_h", None):
            lib().nemo_delete_network(self._h)
            self._h = None

    def add_neuron_type(self, name):
        t = ctypes.c_uint(0)
        check(lib().nemo_add_neuron_type(self._h, name.encode(), ctypes.byref(t)))
        self._types[t.value] = name
        return t.value

    def add_neuron(self, neuron_type, neuron_idx, *args, paramDict=None, stateDict=None):
        """Add one neuron or a vector of neurons.  As in NeMo's Python module
        (wrappers.py:72-155), the model's parameters and state variables are
        given as two dictionaries of named values, positionally or by
        keyword: add_neuron(type, idx, paramDict, stateDict).  Alternatively
        args are the parameters followed by the state variables, in index
        order.  Each value is a scalar or a sequence as long as neuron_idx."""
        if neuron_type not in self._types:
            raise RuntimeError("That neuron type is not in this network. Use add_neuron_type first.")
        if args and isinstance(args[0], dict):
            paramDict = args[0]
            stateDict = args[1] if len(args) > 1 else {}
            args = ()
        if paramDict is not None or stateDict is not None:
            pnames, snames = _ini_names(self._types[neuron_type])
            paramDict, stateDict = paramDict or {}, stateDict or {}
            missing = [k for k in pnames if k not in paramDict] + [k for k in snames if k not in stateDict]
            if missing:
                raise KeyError("missing neuron parameters or state variables: " + ", ".join(missing))
            args = tuple(paramDict[k] for k in pnames) + tuple(stateDict[k] for k in snames)
        idx = np.atleast_1d(_u32(neuron_idx))
        n = idx.size
        cols = [np.broadcast_to(np.asarray(a, dtype=np.float32), (n,)) for a in args]
        mat = np.ascontiguousarray(np.stack(cols, axis=1) if cols else np.zeros((n, 0), np.float32))
        check(lib().nemo_add_neurons(self._h, neuron_type, n, _ptr(idx, ctypes.c_uint), len(cols),
                                     _ptr(mat, ctypes.c_float)))

    def add_synapse(self, source, target, delay, weight, plastic):
        """Add one synapse (returns its id) or vectors of synapses (returns a
        numpy array of ids); scalars broadcast against vectors."""
        scalar = all(np.isscalar(x) for x in (source, target, delay, weight, plastic))
        as_list = any(isinstance(x, (list, tuple, range)) for x in (source, target, delay, weight, plastic))
        n = max(np.size(x) for x in (source, target, delay, weight, plastic))
        s = np.ascontiguousarray(np.broadcast_to(_u32(source), (n,)))
        t = np.ascontiguousarray(np.broadcast_to(_u32(target), (n,)))
        d = np.ascontiguousarray(np.broadcast_to(_u32(delay), (n,)))
        w = np.ascontiguousarray(np.broadcast_to(_f32(weight), (n,)))
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(plastic, dtype=np.uint8), (n,)))
        ids = np.empty(n, dtype=np.uint64)
        check(lib().nemo_add_synapses(self._h, n, _ptr(s, ctypes.c_uint), _ptr(t, ctypes.c_uint),
                                      _ptr(d, ctypes.c_uint), _ptr(w, ctypes.c_float),
                                      _ptr(p, ctypes.c_ubyte), _ptr(ids, ctypes.c_uint64)))
        if scalar:
            return int(ids[0])
        return ids.tolist() if as_list else ids

    def neuron_count(self):
        n = ctypes.c_uint(0)
        check(lib().nemo_neuron_count(self._h, ctypes.byref(n)))
        return n.value

    def _get1(self, fn, idx, k):
        v = ctypes.c_float()
        check(fn(self._h, int(idx), k, ctypes.byref(v)))
        return v.value

    def get_neuron_state(self, idx, var):
        """State variable var of a neuron, or a list of them for a list of
        neurons."""
        if _is_seq(idx):
            return [self._get1(lib().nemo_get_neuron_state_n, i, var) for i in idx]
        return self._get1(lib().nemo_get_neuron_state_n, idx, var)

    def get_neuron_parameter(self, idx, param):
        if _is_seq(idx):
            return [self._get1(lib().nemo_get_neuron_parameter_n, i, param) for i in idx]
        return self._get1(lib().nemo_get_neuron_parameter_n, idx, param)

    def set_neuron(self, idx, *args):
        """Replace the parameters and state of one neuron, or of a list of
        neurons (each argument then a scalar or a list as long)."""
        idx, mat = _neuron_rows(idx, args)
        for i, row in zip(idx, mat):
            check(lib().nemo_set_neuron_n(self._h, int(i), row.size, _ptr(row, ctypes.c_float)))

    def set_neuron_state(self, idx, var, val):
        idx, vals = _pairs(idx, val)
        for i, v in zip(idx, vals):
            check(lib().nemo_set_neuron_state_n(self._h, int(i), var, float(v)))

    def set_neuron_parameter(self, idx, param, val):
        idx, vals = _pairs(idx, val)
        for i, v in zip(idx, vals):
            check(lib().nemo_set_neuron_parameter_n(self._h, int(i), param, float(v)))


def _neuron_rows(idx, args):
    """(indices, n x nargs float32 rows) from a neuron or list of neurons and
    per-argument scalars or lists."""
    idx = np.atleast_1d(_u32(idx))
    cols = [np.broadcast_to(np.asarray(a, dtype=np.float32), (idx.size,)) for a in args]
    mat = np.ascontiguousarray(np.stack(cols, axis=1) if cols else np.zeros((idx.size, 0), np.float32))
    return idx, mat


def _pairs(idx, val):
    idx = np.atleast_1d(_u32(idx))
    return idx, np.ascontiguousarray(np.broadcast_to(np.asarray(val, dtype=np.float32), (idx.size,)))


class Simulation(_SynapseQueries):
    """A HIP simulation of a network (src/nemo/Simulation.hpp:38-218)."""

    _sfx = "s"

    def __init__(self, net, conf):
        h = lib().nemo_new_simulation(net._h, conf._conf)
        if not h:
            raise NemoError(-1, lib().nemo_strerror().decode())
        self._adopt(h)

    def _adopt(self, h):
        self._h = h
        n = ctypes.c_size_t()
        s = ctypes.c_size_t()
        p = ctypes.c_size_t()
        check(lib().nemo_hip_counts(self._h, ctypes.byref(n), ctypes.byref(s), ctypes.byref(p)))
        self.neuron_count, self.synapse_count, self.plastic_count = n.value, s.value, p.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().nemo_delete_simulation(self._h)
            self._h = None

    def step(self, fstim=None, istim=None):
        """One 1 ms step; fstim: neurons forced to fire, istim: list of
        (neuron, current) pairs.  Returns the fired neurons in simulation order
        (wrappers.py:252-276)."""
        if istim and (not isinstance(istim[0], tuple) or len(istim[0]) != 2):
            raise ValueError("Current input must be specified as a list of (idx, I) pairs, "
                             "e.g. [(idx1, I1), (idx2, I2), ...]")
        f = _u32(fstim if fstim is not None else [])
        if istim:
            ii = _u32([x[0] for x in istim])
            iv = _f32([x[1] for x in istim])
        else:
            ii, iv = _u32([]), _f32([])
        out = ctypes.POINTER(ctypes.c_uint)()
        n = ctypes.c_size_t(0)
        check(lib().nemo_step(self._h, _ptr(f, ctypes.c_uint), f.size, _ptr(ii, ctypes.c_uint),
                              _ptr(iv, ctypes.c_float), ii.size, ctypes.byref(out), ctypes.byref(n)))
        return [out[i] for i in range(n.value)]

    def step_noinput(self):
        return self.step()

    def step_f(self, fstim):
        return self.step(fstim=fstim)

    def step_i(self, istim):
        return self.step(istim=istim)

    def step_fi(self, fstim, istim):
        return self.step(fstim=fstim, istim=istim)

    def run(self, nsteps, stdp_period=0, reward=1.0, raster=True):
        """Run nsteps steps on the device (no stimulus), applying STDP with the
        given reward every stdp_period steps.  Returns (counts, fired, events):
        per-step fired counts, the concatenated fired neurons (each step in
        simulation order) and the synaptic events delivered."""
        counts = np.zeros(nsteps, dtype=np.uint32) if raster else None
        fired, events = self._run(nsteps, stdp_period, reward, counts)
        return counts, fired, events

    def _run(self, nsteps, stdp_period, reward, counts):
        """nemo_hip_run, or with counts (the raster read back into it)
        nemo_hip_run_into: the fired neurons are decoded on the host beside
        the later steps straight into an array sized from the previous run's
        firing rate, and any beyond it are appended from the simulation's own
        list.  Returns (fired or None, events)."""
        ev = ctypes.c_uint64(0)
        if counts is None:
            out = ctypes.POINTER(ctypes.c_uint)()
            n = ctypes.c_size_t(0)
            check(lib().nemo_hip_run(self._h, nsteps, stdp_period, reward, None, ctypes.byref(out),
                                     ctypes.byref(n), ctypes.byref(ev)))
            return None, ev.value
        rate = getattr(self, "_fired_per_step", 0.0)
        cap = int(1.25 * rate * nsteps) + 1024 if rate else 0
        buf = np.empty(cap, np.uint32)
        used = ctypes.c_size_t(0)
        spill = ctypes.POINTER(ctypes.c_uint)()
        nspill = ctypes.c_size_t(0)
        check(lib().nemo_hip_run_into(self._h, nsteps, stdp_period, reward, _ptr(counts, ctypes.c_uint),
                                      _ptr(buf, ctypes.c_uint), cap, ctypes.byref(used), ctypes.byref(spill),
                                      ctypes.byref(nspill), ctypes.byref(ev)))
        fired = buf[:used.value]
        if nspill.value:
            fired = np.concatenate([fired, np.ctypeslib.as_array(spill, shape=(nspill.value,))])
        if nsteps:
            self._fired_per_step = (used.value + nspill.value) / nsteps
        return fired, ev.value

    def apply_stdp(self, reward):
        check(lib().nemo_apply_stdp(self._h, reward))

    def propagate(self, fired):
        """Brian's propagate (Simulation.hpp:93-113): the given neurons fire
        in the next step instead of the model deciding; their spikes are
        delivered and the excitatory and inhibitory input of the step after
        is returned, two float32 arrays in simulation order.  Neuron state is
        untouched.  Without STDP only."""
        f = _u32(fired if fired is not None else [])
        n = self.neuron_count
        E = np.zeros(n, np.float32)
        I = np.zeros(n, np.float32)
        check(lib().nemo_hip_propagate(self._h, _ptr(f, ctypes.c_uint), f.size, _ptr(E, ctypes.c_float),
                                       _ptr(I, ctypes.c_float), n))
        return E, I

    def prepare(self, nsteps, stdp_period=0):
        """Capture the step graphs run(nsteps, stdp_period) uses (so that a timed
        run excludes graph construction)."""
        check(lib().nemo_hip_prepare(self._h, nsteps, stdp_period))

    @property
    def launch_shape(self):
        """(delivery blocks for delays 2..D -- all delays when the next is 0 --,
        delivery blocks for delay 1, bin capacity, STDP mode (0 own branch, 1
        inside the updates, 2 in the delivery launch, 3 own branch joined one
        step later, 4 batched after each run of up to 64 steps), bitmap words
        per STDP block, and how the delays 2..D are
        delivered beside the previous update: 0 not, 1 on a graph branch, 2 as
        extra blocks of the update launch)."""
        a = (ctypes.c_uint * 6)()
        check(lib().nemo_hip_get_launch_shape(self._h, a))
        return tuple(a)

    @launch_shape.setter
    def launch_shape(self, shape):
        a = (ctypes.c_uint * 6)(*shape)
        check(lib().nemo_hip_set_launch_shape(self._h, a))

    def autotune(self, steps_per_trial=128):
        """Pick the fastest launch shape by timing candidates from a snapshot of
        the current state (restored afterwards); returns the chosen shape."""
        a = (ctypes.c_uint * 6)()
        check(lib().nemo_hip_autotune(self._h, steps_per_trial, a))
        return tuple(a)

    def get_membrane_potential(self, idx):
        v = ctypes.c_float()
        check(lib().nemo_get_membrane_potential(self._h, idx, ctypes.byref(v)))
        return v.value

    def _get_values(self, fn, idx, k):
        seq = _is_seq(idx)
        ii = np.atleast_1d(_u32(idx))
        out = np.zeros(ii.size, np.float32)
        check(fn(self._h, ii.size, _ptr(ii, ctypes.c_uint), k, _ptr(out, ctypes.c_float)))
        return [float(x) for x in out] if seq else float(out[0])

    def get_neuron_state(self, idx, var):
        """State variable var of a neuron, or a list of them for a list of
        neurons."""
        return self._get_values(lib().nemo_get_neuron_states_s, idx, var)

    def get_neuron_parameter(self, idx, param):
        return self._get_values(lib().nemo_get_neuron_parameters_s, idx, param)

    def set_neuron(self, idx, *args):
        """Replace the parameters and state of one neuron, or of a list of
        neurons (each argument then a scalar or a list as long)."""
        ii, mat = _neuron_rows(idx, args)
        check(lib().nemo_set_neurons_s(self._h, ii.size, _ptr(ii, ctypes.c_uint), mat.shape[1],
                                       _ptr(mat, ctypes.c_float)))

    def set_neuron_state(self, idx, var, val):
        ii, vals = _pairs(idx, val)
        check(lib().nemo_set_neuron_states_s(self._h, ii.size, _ptr(ii, ctypes.c_uint), var,
                                             _ptr(vals, ctypes.c_float)))

    def set_neuron_parameter(self, idx, param, val):
        ii, vals = _pairs(idx, val)
        check(lib().nemo_set_neuron_parameters_s(self._h, ii.size, _ptr(ii, ctypes.c_uint), param,
                                                 _ptr(vals, ctypes.c_float)))

    def elapsed_wallclock(self):
        v = ctypes.c_ulong()
        check(lib().nemo_elapsed_wallclock(self._h, ctypes.byref(v)))
        return v.value

    def elapsed_simulation(self):
        v = ctypes.c_ulong()
        check(lib().nemo_elapsed_simulation(self._h, ctypes.byref(v)))
        return v.value

    def reset_timer(self):
        check(lib().nemo_reset_timer(self._h))

    def profile(self, nsteps):
        """Mean time (us) of the launches of a step -- "propagate" (delivery
        of the step and STDP accumulation of the previous one, per launch) and
        "update" (all neuron models, per step; with launch-shape overlap 2 or
        3 it also delivers the next step) -- and "propagate_launches" per step
        (below 1 with overlap 3, whose steps after the first of each run of 64
        need none), over nsteps eager steps measured with HIP events on the
        simulation's stream, and the work done."""
        avg = (ctypes.c_double * 4)()
        cnt = (ctypes.c_uint64 * 4)()
        check(lib().nemo_hip_profile(self._h, nsteps, avg, cnt))
        keys = ("propagate", "update", "propagate_launches")
        return ({k: avg[i] for i, k in enumerate(keys)},
                {"events": cnt[0], "fired": cnt[1], "stdp_synapses": cnt[2], "steps": cnt[3]})

    # ---- parity inspection (extensions) ----
    def set_debug(self, on=True):
        check(lib().nemo_hip_set_debug(self._h, 1 if on else 0))

    def currents(self):
        n = self.neuron_count
        e = np.zeros(n, np.float32)
        i = np.zeros(n, np.float32)
        check(lib().nemo_hip_get_currents(self._h, _ptr(e, ctypes.c_float), _ptr(i, ctypes.c_float), n))
        return e, i

    def local_to_global(self):
        g = np.zeros(self.neuron_count, np.uint32)
        check(lib().nemo_hip_local_to_global(self._h, _ptr(g, ctypes.c_uint), g.size))
        return g

    def recent_firing(self):
        h = np.zeros(self.neuron_count, np.uint64)
        check(lib().nemo_hip_get_recent_firing(self._h, _ptr(h, ctypes.c_uint64), h.size))
        return h

    def rng_state(self):
        s = np.zeros(4 * self.neuron_count, np.uint32)
        check(lib().nemo_hip_get_rng_state(self._h, _ptr(s, ctypes.c_uint32), self.neuron_count))
        return s.reshape(-1, 4)

    def neuron_states(self, var):
        """State variable var of every neuron, in simulation order (NaN where
        the neuron's model has fewer state variables)."""
        out = np.zeros(self.neuron_count, np.float32)
        check(lib().nemo_hip_get_neuron_states(self._h, var, _ptr(out, ctypes.c_float), out.size))
        return out

    def fixed_weights(self):
        w = np.zeros(self.synapse_count, np.int32)
        check(lib().nemo_hip_get_fixed_weights(self._h, _ptr(w, ctypes.c_int32), w.size))
        return w

    def stdp_accumulators(self):
        a = np.zeros(self.synapse_count, np.int32)
        check(lib().nemo_hip_get_stdp_accumulators(self._h, _ptr(a, ctypes.c_int32), a.size))
        return a


class Batch(Simulation):
    """Independent networks advanced together in the same kernel launches
    (nemo_hip_new_batch).  Every network behaves exactly as in a Simulation of
    its own: same RNG streams, same firing."""

    def __init__(self, nets, conf):
        arr = (ctypes.c_void_p * len(nets))(*[n._h for n in nets])
        h = ctypes.c_void_p()
        check(lib().nemo_hip_new_batch(arr, len(nets), conf._conf, ctypes.byref(h)))
        self._adopt(h.value)
        k = ctypes.c_uint()
        check(lib().nemo_hip_batch_size(self._h, ctypes.byref(k)))
        self.size = k.value

    def run(self, nsteps, stdp_period=0, reward=1.0, raster=True):
        """As Simulation.run; counts has shape (nsteps, size) and fired lists,
        step by step, each network's fired neurons in turn."""
        counts = np.zeros((nsteps, self.size), dtype=np.uint32) if raster else None
        fired, events = self._run(nsteps, stdp_period, reward, counts)
        return counts, fired, events

    @staticmethod
    def split(counts, fired):
        """Per network (counts, fired) in the format of Simulation.run."""
        flat = counts.ravel()
        ends = np.cumsum(flat.astype(np.int64))
        starts = ends - flat
        k = counts.shape[1]
        out = []
        for j in range(k):
            idx = np.arange(j, flat.size, k)
            parts = [fired[starts[i]:ends[i]] for i in idx]
            out.append((counts[:, j].copy(), np.concatenate(parts) if parts else np.zeros(0, np.uint32)))
        return out

    def get_membrane_potential(self, idx, net=0):
        v = ctypes.c_float()
        check(lib().nemo_hip_batch_get_membrane_potential(self._h, net, idx, ctypes.byref(v)))
        return v.value

    def get_neuron_state(self, idx, var, net=0):
        v = ctypes.c_float()
        check(lib().nemo_hip_batch_get_neuron_state(self._h, net, idx, var, ctypes.byref(v)))
        return v.value

    def get_neuron_parameter(self, idx, param, net=0):
        v = ctypes.c_float()
        check(lib().nemo_hip_batch_get_neuron_parameter(self._h, net, idx, param, ctypes.byref(v)))
        return v.value
