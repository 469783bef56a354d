"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so) with the
same Python surface as nemo_amd (Network, Configuration, Simulation).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# NEMO_ORACLE_LIB: another build of the oracle (e.g. one with a planted
# mistake, to check that the pinning tests catch it)
ORACLE_LIB = os.environ.get("NEMO_ORACLE_LIB", os.path.join(ORACLE_DIR, "_build", "liboracle.so"))

_lib = None

P = ctypes.POINTER
vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-C", ORACLE_DIR], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        l = ctypes.CDLL(ORACLE_LIB)
        protos = [
            ("oracle_strerror", ctypes.c_char_p, []),
            ("oracle_new_network", vp, []),
            ("oracle_delete_network", None, [vp]),
            ("oracle_add_neuron_type", ctypes.c_int, [vp, ctypes.c_char_p, P(ctypes.c_uint)]),
            ("oracle_add_neurons", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_size_t, P(ctypes.c_uint),
                                                  ctypes.c_uint, P(ctypes.c_float)]),
            ("oracle_add_synapses", ctypes.c_int, [vp, ctypes.c_size_t, P(ctypes.c_uint), P(ctypes.c_uint),
                                                   P(ctypes.c_uint), P(ctypes.c_float), P(ctypes.c_ubyte)]),
            ("oracle_neuron_count", ctypes.c_int, [vp, P(ctypes.c_uint)]),
            ("oracle_new_configuration", vp, []),
            ("oracle_delete_configuration", None, [vp]),
            ("oracle_set_stdp_function", ctypes.c_int, [vp, P(ctypes.c_float), ctypes.c_size_t, P(ctypes.c_float),
                                                        ctypes.c_size_t] + [ctypes.c_float] * 4),
            ("oracle_new_simulation", ctypes.c_int, [vp, vp, P(vp)]),
            ("oracle_delete_simulation", None, [vp]),
            ("oracle_step", ctypes.c_int, [vp, P(ctypes.c_uint), ctypes.c_size_t, P(ctypes.c_uint),
                                           P(ctypes.c_float), ctypes.c_size_t, P(P(ctypes.c_uint)),
                                           P(ctypes.c_size_t)]),
            ("oracle_apply_stdp", ctypes.c_int, [vp, ctypes.c_float]),
            ("oracle_get_membrane_potential", ctypes.c_int, [vp, ctypes.c_uint, P(ctypes.c_float)]),
            ("oracle_get_neuron_state", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, P(ctypes.c_float)]),
            ("oracle_get_neuron_parameter", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, P(ctypes.c_float)]),
            ("oracle_set_neuron_state", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
            ("oracle_set_neuron_parameter", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
            ("oracle_get_synapses_from", ctypes.c_int, [vp, ctypes.c_uint, P(P(ctypes.c_uint64)),
                                                        P(ctypes.c_size_t)]),
            ("oracle_get_synapse_target", ctypes.c_int, [vp, ctypes.c_uint64, P(ctypes.c_uint)]),
            ("oracle_get_synapse_delay", ctypes.c_int, [vp, ctypes.c_uint64, P(ctypes.c_uint)]),
            ("oracle_get_synapse_weight", ctypes.c_int, [vp, ctypes.c_uint64, P(ctypes.c_float)]),
            ("oracle_get_synapse_plastic", ctypes.c_int, [vp, ctypes.c_uint64, P(ctypes.c_ubyte)]),
            ("oracle_elapsed_simulation", ctypes.c_int, [vp, P(ctypes.c_ulong)]),
            ("oracle_reset_timer", ctypes.c_int, [vp]),
            ("oracle_local_to_global", ctypes.c_int, [vp, P(ctypes.c_uint), ctypes.c_size_t]),
            ("oracle_get_currents", ctypes.c_int, [vp, P(ctypes.c_float), P(ctypes.c_float), ctypes.c_size_t]),
            ("oracle_get_fixed_weights", ctypes.c_int, [vp, P(ctypes.c_int32), ctypes.c_size_t]),
            ("oracle_get_stdp_accumulators", ctypes.c_int, [vp, P(ctypes.c_int32), ctypes.c_size_t]),
            ("oracle_get_recent_firing", ctypes.c_int, [vp, P(ctypes.c_uint64), ctypes.c_size_t]),
            ("oracle_get_rng_state", ctypes.c_int, [vp, P(ctypes.c_uint32), ctypes.c_size_t]),
            ("oracle_delivered_events", ctypes.c_uint64, [vp]),
            ("oracle_num_threads", ctypes.c_int, []),
            ("oracle_rng_seeds", None, [ctypes.c_uint, ctypes.c_uint, P(ctypes.c_uint32)]),
            ("oracle_urand", ctypes.c_uint32, [P(ctypes.c_uint32)]),
            ("oracle_nrand", ctypes.c_float, [P(ctypes.c_uint32)]),
            ("oracle_nrand_n", None, [P(ctypes.c_uint32), ctypes.c_size_t, P(ctypes.c_float)]),
            ("oracle_get_neuron_states", ctypes.c_int, [vp, ctypes.c_uint, P(ctypes.c_float), ctypes.c_size_t]),
            ("oracle_propagate", ctypes.c_int, [vp, P(ctypes.c_uint), ctypes.c_size_t, P(P(ctypes.c_float)),
                                                P(P(ctypes.c_float))]),
            ("oracle_mt19937_nth", ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint]),
            ("oracle_fx_tofix", ctypes.c_int, [ctypes.c_float, ctypes.c_uint, P(ctypes.c_int32)]),
            ("oracle_wfx_tofloat", ctypes.c_float, [ctypes.c_int64, ctypes.c_uint]),
        ]
        for name, res, args in protos:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def check(status):
    if status != 0:
        raise OracleError(status, lib().oracle_strerror().decode())


def _ptr(a, t):
    return a.ctypes.data_as(P(t))


class Configuration:
    def __init__(self):
        self._conf = lib().oracle_new_configuration()

    def __del__(self):
        if getattr(self, "_conf", None):
            lib().oracle_delete_configuration(self._conf)
            self._conf = None

    def set_stdp_function(self, prefire, postfire, minWeight, maxWeight, minInhibitory=None, maxInhibitory=None):
        if minInhibitory is None:
            minE, maxE, minI, maxI = 0.0, maxWeight, 0.0, minWeight
        else:
            minE, maxE, minI, maxI = minWeight, maxWeight, minInhibitory, maxInhibitory
        pre = np.ascontiguousarray(prefire, dtype=np.float32)
        post = np.ascontiguousarray(postfire, dtype=np.float32)
        check(lib().oracle_set_stdp_function(self._conf, _ptr(pre, ctypes.c_float), pre.size,
                                             _ptr(post, ctypes.c_float), post.size, minE, maxE, minI, maxI))

    def set_hip_backend(self, device=-1):
        pass

    def set_cpu_backend(self):
        pass


class Network:
    def __init__(self):
        self._h = lib().oracle_new_network()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_delete_network(self._h)
            self._h = None

    def add_neuron_type(self, name):
        t = ctypes.c_uint(0)
        check(lib().oracle_add_neuron_type(self._h, name.encode(), ctypes.byref(t)))
        return t.value

    def add_neuron(self, neuron_type, neuron_idx, *args):
        idx = np.atleast_1d(np.ascontiguousarray(neuron_idx, dtype=np.uint32))
        n = idx.size
        cols = [np.broadcast_to(np.asarray(a, dtype=np.float32), (n,)) for a in args]
        mat = np.ascontiguousarray(np.stack(cols, axis=1) if cols else np.zeros((n, 0), np.float32))
        check(lib().oracle_add_neurons(self._h, neuron_type, n, _ptr(idx, ctypes.c_uint), len(cols),
                                       _ptr(mat, ctypes.c_float)))

    def add_synapse(self, source, target, delay, weight, plastic):
        n = max(np.size(x) for x in (source, target, delay, weight, plastic))
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(source, np.uint32), (n,)))
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(target, np.uint32), (n,)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(delay, np.uint32), (n,)))
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(weight, np.float32), (n,)))
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(plastic, np.uint8), (n,)))
        check(lib().oracle_add_synapses(self._h, n, _ptr(s, ctypes.c_uint), _ptr(t, ctypes.c_uint),
                                        _ptr(d, ctypes.c_uint), _ptr(w, ctypes.c_float), _ptr(p, ctypes.c_ubyte)))

    def neuron_count(self):
        n = ctypes.c_uint(0)
        check(lib().oracle_neuron_count(self._h, ctypes.byref(n)))
        return n.value


class Simulation:
    def __init__(self, net, conf):
        h = vp()
        check(lib().oracle_new_simulation(net._h, conf._conf, ctypes.byref(h)))
        self._h = h
        self.neuron_count = net.neuron_count()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_delete_simulation(self._h)
            self._h = None

    def step(self, fstim=None, istim=None):
        f = np.ascontiguousarray(fstim if fstim is not None else [], dtype=np.uint32)
        if istim:
            ii = np.ascontiguousarray([x[0] for x in istim], dtype=np.uint32)
            iv = np.ascontiguousarray([x[1] for x in istim], dtype=np.float32)
        else:
            ii = np.zeros(0, np.uint32)
            iv = np.zeros(0, np.float32)
        out = P(ctypes.c_uint)()
        n = ctypes.c_size_t(0)
        check(lib().oracle_step(self._h, _ptr(f, ctypes.c_uint), f.size, _ptr(ii, ctypes.c_uint),
                                _ptr(iv, ctypes.c_float), ii.size, ctypes.byref(out), ctypes.byref(n)))
        return [out[i] for i in range(n.value)]

    def run(self, nsteps, stdp_period=0, reward=1.0, raster=True):
        counts = np.zeros(nsteps, np.uint32)
        fired = []
        ev0 = self.delivered_events()
        since = 0
        for s in range(nsteps):
            f = self.step()
            counts[s] = len(f)
            fired.extend(f)
            since += 1
            if stdp_period and since == stdp_period:
                self.apply_stdp(reward)
                since = 0
        return counts, np.asarray(fired, dtype=np.uint32), self.delivered_events() - ev0

    def apply_stdp(self, reward):
        check(lib().oracle_apply_stdp(self._h, reward))

    def get_membrane_potential(self, idx):
        v = ctypes.c_float()
        check(lib().oracle_get_membrane_potential(self._h, idx, ctypes.byref(v)))
        return v.value

    def get_neuron_state(self, idx, var):
        v = ctypes.c_float()
        check(lib().oracle_get_neuron_state(self._h, idx, var, ctypes.byref(v)))
        return v.value

    def get_neuron_parameter(self, idx, param):
        v = ctypes.c_float()
        check(lib().oracle_get_neuron_parameter(self._h, idx, param, ctypes.byref(v)))
        return v.value

    def set_neuron_state(self, idx, var, val):
        check(lib().oracle_set_neuron_state(self._h, idx, var, val))

    def set_neuron_parameter(self, idx, param, val):
        check(lib().oracle_set_neuron_parameter(self._h, idx, param, val))

    def get_synapses_from(self, source):
        ptr = P(ctypes.c_uint64)()
        n = ctypes.c_size_t(0)
        check(lib().oracle_get_synapses_from(self._h, source, ctypes.byref(ptr), ctypes.byref(n)))
        return [ptr[i] for i in range(n.value)]

    def _get(self, fn, ctype, sid):
        v = ctype()
        check(getattr(lib(), fn)(self._h, int(sid), ctypes.byref(v)))
        return v.value

    def get_synapse_target(self, sid):
        return self._get("oracle_get_synapse_target", ctypes.c_uint, sid)

    def get_synapse_delay(self, sid):
        return self._get("oracle_get_synapse_delay", ctypes.c_uint, sid)

    def get_synapse_weight(self, sid):
        return self._get("oracle_get_synapse_weight", ctypes.c_float, sid)

    def get_synapse_plastic(self, sid):
        return self._get("oracle_get_synapse_plastic", ctypes.c_ubyte, sid)

    def elapsed_simulation(self):
        v = ctypes.c_ulong()
        check(lib().oracle_elapsed_simulation(self._h, ctypes.byref(v)))
        return v.value

    def reset_timer(self):
        check(lib().oracle_reset_timer(self._h))

    def currents(self):
        n = self.neuron_count
        e = np.zeros(n, np.float32)
        i = np.zeros(n, np.float32)
        check(lib().oracle_get_currents(self._h, _ptr(e, ctypes.c_float), _ptr(i, ctypes.c_float), n))
        return e, i

    def local_to_global(self):
        g = np.zeros(self.neuron_count, np.uint32)
        check(lib().oracle_local_to_global(self._h, _ptr(g, ctypes.c_uint), g.size))
        return g

    def recent_firing(self):
        h = np.zeros(self.neuron_count, np.uint64)
        check(lib().oracle_get_recent_firing(self._h, _ptr(h, ctypes.c_uint64), h.size))
        return h

    def rng_state(self):
        s = np.zeros(4 * self.neuron_count, np.uint32)
        check(lib().oracle_get_rng_state(self._h, _ptr(s, ctypes.c_uint32), self.neuron_count))
        return s.reshape(-1, 4)

    def fixed_weights(self, nsynapses):
        w = np.zeros(nsynapses, np.int32)
        check(lib().oracle_get_fixed_weights(self._h, _ptr(w, ctypes.c_int32), nsynapses))
        return w

    def stdp_accumulators(self, nsynapses):
        a = np.zeros(nsynapses, np.int32)
        check(lib().oracle_get_stdp_accumulators(self._h, _ptr(a, ctypes.c_int32), nsynapses))
        return a

    def delivered_events(self):
        return lib().oracle_delivered_events(self._h)

    def neuron_states(self, var):
        out = np.zeros(self.neuron_count, np.float32)
        check(lib().oracle_get_neuron_states(self._h, var, _ptr(out, ctypes.c_float), out.size))
        return out

    def propagate(self, fired):
        """cpu/Simulation.cpp:115-137 with the fired neurons given by neuron
        index (mapped to local indices here); returns copies of E and I in
        local order."""
        if not hasattr(self, "_g2l"):
            g = self.local_to_global()
            self._g2l = {int(x): i for i, x in enumerate(g)}
        f = np.ascontiguousarray([self._g2l[int(x)] for x in fired], dtype=np.uint32)
        e, i = P(ctypes.c_float)(), P(ctypes.c_float)()
        check(lib().oracle_propagate(self._h, _ptr(f, ctypes.c_uint), f.size, ctypes.byref(e), ctypes.byref(i)))
        n = self.neuron_count
        return (np.ctypeslib.as_array(e, shape=(n,)).copy(), np.ctypeslib.as_array(i, shape=(n,)).copy())


def rng_seeds(first, last):
    out = np.zeros(4 * (last - first + 1), np.uint32)
    lib().oracle_rng_seeds(first, last, _ptr(out, ctypes.c_uint32))
    return out.reshape(-1, 4)


def urand(state):
    s = np.ascontiguousarray(state, dtype=np.uint32)
    v = lib().oracle_urand(_ptr(s, ctypes.c_uint32))
    state[:] = s
    return v


def nrand(state):
    s = np.ascontiguousarray(state, dtype=np.uint32)
    v = lib().oracle_nrand(_ptr(s, ctypes.c_uint32))
    state[:] = s
    return v


def nrand_n(states):
    """nrand of every generator in states (n x 4, advanced in place)."""
    s = np.ascontiguousarray(states, dtype=np.uint32)
    out = np.zeros(s.shape[0], np.float32)
    lib().oracle_nrand_n(_ptr(s, ctypes.c_uint32), s.shape[0], _ptr(out, ctypes.c_float))
    states[...] = s
    return out


def mt19937_nth(seed, n):
    return lib().oracle_mt19937_nth(seed, n)


def fx_tofix(f, fbits=20):
    out = ctypes.c_int32()
    check(lib().oracle_fx_tofix(f, fbits, ctypes.byref(out)))
    return out.value


def wfx_tofloat(v, fbits=20):
    return lib().oracle_wfx_tofloat(v, fbits)
