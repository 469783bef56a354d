"""Loader and ctypes prototypes for libnemo_hip.so (include/nemo_hip.h).

The library is built in-tree by nemo_amd.build; importing a simulation
without it fails loudly -- there is no fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnemo_hip.so")

_lib = None

c_uint_p = ctypes.POINTER(ctypes.c_uint)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_uchar_p = ctypes.POINTER(ctypes.c_ubyte)
c_size_p = ctypes.POINTER(ctypes.c_size_t)
vp = ctypes.c_void_p

# (name, restype, argtypes)
_PROTOTYPES = [
    ("nemo_version", ctypes.c_char_p, []),
    ("nemo_strerror", ctypes.c_char_p, []),
    ("nemo_add_plugin_path", ctypes.c_int, [ctypes.c_char_p]),
    ("nemo_hip_device_count", ctypes.c_int, [c_uint_p]),
    ("nemo_hip_device_description", ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(ctypes.c_char_p)]),
    ("nemo_new_configuration", vp, []),
    ("nemo_delete_configuration", None, [vp]),
    ("nemo_log_stdout", ctypes.c_int, [vp]),
    ("nemo_set_stdp_function", ctypes.c_int, [vp, c_float_p, ctypes.c_size_t, c_float_p, ctypes.c_size_t,
                                              ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("nemo_set_cpu_backend", ctypes.c_int, [vp]),
    ("nemo_set_hip_backend", ctypes.c_int, [vp, ctypes.c_int]),
    ("nemo_hip_device", ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
    ("nemo_backend", ctypes.c_int, [vp, c_uint_p]),
    ("nemo_backend_description", ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_char_p)]),
    ("nemo_set_write_only_synapses", ctypes.c_int, [vp]),
    ("nemo_new_network", vp, []),
    ("nemo_delete_network", None, [vp]),
    ("nemo_add_neuron_type", ctypes.c_int, [vp, ctypes.c_char_p, c_uint_p]),
    ("nemo_add_neuron_iz", ctypes.c_int, [vp, ctypes.c_uint] + [ctypes.c_float] * 7),
    ("nemo_add_neuron", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_add_neurons", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_size_t, c_uint_p, ctypes.c_uint, c_float_p]),
    ("nemo_add_synapse", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_float,
                                        ctypes.c_ubyte, c_u64_p]),
    ("nemo_add_synapses", ctypes.c_int, [vp, ctypes.c_size_t, c_uint_p, c_uint_p, c_uint_p, c_float_p,
                                         c_uchar_p, c_u64_p]),
    ("nemo_neuron_count", ctypes.c_int, [vp, c_uint_p]),
    ("nemo_get_neuron_state_n", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_get_neuron_parameter_n", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_set_neuron_n", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_set_neuron_state_n", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    ("nemo_set_neuron_iz_n", ctypes.c_int, [vp, ctypes.c_uint] + [ctypes.c_float] * 7),
    ("nemo_set_neuron_parameter_n", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    ("nemo_get_synapse_source_n", ctypes.c_int, [vp, ctypes.c_uint64, c_uint_p]),
    ("nemo_get_synapse_target_n", ctypes.c_int, [vp, ctypes.c_uint64, c_uint_p]),
    ("nemo_get_synapse_delay_n", ctypes.c_int, [vp, ctypes.c_uint64, c_uint_p]),
    ("nemo_get_synapse_weight_n", ctypes.c_int, [vp, ctypes.c_uint64, c_float_p]),
    ("nemo_get_synapse_plastic_n", ctypes.c_int, [vp, ctypes.c_uint64, c_uchar_p]),
    ("nemo_get_synapses_from_n", ctypes.c_int, [vp, ctypes.c_uint, ctypes.POINTER(c_u64_p), c_size_p]),
    ("nemo_new_simulation", vp, [vp, vp]),
    ("nemo_delete_simulation", None, [vp]),
    ("nemo_step", ctypes.c_int, [vp, c_uint_p, ctypes.c_size_t, c_uint_p, c_float_p, ctypes.c_size_t,
                                 ctypes.POINTER(c_uint_p), c_size_p]),
    ("nemo_apply_stdp", ctypes.c_int, [vp, ctypes.c_float]),
    ("nemo_get_membrane_potential", ctypes.c_int, [vp, ctypes.c_uint, c_float_p]),
    ("nemo_get_neuron_state_s", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_get_neuron_parameter_s", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_set_neuron_s", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_set_neuron_state_s", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    ("nemo_set_neuron_iz_s", ctypes.c_int, [vp, ctypes.c_uint] + [ctypes.c_float] * 7),
    ("nemo_set_neurons_s", ctypes.c_int, [vp, ctypes.c_size_t, c_uint_p, ctypes.c_uint, c_float_p]),
    ("nemo_get_neuron_states_s", ctypes.c_int, [vp, ctypes.c_size_t, c_uint_p, ctypes.c_uint, c_float_p]),
    ("nemo_get_neuron_parameters_s", ctypes.c_int, [vp, ctypes.c_size_t, c_uint_p, ctypes.c_uint, c_float_p]),
    ("nemo_set_neuron_states_s", ctypes.c_int, [vp, ctypes.c_size_t, c_uint_p, ctypes.c_uint, c_float_p]),
    ("nemo_set_neuron_parameters_s", ctypes.c_int, [vp, ctypes.c_size_t, c_uint_p, ctypes.c_uint, c_float_p]),
    ("nemo_set_neuron_parameter_s", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    ("nemo_get_synapse_source_s", ctypes.c_int, [vp, ctypes.c_uint64, c_uint_p]),
    ("nemo_get_synapse_target_s", ctypes.c_int, [vp, ctypes.c_uint64, c_uint_p]),
    ("nemo_get_synapse_delay_s", ctypes.c_int, [vp, ctypes.c_uint64, c_uint_p]),
    ("nemo_get_synapse_weight_s", ctypes.c_int, [vp, ctypes.c_uint64, c_float_p]),
    ("nemo_get_synapse_plastic_s", ctypes.c_int, [vp, ctypes.c_uint64, c_uchar_p]),
    ("nemo_get_synapses_from_s", ctypes.c_int, [vp, ctypes.c_uint, ctypes.POINTER(c_u64_p), c_size_p]),
    ("nemo_elapsed_wallclock", ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_ulong)]),
    ("nemo_elapsed_simulation", ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_ulong)]),
    ("nemo_reset_timer", ctypes.c_int, [vp]),
    ("nemo_hip_run", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float, c_uint_p,
                                    ctypes.POINTER(c_uint_p), c_size_p, c_u64_p]),
    ("nemo_hip_run_into", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_float, c_uint_p, c_uint_p,
                                         ctypes.c_size_t, c_size_p, ctypes.POINTER(c_uint_p), c_size_p, c_u64_p]),
    ("nemo_hip_new_batch", ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_size_t, vp, ctypes.POINTER(vp)]),
    ("nemo_hip_batch_size", ctypes.c_int, [vp, c_uint_p]),
    ("nemo_hip_batch_get_neuron_state", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_hip_batch_get_neuron_parameter", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                                           c_float_p]),
    ("nemo_hip_batch_get_membrane_potential", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    ("nemo_hip_prepare", ctypes.c_int, [vp, ctypes.c_uint, ctypes.c_uint]),
    ("nemo_hip_get_launch_shape", ctypes.c_int, [vp, c_uint_p]),
    ("nemo_hip_set_launch_shape", ctypes.c_int, [vp, c_uint_p]),
    ("nemo_hip_autotune", ctypes.c_int, [vp, ctypes.c_uint, c_uint_p]),
    ("nemo_hip_new_float_network", ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp)]),
    ("nemo_hip_delete_float_network", None, [vp]),
    ("nemo_hip_float_network_sizes", ctypes.c_int, [vp, c_size_p, c_size_p, c_uint_p]),
    ("nemo_hip_float_forward", ctypes.c_int, [vp, vp, ctypes.c_uint, ctypes.c_uint] + [vp] * 11),
    ("nemo_hip_float_forward_static", ctypes.c_int, [vp, vp, ctypes.c_uint, ctypes.c_uint] + [vp] * 11),
    ("nemo_hip_float_backward", ctypes.c_int, [vp, vp, ctypes.c_uint, ctypes.c_uint] + [vp] * 17),
    ("nemo_hip_set_debug", ctypes.c_int, [vp, ctypes.c_int]),
    ("nemo_hip_get_currents", ctypes.c_int, [vp, c_float_p, c_float_p, ctypes.c_size_t]),
    ("nemo_hip_local_to_global", ctypes.c_int, [vp, c_uint_p, ctypes.c_size_t]),
    ("nemo_hip_get_recent_firing", ctypes.c_int, [vp, c_u64_p, ctypes.c_size_t]),
    ("nemo_hip_get_rng_state", ctypes.c_int, [vp, c_u32_p, ctypes.c_size_t]),
    ("nemo_hip_get_neuron_states", ctypes.c_int, [vp, ctypes.c_uint, c_float_p, ctypes.c_size_t]),
    ("nemo_hip_get_fixed_weights", ctypes.c_int, [vp, c_i32_p, ctypes.c_size_t]),
    ("nemo_hip_get_stdp_accumulators", ctypes.c_int, [vp, c_i32_p, ctypes.c_size_t]),
    ("nemo_hip_counts", ctypes.c_int, [vp, c_size_p, c_size_p, c_size_p]),
    ("nemo_hip_profile", ctypes.c_int, [vp, ctypes.c_uint, ctypes.POINTER(ctypes.c_double), c_u64_p]),
    ("nemo_hip_eval_math", ctypes.c_int, [ctypes.c_int, c_float_p, c_float_p, ctypes.c_size_t]),
    ("nemo_hip_propagate", ctypes.c_int, [vp, c_uint_p, ctypes.c_size_t, c_float_p, c_float_p, ctypes.c_size_t]),
    ("nemo_hip_propagate_device", ctypes.c_int, [vp, vp, ctypes.POINTER(c_float_p), ctypes.POINTER(c_float_p)]),
    ("nemo_hip_test_device", ctypes.c_int, [ctypes.c_uint]),
]


class NemoError(RuntimeError):
    """An error reported by the library, with its NEMO_* status code."""

    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def lib():
    """Return the loaded library, loading it on first use."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libnemo_hip.so is not built (%s); run `python -m nemo_amd.build`" % LIB_PATH)
        # One HIP runtime per process: PyTorch-ROCm ships its own
        # libamdhip64.so.7, and loading it first makes this library bind to
        # that copy (same soname), so that torch tensors and this library's
        # kernels share one runtime.  Without torch, the system runtime is used.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, res, args in _PROTOTYPES:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(status):
    if status != 0:
        raise NemoError(status, lib().nemo_strerror().decode())
