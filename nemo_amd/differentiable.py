"""Float-weight mode of an Izhikevich network as a torch.autograd.Function
(PyTorch-ROCm), with the hand-written HIP adjoint of libnemo_hip.so.

    fnet = FloatNetwork(network)            # one Izhikevich type, indices 0..N-1
    S, V = fnet(w, a, b, d, x, c=c, v0=v0, u0=u0, checkpoint_every=50)

w: FP32 weights in network order (source ascending, then insertion order);
a, b, c, d, v0, u0: per neuron; x: [T][N] external current (e.g. Gaussian
input); S: [T][N] spikes (0/1), V: [T][N] membrane potential after each step.
Gradients flow to w, a, b, d, x, v0 and u0; where spikes feed synapses and
the loss, the threshold is differentiated through the surrogate
1 / (1 + 10 |v - 30|)^2, while the reset and the substep stops are held
constant (model: include/nemo_hip.h, nemo_hip_float_forward).  All tensors live on the HIP
device; there is no CPU path.
"""
import ctypes

import torch

from ._lib import NemoError, check, lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dev_f32(t, name, shape, device):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a tensor" % name)
    if t.dtype != torch.float32 or t.device != device or tuple(t.shape) != tuple(shape):
        raise ValueError("%s must be float32 on %s with shape %s (got %s %s %s)"
                         % (name, device, tuple(shape), t.dtype, t.device, tuple(t.shape)))
    return t.contiguous()


class _Spiking(torch.autograd.Function):

    @staticmethod
    def forward(ctx, w, a, b, d, x, v0, u0, c, fnet, K):
        T, N = x.shape
        dev = x.device
        S = torch.empty((T, N), dtype=torch.float32, device=dev)
        V = torch.empty((T, N), dtype=torch.float32, device=dev)
        ck = torch.empty(((T + K - 1) // K, 2, N), dtype=torch.float64, device=dev)
        check(lib().nemo_hip_float_forward(fnet._h, _stream(dev), T, K, _ptr(w), _ptr(a), _ptr(b), _ptr(c),
                                           _ptr(d), _ptr(v0), _ptr(u0), _ptr(x), _ptr(S), _ptr(V), _ptr(ck)))
        ctx.save_for_backward(w, a, b, d, x, c, S, ck)
        ctx.fnet = fnet
        ctx.K = K
        return S, V

    @staticmethod
    def backward(ctx, gS, gV):
        w, a, b, d, x, c, S, ck = ctx.saved_tensors
        T, N = x.shape
        dev = x.device
        gS = gS.contiguous() if gS is not None else None
        gV = gV.contiguous() if gV is not None else None
        gw = torch.empty_like(w)
        ga = torch.empty_like(a)
        gb = torch.empty_like(b)
        gd = torch.empty_like(d)
        gx = torch.empty_like(x) if ctx.needs_input_grad[4] else None
        gv0 = torch.empty_like(a) if ctx.needs_input_grad[5] else None
        gu0 = torch.empty_like(a) if ctx.needs_input_grad[6] else None
        check(lib().nemo_hip_float_backward(ctx.fnet._h, _stream(dev), T, ctx.K, _ptr(w), _ptr(a), _ptr(b),
                                            _ptr(c), _ptr(d), _ptr(x), _ptr(S), _ptr(ck), _ptr(gS), _ptr(gV),
                                            _ptr(gw), _ptr(ga), _ptr(gb), _ptr(gd), _ptr(gx), _ptr(gv0),
                                            _ptr(gu0)))
        return gw, ga, gb, gd, gx, gv0, gu0, None, None, None


class FloatNetwork:
    """A network's connectivity on the device, for float-weight simulation
    and its adjoint."""

    def __init__(self, net, device=0):
        h = ctypes.c_void_p()
        check(lib().nemo_hip_new_float_network(net._h, int(device), ctypes.byref(h)))
        self._h = h.value
        n = ctypes.c_size_t()
        s = ctypes.c_size_t()
        dmax = ctypes.c_uint()
        check(lib().nemo_hip_float_network_sizes(self._h, ctypes.byref(n), ctypes.byref(s), ctypes.byref(dmax)))
        self.neurons, self.synapses, self.max_delay = n.value, s.value, dmax.value
        self.device = torch.device("cuda", int(device))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().nemo_hip_delete_float_network(self._h)
            self._h = None

    def __call__(self, w, a, b, d, x, c, v0, u0, checkpoint_every=50):
        """Simulate x.shape[0] steps; returns (S, V), both [T][N]."""
        if not torch.cuda.is_available():
            raise NemoError(-1, "the float-weight mode needs a HIP device")
        N, dev = self.neurons, self.device
        T = x.shape[0]
        if T == 0 or checkpoint_every <= 0:
            raise ValueError("need at least one step and a positive checkpoint interval")
        w = _dev_f32(w, "w", (self.synapses,), dev)
        a = _dev_f32(a, "a", (N,), dev)
        b = _dev_f32(b, "b", (N,), dev)
        d = _dev_f32(d, "d", (N,), dev)
        c = _dev_f32(c, "c", (N,), dev)
        v0 = _dev_f32(v0, "v0", (N,), dev)
        u0 = _dev_f32(u0, "u0", (N,), dev)
        x = _dev_f32(x, "x", (T, N), dev)
        return _Spiking.apply(w, a, b, d, x, v0, u0, c, self, int(checkpoint_every))
