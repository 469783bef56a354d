"""The allocation-free float forward (nemo_hip_float_forward_static) against
nemo_hip_float_forward, through the C ABI: two calls in a row, the second
with fewer steps, each bit-identical in spikes, membrane potentials and
checkpoints; a call whose workspace exceeds the static array is refused."""
import ctypes

import numpy as np
import pytest

import nemo_amd
from nemo_amd import generators as gen
from nemo_amd._lib import check, lib

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu(require_gpu):
    pass


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(entry, h, spec, T, K, seed):
    import torch
    dev = torch.device("cuda", 0)
    n = spec.neuron_count
    a, b, c, d, sg, u0, v0 = [torch.tensor(np.asarray(x), device=dev) for x in spec.blocks[0].args]
    w = torch.tensor(spec.weight, device=dev)
    g = torch.Generator(device=dev).manual_seed(seed)
    x = sg * torch.randn(T, n, device=dev, generator=g)
    S = torch.full((T, n), -1.0, dtype=torch.float32, device=dev)
    V = torch.full((T, n), -1.0, dtype=torch.float32, device=dev)
    ck = torch.full(((T + K - 1) // K, 2, n), -1.0, dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    check(entry(h, stream, T, K, _ptr(w), _ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(v0), _ptr(u0), _ptr(x),
                _ptr(S), _ptr(V), _ptr(ck)))
    torch.cuda.synchronize()
    return S.cpu().numpy(), V.cpu().numpy(), ck.cpu().numpy()


def _float_network(spec):
    net = nemo_amd.Network()
    spec.load(net)
    h = ctypes.c_void_p()
    check(lib().nemo_hip_new_float_network(net._h, 0, ctypes.byref(h)))
    return net, h


def test_static_forward_matches_forward_twice():
    spec = gen.random_izhikevich(3000, 100, dmax=10, seed=71)
    net, h = _float_network(spec)
    try:
        for T, K, seed in ((300, 50, 1), (120, 25, 2)):
            want = _run(lib().nemo_hip_float_forward, h, spec, T, K, seed)
            got = _run(lib().nemo_hip_float_forward_static, h, spec, T, K, seed)
            assert want[0].sum() > 0
            for x, y in zip(want, got):
                assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (T, K)
    finally:
        lib().nemo_hip_delete_float_network(h)


def test_static_forward_refuses_oversized_workspace():
    """10k neurons x 200 synapses: 8 MB of weights and state plus 1.3 KB of
    bitmaps per step, so 60000 steps exceed the 64 MiB array."""
    spec = gen.random_izhikevich(10000, 200, dmax=10, seed=72)
    net, h = _float_network(spec)
    try:
        import torch
        dev = torch.device("cuda", 0)
        one = torch.zeros(1, device=dev)
        args = [h, None, 60000, 50] + [_ptr(one)] * 11
        assert lib().nemo_hip_float_forward_static(*args) == 6  # NEMO_INVALID_INPUT
        assert b"static workspace" in lib().nemo_strerror()
    finally:
        lib().nemo_hip_delete_float_network(h)
