"""Float-weight mode and its adjoint against the FP64 PyTorch autograd of the
same surrogate model (oracle/surrogate_ref.py), and the reference itself
against finite differences.

Bars: spike rasters identical to the FP64 forward; gradients within 1e-4
relative (2-norm over each parameter group) of FP64 autograd; in a spike-free
regime (weights x0.01, input sigma 1) the GPU adjoint within 1e-4 of central
finite differences of the FP64 forward, which is where the surrogate cannot
matter.
"""
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import surrogate_ref as ref  # noqa: E402

from nemo_amd import generators as gen  # noqa: E402


def instance(n, m, dmax, seed, sigma, wscale=1.0, steps=100):
    """Izhikevich 80/20 network (random_izhikevich) as float64 CPU tensors."""
    spec = gen.random_izhikevich(n, m, dmax=dmax, seed=seed)
    a, b, c, d, sg, u, v = [torch.tensor(np.asarray(x), dtype=torch.float64) for x in spec.blocks[0].args]
    g = torch.Generator().manual_seed(seed)
    x = (sg * torch.randn(steps, n, generator=g, dtype=torch.float64)) * (sigma / 5.0)
    x = x.float().double()  # the GPU takes float inputs: give both sides the same values
    # network order = source ascending, insertion order: the generator's order
    src = torch.tensor(spec.source.astype(np.int64))
    tgt = torch.tensor(spec.target.astype(np.int64))
    dly = torch.tensor(spec.delay.astype(np.int64))
    w = torch.tensor(spec.weight.astype(np.float64)) * wscale
    return spec, dict(src=src, tgt=tgt, delay=dly, w=w, a=a, b=b, c=c, d=d, x=x, v0=v, u0=u)


# ------------------------------------------------------ reference (CPU)
def test_surrogate_closed_form():
    z = torch.tensor([-3.0, -0.1, 0.0, 0.05, 2.0], dtype=torch.float64, requires_grad=True)
    s = ref.Heaviside.apply(z)
    assert s.tolist() == [0.0, 0.0, 1.0, 1.0, 1.0]
    s.sum().backward()
    assert torch.allclose(z.grad, 1.0 / (1.0 + 10.0 * z.detach().abs()) ** 2, rtol=0, atol=0)


def test_reference_delay_indexing():
    """A neuron driven over threshold at step 0 reaches its target through a
    3 ms synapse of weight 1000: the target fires at step 3, not before."""
    N, T = 2, 6
    x = torch.zeros(T, N, dtype=torch.float64)
    x[0, 0] = 1000.0
    one = torch.ones(N, dtype=torch.float64)
    S, V = ref.simulate(torch.tensor([0]), torch.tensor([1]), torch.tensor([3]),
                        torch.tensor([1000.0], dtype=torch.float64), 0.02 * one, 0.2 * one, -65.0 * one,
                        8.0 * one, x, -65.0 * one, -13.0 * one)
    assert S[:, 0].tolist() == [1, 0, 0, 0, 0, 0]
    assert S[:, 1].tolist() == [0, 0, 0, 1, 0, 0]


def _fd_setup(n=60, m=12, steps=40, seed=3):
    spec, p = instance(n, m, 5, seed, sigma=1.0, wscale=0.01, steps=steps)
    r = torch.randn(steps, n, generator=torch.Generator().manual_seed(seed + 1), dtype=torch.float64)
    return p, r


def _loss_v(p, r, **over):
    q = dict(p)
    q.update(over)
    S, V = ref.simulate(q["src"], q["tgt"], q["delay"], q["w"], q["a"], q["b"], q["c"], q["d"], q["x"],
                        q["v0"], q["u0"])
    assert float(S.sum()) == 0.0
    return (r * V).sum()


@pytest.mark.parametrize("name", ["w", "a", "b", "d", "x"])
def test_reference_gradient_matches_finite_differences(name):
    """Spike-free regime: FP64 autograd of the reference against central
    differences along a random direction (pins the masked substeps, the
    reset algebra and the delayed delivery of the reference itself)."""
    p, r = _fd_setup()
    leaf = p[name].clone().requires_grad_(True)
    _loss_v(p, r, **{name: leaf}).backward()
    g = torch.Generator().manual_seed(11)
    dirn = torch.randn(leaf.shape, generator=g, dtype=torch.float64)
    h = 1e-6 * max(1.0, float(p[name].abs().max()))
    with torch.no_grad():
        lp = _loss_v(p, r, **{name: p[name] + h * dirn})
        lm = _loss_v(p, r, **{name: p[name] - h * dirn})
    fd = float((lp - lm) / (2 * h))
    an = float((leaf.grad * dirn).sum())
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (name, fd, an)


# -------------------------------------------------------------- GPU
def _gpu_or_fail():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")


def _float_network(spec):
    import nemo_amd
    from nemo_amd.differentiable import FloatNetwork
    net = nemo_amd.Network()
    spec.load(net)
    return FloatNetwork(net, 0)


def _to_dev(p, names, grad=()):
    out = {}
    for k in names:
        t = p[k].to(torch.float32).to("cuda")
        out[k] = t.requires_grad_(True) if k in grad else t
    return out


PARAMS = ("w", "a", "b", "d", "x")


def _gpu_run(fnet, q, K):
    return fnet(q["w"], q["a"], q["b"], q["d"], q["x"], c=q["c"], v0=q["v0"], u0=q["u0"], checkpoint_every=K)


def _rate_loss(S, target, last):
    rates = S[-last:].sum(0) / (last * 1e-3)
    return ((rates - target) ** 2).mean()


def _rel(a, b):
    return float((a - b).norm() / max(float(b.norm()), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("stage", ["auto", "0"])
def test_forward_matches_fp64_reference(stage, monkeypatch):
    """Spike bitmaps staged in LDS (auto for this size) and read from global
    memory (NEMO_HIP_FLOAT_STAGE=0)."""
    _gpu_or_fail()
    if stage == "0":
        monkeypatch.setenv("NEMO_HIP_FLOAT_STAGE", "0")
    spec, p = instance(1000, 100, 10, seed=5, sigma=5.0, steps=100)
    fnet = _float_network(spec)
    q = _to_dev(p, ("w", "a", "b", "c", "d", "x", "v0", "u0"))
    S, V = _gpu_run(fnet, q, 50)
    Sr, Vr = ref.simulate(p["src"], p["tgt"], p["delay"], p["w"], p["a"], p["b"], p["c"], p["d"], p["x"],
                          p["v0"], p["u0"])
    assert torch.equal(S.cpu().double(), Sr), "raster differs from the FP64 forward"
    assert float(Sr.sum()) > 100
    # state is FP32 between steps: per-step rounding (~4e-6 relative) grows
    # near threshold, so bound the typical and the worst deviation
    dv = (V.cpu().double() - Vr).abs()
    assert float(dv.quantile(0.999)) < 1e-3
    assert float(dv.max()) < 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("stage", ["auto", "0"])
def test_gradients_match_fp64_autograd(stage, monkeypatch):
    """1k neurons, 100 synapses each, 100 steps, spike-count loss over the
    last 50 steps against random target rates; spike bitmaps staged and not."""
    _gpu_or_fail()
    if stage == "0":
        monkeypatch.setenv("NEMO_HIP_FLOAT_STAGE", "0")
    T, last = 100, 50
    spec, p = instance(1000, 100, 10, seed=5, sigma=5.0, steps=T)
    target = torch.rand(1000, generator=torch.Generator().manual_seed(2), dtype=torch.float64) * 18 + 2
    fnet = _float_network(spec)
    q = _to_dev(p, ("w", "a", "b", "c", "d", "x", "v0", "u0"), grad=PARAMS)
    S, V = _gpu_run(fnet, q, 50)
    _rate_loss(S, target.float().cuda(), last).backward()

    leaves = {k: p[k].clone().requires_grad_(True) for k in PARAMS}
    pr = dict(p)
    pr.update(leaves)
    Sr, Vr = ref.simulate(pr["src"], pr["tgt"], pr["delay"], pr["w"], pr["a"], pr["b"], pr["c"], pr["d"], pr["x"],
                          pr["v0"], pr["u0"])
    assert torch.equal(S.detach().cpu().double(), Sr.detach())
    _rate_loss(Sr, target, last).backward()
    for k in PARAMS:
        err = _rel(q[k].grad.cpu().double(), leaves[k].grad)
        assert err < 1e-4, (k, err)


@pytest.mark.gpu
def test_gradients_match_finite_differences_spike_free():
    """Weights x0.01, input sigma 1: no neuron fires, the loss is a random
    projection of the membrane trace; the GPU adjoint against central
    differences of the FP64 forward along random directions."""
    _gpu_or_fail()
    T = 100
    spec, p = instance(1000, 100, 10, seed=6, sigma=1.0, wscale=0.01, steps=T)
    r = torch.randn(T, 1000, generator=torch.Generator().manual_seed(9), dtype=torch.float64)
    spec.weight = (spec.weight * np.float32(0.01)).astype(np.float32)
    fnet = _float_network(spec)
    q = _to_dev(p, ("w", "a", "b", "c", "d", "x", "v0", "u0"), grad=PARAMS)
    S, V = _gpu_run(fnet, q, 50)
    assert float(S.sum()) == 0.0
    (r.float().cuda() * V).sum().backward()
    for k in PARAMS:
        dirn = torch.randn(p[k].shape, generator=torch.Generator().manual_seed(13), dtype=torch.float64)
        h = 1e-6 * max(1.0, float(p[k].abs().max()))
        with torch.no_grad():
            lp = _loss_v(p, r, **{k: p[k] + h * dirn})
            lm = _loss_v(p, r, **{k: p[k] - h * dirn})
        fd = float((lp - lm) / (2 * h))
        an = float((q[k].grad.cpu().double() * dirn).sum())
        assert abs(fd - an) <= 1e-4 * abs(fd), (k, fd, an)


@pytest.mark.gpu
@pytest.mark.parametrize("n,m,dmax,T", [(800, 60, 10, 60), (300, 40, 64, 150)])
def test_checkpoint_interval_does_not_change_gradients(n, m, dmax, T):
    """The adjoint replays each checkpoint segment in one launch from the
    recorded spikes (16 steps at a time, each synapse's spikes from 64-step
    history words); gradients must not depend on the segment length, so the
    replayed state equals the forward's bit for bit.  The second case has
    delays up to 64 and segments that cross history words."""
    _gpu_or_fail()
    spec, p = instance(n, m, dmax, seed=7, sigma=5.0, steps=T)
    fnet = _float_network(spec)
    target = torch.full((n,), 10.0, device="cuda")
    grads = []
    for K in (1, 7, 50, T):
        q = _to_dev(p, ("w", "a", "b", "c", "d", "x", "v0", "u0"), grad=PARAMS)
        S, V = _gpu_run(fnet, q, K)
        (_rate_loss(S, target, 30) + 1e-3 * V.mean()).backward()
        grads.append([q[k].grad.cpu() for k in PARAMS])
    for g in grads[1:]:
        for x, y in zip(grads[0], g):
            assert torch.equal(x, y)
