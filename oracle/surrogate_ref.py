"""FP64 PyTorch reference of the float-weight Izhikevich model and its
surrogate gradient -- TEST INFRASTRUCTURE ONLY (imported by tests and by
nothing that ships).

The float-weight mode has no counterpart in NeMo; its neuron step is NeMo's
Izhikevich step (src/nemo/cpu/plugins/Izhikevich.cpp:14-98: four 0.25 ms
substeps, each skipped once v has reached 30, then reset) with FP32 weights
instead of fixed point.  Gradients come from torch autograd in float64 over
the same expression:
  * the substep masks ("not yet fired") are constants of the forward pass
    (torch.where with a boolean condition);
  * the reset v <- v + S (c - v), u <- u + S d uses the spike as a constant
    (detached), so the threshold is differentiated only where the spike
    feeds synapses and the loss;
  * there the spike S = H(v - 30) has the fast-sigmoid surrogate derivative
    1 / (1 + 10 |v - 30|)^2.
Synaptic input: I[t] = x[t] + index_add over synapses of w * S[t - delay][src].

Pinned by: central finite differences in a spike-free regime (where the
surrogate never enters the loss), by the surrogate's closed form, and by the
forward pass reducing to the plain model's neuron step.
"""
import torch


class Heaviside(torch.autograd.Function):
    """H(z) = [z >= 0] with dH/dz := 1 / (1 + 10 |z|)^2."""

    @staticmethod
    def forward(ctx, z):
        ctx.save_for_backward(z)
        return (z >= 0).to(z.dtype)

    @staticmethod
    def backward(ctx, g):
        (z,) = ctx.saved_tensors
        return g / (1.0 + 10.0 * z.abs()) ** 2


def simulate(src, tgt, delay, w, a, b, c, d, x, v0, u0):
    """src, tgt, delay: int64 [S]; w: [S]; a, b, c, d, v0, u0: [N]; x: [T][N]
    (all float64, CPU).  Returns S, V ([T][N]) as differentiable tensors."""
    T, N = x.shape
    h = 0.25
    v, u = v0, u0
    spikes, volts = [], []
    for t in range(T):
        valid = delay <= t
        I = x[t]
        if bool(valid.any()):
            hist = torch.stack(spikes) if spikes else None
            sv, tv, dv, wv = src[valid], tgt[valid], delay[valid], w[valid]
            arriving = hist[t - dv, sv]
            I = I + torch.zeros(N, dtype=x.dtype).index_add(0, tv, wv * arriving)
        m = torch.ones(N, dtype=torch.bool)
        for _ in range(4):
            vn = v + h * (0.04 * v * v + 5.0 * v + 140.0 - u + I)
            v = torch.where(m, vn, v)
            un = u + h * a * (b * v - u)
            u = torch.where(m, un, u)
            m = m & (v < 30.0).detach()
        s = Heaviside.apply(v - 30.0)
        r = s.detach()
        v = v + r * (c - v)
        u = u + r * d
        spikes.append(s)
        volts.append(v)
    return torch.stack(spikes), torch.stack(volts)
