#!/usr/bin/env python3
"""Benchmark of the NeMo HIP backend on one MI355X.

Workload (the single-GPU benchmark network): 100,000 Izhikevich neurons, 80/20
excitatory/inhibitory, 1,000 synapses per neuron (1e8 synapses) to uniformly
random targets with delays U{1..20} ms, Gaussian input sigma 5/2, STDP on the
excitatory synapses (A+ = 0.1 exp(-dt/20), A- = -0.12 exp(-dt/20), 20 ms
windows, weights in [0, 10]) with applyStdp(1.0) every 1000 steps.  Fixed-point
Q11.20 currents, FP32 state (FP64 v update), 1 ms steps.

A "step" is one simulated millisecond: spike delivery, neuron update, STDP
accumulation, with the step's firing read back to the host (every step's fired
list is returned, in graphs of up to 64 steps).  The timed region starts with
the network resident in HBM and ends after the last readout.

Prints one JSON line.  --gpus N with N > 1 exits: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "synaptic events/s: plain vs batched-64 nets; fwd+adjoint ms/step, 1 MI355X"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes (lower bounds; intermediates excluded)
BYTES_PER_EVENT = 8            # 4 B target + 4 B Q11.20 weight
BYTES_PER_IZH_NEURON = 84      # 20 B params + 16 B u,v r/w + 32 B RNG r/w + 16 B history r/w
BYTES_PER_STDP_SYNAPSE = 20    # 4 B packed source/delay + 8 B presynaptic history + 8 B accumulator r/w


def traffic_key(shape):
    """The launch-shape fields the launches' traffic depends on: delivery
    blocks of both sets, bin capacity and STDP mode, the STDP words when STDP
    is fused into a launch (mode 2), and the overlap when the sets are split
    (it moves delivery into the update launch)."""
    return list(shape[:4]) + ([shape[4]] if shape[3] == 2 else []) + [shape[5] if shape[1] else 0]


def launch_bytes(shape, kern, work, n, spec):
    """Algorithmic bytes per launch of the propagate and update launches of
    the plain workload under a launch shape, and each launch's time per step
    (us).  Overlap 3 (shape[5], split sets): the update launch of step t also
    delivers all events of step t+1 (delays 2..D as carried blocks, delay 1
    pushed by its tiles), and only the first step of each run of 64 has a
    propagate launch.  Overlap 2: the update launch delivers delays 2..D,
    their share of events taken as their share of the synapses (uniform
    firing).  The STDP accumulation counts towards a launch only in STDP
    mode 2, where it shares that launch; otherwise it runs in launches of its
    own (mode 4: one batch per run of up to 64 steps)."""
    steps = work["steps"]
    ev = BYTES_PER_EVENT * work["events"] / steps
    stdp = BYTES_PER_STDP_SYNAPSE * work["stdp_synapses"] / steps if shape[3] == 2 else 0.0
    neurons = BYTES_PER_IZH_NEURON * n
    split = shape[1] > 0
    overlap = shape[5] if split else 0
    if overlap == 3:
        per_launch = {"propagate": ev, "update": neurons + ev + stdp}
    elif overlap == 2:
        late = float(np.count_nonzero(spec.delay == 1)) / max(1, spec.delay.size)
        per_launch = {"propagate": ev * late, "update": neurons + ev * (1.0 - late) + stdp}
    else:
        per_launch = {"propagate": ev + stdp, "update": neurons}
    per_step_us = {"propagate": kern["propagate"] * kern.get("propagate_launches", 1.0), "update": kern["update"]}
    return per_launch, per_step_us


def measured_traffic(kernel, shape):
    """HBM bytes per launch of `kernel` from the committed rocprof counter
    summaries of this workload (FETCH_SIZE and WRITE_SIZE passes scaled by the
    copy-kernel calibration), from the one measured at a launch shape that
    agrees with `shape` in the fields that matter (traffic_key); None if
    there is none (the traffic of a launch depends on its shape)."""
    split = shape[1] > 0
    names = {"propagate": "propagate_sets_kernel" if split else "propagate_kernel", "update": "izhikevich_update"}
    for name in ("update_delivery_izh100k_stdp.md", "propagate_update_izh100k_stdp.md"):
        try:
            measured_shape = None
            fields = {}
            for line in open(os.path.join(ROOT, "profiles", name)):
                if line.startswith("traffic_launch_shape:"):
                    measured_shape = [int(x) for x in line.split(":", 1)[1].split()]
                if line.startswith("traffic_bytes_per_launch:"):
                    fields = dict(f.split("=") for f in line.split(":", 1)[1].split())
            if measured_shape is not None and traffic_key(measured_shape) == traffic_key(shape):
                return float(fields[names[kernel]])
        except (OSError, KeyError, ValueError):
            pass
    return None


def host_description():
    """CPU model and C library of this host (the CPU baseline's libm is part
    of the numerics)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        libc = os.confstr("CS_GNU_LIBC_VERSION")
    except (ValueError, OSError):
        libc = "unknown"
    return model, libc


def stdp_function():
    from nemo_amd import generators as gen
    pre, post = gen.stdp_exp_function(0.1, -0.12, 20.0, 20)
    return pre, post, 0.0, 10.0, 0.0, -10.0


def build_network(api, spec, stdp):
    net = api.Network()
    spec.load(net)
    conf = api.Configuration()
    if stdp:
        conf.set_stdp_function(*stdp_function())
    return api.Simulation(net, conf)


def cpu_baseline(spec, warmup, steps, seconds, period=1000):
    """The CPU oracle (restated reference CPU backend) on the same network and
    the same step window as the GPU's timed run: `warmup` untimed steps from
    construction, then up to `steps` timed steps (fewer if they would take
    more than `seconds`), applyStdp(1.0) every `period` steps of each of the
    two runs as in the GPU's, on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa
    sim = build_network(oa, spec, True)
    t_w = time.perf_counter()
    for s in range(warmup):
        sim.step()
        if (s + 1) % period == 0:
            sim.apply_stdp(1.0)
    t_w = time.perf_counter() - t_w
    ev0 = sim.delivered_events()
    t0 = time.perf_counter()
    done = 0
    while done < steps and time.perf_counter() - t0 < seconds:
        sim.step()
        done += 1
        if done % period == 0:
            sim.apply_stdp(1.0)
    wall = time.perf_counter() - t0
    events = sim.delivered_events() - ev0
    model, libc = host_description()
    return {
        "value": events / wall,
        "unit": "events/s",
        "cores": int(oa.lib().oracle_num_threads()),
        "kind": "port",
        "sample": "steps %d..%d of the same network (the GPU's timed window is steps %d..%d), %d untimed "
                  "warm-up steps first (%.1f s); %.1f s wall, %.2f ms/step"
                  % (warmup, warmup + done, warmup, warmup + steps, warmup, t_w, wall, 1000.0 * wall / max(1, done)),
        "ms_per_step": 1000.0 * wall / max(1, done),
        "cpu_model": model,
        "libc": libc,
    }


def batched(api, gen, count, steps, autotune_steps, warmup=200):
    """The batched workload: `count` independent networks (log-uniform sizes
    1k..50k, U{100..1000} synapses per neuron, STDP on every other one) in one
    simulation, against the same networks run one after another as plain
    simulations.  Both run `warmup` untimed steps after construction (the
    first steps of a new simulation run about 10% slower; the warm-up reads
    the firing back too, which sizes the fired-list buffer of the timed run),
    then `steps` timed steps, firing read back every step, graphs captured before the timed
    region; each network's per-step firing counts are compared with its solo
    run."""
    t = time.perf_counter()
    specs = gen.batch_izhikevich(count=count)
    nets = []
    for spec, _ in specs:
        net = api.Network()
        spec.load(net)
        nets.append(net)
    nsyn = sum(spec.synapse_count for spec, _ in specs)
    nneu = sum(spec.neuron_count for spec, _ in specs)
    del specs[:]
    conf = api.Configuration()
    conf.set_stdp_function(*stdp_function())
    batch = api.Batch(nets, conf)
    print("batch of %d networks: %d neurons, %d synapses, built in %.1f s"
          % (count, nneu, nsyn, time.perf_counter() - t), file=sys.stderr, flush=True)
    default_shape = batch.launch_shape
    shape = batch.autotune(autotune_steps) if autotune_steps else default_shape
    if warmup:
        batch.run(warmup, stdp_period=1000, reward=1.0)
    batch.prepare(steps, stdp_period=1000)
    t0 = time.perf_counter()
    bcounts, _, bevents = batch.run(steps, stdp_period=1000, reward=1.0)
    bwall = time.perf_counter() - t0
    del batch
    print("batched run %.3f s" % bwall, file=sys.stderr, flush=True)

    swall = 0.0
    sevents = 0
    exact = True
    stdp_conf = api.Configuration()
    stdp_conf.set_stdp_function(*stdp_function())
    plain_conf = api.Configuration()
    for k, net in enumerate(nets):
        stdp = k % 2 == 0
        sim = api.Simulation(net, stdp_conf if stdp else plain_conf)
        if warmup:
            sim.run(warmup, stdp_period=1000 if stdp else 0, reward=1.0)
        sim.prepare(steps, stdp_period=1000 if stdp else 0)
        t0 = time.perf_counter()
        c, _, e = sim.run(steps, stdp_period=1000 if stdp else 0, reward=1.0)
        swall += time.perf_counter() - t0
        sevents += e
        exact = exact and np.array_equal(c, bcounts[:, k])
        del sim
    print("sequential runs %.3f s" % swall, file=sys.stderr, flush=True)
    return {
        "workload": "izhikevich_80_20_batch_%d_nets" % count,
        "networks": count, "neurons": nneu, "synapses": nsyn, "steps": steps, "warmup": warmup,
        "events_per_s": bevents / bwall,
        "ms_per_step": 1000.0 * bwall / steps,
        "sequential_events_per_s": sevents / swall,
        "sequential_ms_per_step": 1000.0 * swall / steps,
        "speedup_vs_sequential": (bevents / bwall) / (sevents / swall),
        "events_equal": bevents == sevents,
        "firing_counts_equal_solo": bool(exact),
        "launch_shape": list(shape),
        "default_launch_shape": list(default_shape),
        "note": "sequential = the same networks as separate plain simulations (default launch shapes), "
                "one after another; construction, warm-up steps and graph capture excluded from both",
    }


def mixed(api, gen, steps=2000, warmup=200, autotune_steps=128, profile_steps=300):
    """The mixed-model workload: 50k neurons on the unit torus (40k
    IF_curr_exp, 8k QIF, 2k HH) plus 1k Poisson sources at 10 Hz, 500
    outgoing synapses each with probability falling with distance (sigma
    0.1) and delay ceil(20 d) + 1 ms, static weights; `warmup` untimed steps,
    then `steps` timed steps with the firing read back every step."""
    t = time.perf_counter()
    spec = gen.torus_mixed()
    net = api.Network()
    spec.load(net)
    nsyn, nneu = int(spec.synapse_count), int(spec.neuron_count)
    del spec
    sim = api.Simulation(net, api.Configuration())
    print("mixed torus: %d neurons, %d synapses, built in %.1f s" % (nneu, nsyn, time.perf_counter() - t),
          file=sys.stderr, flush=True)
    if warmup:
        sim.run(warmup)
    default_shape = sim.launch_shape
    shape = sim.autotune(autotune_steps) if autotune_steps else default_shape
    sim.prepare(steps)
    t0 = time.perf_counter()
    counts, _, events = sim.run(steps)
    wall = time.perf_counter() - t0
    kern, work = sim.profile(profile_steps)
    del sim
    return {
        "workload": "torus_mixed_lif_qif_hh_poisson_51k_x_500",
        "neurons": nneu, "synapses": nsyn, "steps": steps, "warmup": warmup,
        "events_per_s": events / wall,
        "ms_per_step": 1000.0 * wall / steps,
        "sim_to_wall": steps * 1e-3 / wall,
        "firing_rate_hz": float(counts.sum()) / nneu / (steps * 1e-3),
        "events_per_step": events / steps,
        "kernel_us": kern,
        "launch_shape": list(shape), "default_launch_shape": list(default_shape),
    }


def differentiable(api, gen, steps=500, checkpoint_every=50, reps=3):
    """The differentiable workload: 10k Izhikevich neurons x 200 synapses
    (delays U{1..10}), FP32 weights and per-neuron a, b, d trainable, Gaussian
    input sigma 5/2, 500 steps; loss = MSE between the spike rates of the
    last 250 steps and target rates U[2,20] Hz.  Forward and adjoint timed
    separately with events on the stream both run on (median of reps)."""
    import torch
    from nemo_amd.differentiable import FloatNetwork
    n = 10000
    spec = gen.random_izhikevich(n, 200, dmax=10, seed=5)
    net = api.Network()
    spec.load(net)
    fnet = FloatNetwork(net, 0)
    dev = torch.device("cuda", 0)
    a, b, c, d, sg, u, v = [torch.tensor(np.asarray(x), device=dev) for x in spec.blocks[0].args]
    w = torch.tensor(spec.weight, device=dev).requires_grad_(True)
    a.requires_grad_(True)
    b.requires_grad_(True)
    d.requires_grad_(True)
    g = torch.Generator(device=dev).manual_seed(5)
    x = sg * torch.randn(steps, n, device=dev, generator=g)
    target = 2.0 + 18.0 * torch.rand(n, device=dev, generator=g)
    last = steps // 2

    def loss_of(S):
        rates = S[-last:].sum(0) / (last * 1e-3)
        return ((rates - target) ** 2).mean()

    fwd, bwd = [], []
    for r in range(reps + 1):
        for p in (w, a, b, d):
            p.grad = None
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        torch.cuda.synchronize()
        e0.record()
        S, V = fnet(w, a, b, d, x, c=c, v0=v, u0=u, checkpoint_every=checkpoint_every)
        loss = loss_of(S)
        e1.record()
        loss.backward()
        e2.record()
        torch.cuda.synchronize()
        if r:  # the first iteration warms up
            fwd.append(e0.elapsed_time(e1))
            bwd.append(e1.elapsed_time(e2))
    f, bk = float(np.median(fwd)), float(np.median(bwd))
    return {
        "workload": "izhikevich_float_10k_x_200_500_steps",
        "neurons": n, "synapses": int(spec.synapse_count), "steps": steps,
        "checkpoint_every": checkpoint_every,
        "forward_ms_per_step": f / steps,
        "adjoint_ms_per_step": bk / steps,
        "adjoint_over_forward": bk / f,
        "firing_rate_hz": float(S.detach().sum()) / n / (steps * 1e-3),
        "loss": float(loss),
        "grad_norms": {k: float(p.grad.norm()) for k, p in (("w", w), ("a", a), ("b", b), ("d", d))},
        "dtype": "f32 (neuron substeps in f64)",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--profile-steps", type=int, default=300)
    ap.add_argument("--neurons", type=int, default=100000)
    ap.add_argument("--synapses", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-nets", type=int, default=64, help="networks in the batched run (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--autotune-steps", type=int, default=128,
                    help="steps per autotuner trial of the batched and mixed runs (0: default shape)")
    ap.add_argument("--autotune-plain", type=int, default=0, metavar="STEPS",
                    help="autotune the plain run too, with STEPS per trial (default: its default launch shape, "
                         "whose traffic the committed counter profile measured)")
    ap.add_argument("--diff-steps", type=int, default=500, help="steps of the differentiable run (0: skip it)")
    ap.add_argument("--mixed-steps", type=int, default=2000, help="steps of the mixed-model torus run (0: skip it)")
    ap.add_argument("--only", choices=["batched", "differentiable", "mixed"],
                    help="run one of the secondary workloads alone and print its JSON (profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the HIP backend runs one simulation per device and has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import nemo_amd
    from nemo_amd import generators as gen

    if args.copy_probe:
        # counter calibration: torch's device-to-device copy of a known size
        import torch
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(10):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        gbps = 2.0 * 4 * n * 10 / (e0.elapsed_time(e1) * 1e-3) / 1e9
        print(json.dumps({"copy_probe_bytes": 4 * n, "launches": 11, "achieved_GBps": gbps}))
        return

    if args.only == "batched":
        print(json.dumps(batched(nemo_amd, gen, args.batch_nets or 64, args.batch_steps, args.autotune_steps)))
        return
    if args.only == "differentiable":
        print(json.dumps(differentiable(nemo_amd, gen, steps=args.diff_steps or 500)))
        return
    if args.only == "mixed":
        print(json.dumps(mixed(nemo_amd, gen, steps=args.mixed_steps or 2000, autotune_steps=args.autotune_steps)))
        return

    n, m = args.neurons, args.synapses
    t_gen = time.perf_counter()
    spec = gen.random_izhikevich(n, m, dmax=20, seed=2, stdp=True)
    t_build = time.perf_counter()
    sim = build_network(nemo_amd, spec, stdp_function())
    t_ready = time.perf_counter()
    print("built %d neurons, %d synapses (%d plastic) in %.1f s (generation %.1f s)"
          % (n, sim.synapse_count, sim.plastic_count, t_ready - t_build, t_build - t_gen), file=sys.stderr, flush=True)

    default_shape = sim.launch_shape
    shape = sim.autotune(args.autotune_plain) if args.autotune_plain else default_shape
    print("launch shape %s (default %s)" % (shape, default_shape), file=sys.stderr, flush=True)
    # the timed run's graphs are built first, so that the warm-up steps run
    # right before it (an idle GPU clocks down, and a short timed run would
    # pay for the ramp back up)
    sim.prepare(args.steps, stdp_period=1000)
    sim.run(args.warmup, stdp_period=1000, reward=1.0)

    t0 = time.perf_counter()
    counts, fired, events = sim.run(args.steps, stdp_period=1000, reward=1.0)
    wall = time.perf_counter() - t0
    print("timed %d steps in %.3f s" % (args.steps, wall), file=sys.stderr, flush=True)

    kern, work = sim.profile(args.profile_steps)
    steps_p = work["steps"]
    per_launch, per_step_us = launch_bytes(shape, kern, work, n, spec)
    dominant = max(per_step_us, key=per_step_us.get)
    achieved = per_launch[dominant] / (kern[dominant] * 1e-6) / 1e9
    stdp_bytes = BYTES_PER_STDP_SYNAPSE * work["stdp_synapses"] / steps_p
    step_bytes = BYTES_PER_EVENT * work["events"] / steps_p + BYTES_PER_IZH_NEURON * n + stdp_bytes

    result = {
        "metric": METRIC,
        "value": events / wall,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "izhikevich_80_20_stdp_100k_x_1000",
            "neurons": n, "synapses_per_neuron": m, "max_delay_ms": 20, "seed": 2,
            "stdp": "A+=0.1exp(-dt/20) A-=-0.12exp(-dt/20) 20ms windows, reward 1.0 every 1000 steps, w in [0,10]",
            "method": "plain (fixed-point Q11.20, bit-exact with the CPU backend)",
            "launch_shape": list(shape), "default_launch_shape": list(default_shape),
        },
        "sim_to_wall": args.steps * 1e-3 / wall,
        "firing_rate_hz": float(counts.sum()) / n / (args.steps * 1e-3),
        "events_per_step": events / args.steps,
        "kernel_us": kern,
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant, shape),
            "algorithmic_bytes_per_launch": per_launch[dominant],
        },
        "step_roofline_frac": (step_bytes / (wall / args.steps)) / 1e9 / HBM_PEAK_GBS,
    }
    del sim
    if args.batch_nets:
        result["batched"] = batched(nemo_amd, gen, args.batch_nets, args.batch_steps, args.autotune_steps)
    if args.diff_steps:
        result["differentiable"] = differentiable(nemo_amd, gen, steps=args.diff_steps)
        print("differentiable done", file=sys.stderr, flush=True)
    if args.mixed_steps:
        result["mixed"] = mixed(nemo_amd, gen, steps=args.mixed_steps, autotune_steps=args.autotune_steps)
        print("mixed done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cb = cpu_baseline(spec, args.warmup, args.steps, args.cpu_seconds)
        result["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "cpu_model", "libc")}
        result["speedup_vs_cpu"] = result["value"] / cb["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
