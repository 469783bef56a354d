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


def cpu_baseline(spec, seconds, max_steps):
    """The CPU oracle (restated reference CPU backend) on the same network, on
    this host's cores, over a bounded sample of steps from construction."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa
    sim = build_network(oa, spec, True)
    ev0 = sim.delivered_events()
    t0 = time.perf_counter()
    steps = 0
    while steps < max_steps and time.perf_counter() - t0 < seconds:
        sim.step()
        steps += 1
        if steps % 1000 == 0:
            sim.apply_stdp(1.0)
    wall = time.perf_counter() - t0
    events = sim.delivered_events() - ev0
    return {
        "value": events / wall,
        "unit": "events/s",
        "cores": int(oa.lib().oracle_num_threads()),
        "kind": "port",
        "sample": "first %d steps of the same network after construction (%.1f s wall, %.0f ms/step)"
                  % (steps, wall, 1000.0 * wall / max(1, steps)),
        "ms_per_step": 1000.0 * wall / max(1, steps),
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
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the HIP backend runs one simulation per device and this "
              "build has no multi-GPU path (see DESIGN.md on sharding)", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import nemo_amd
    from nemo_amd import generators as gen
    from nemo_amd._lib import check, lib

    if args.copy_probe:
        gbps = ctypes.c_double()
        check(lib().nemo_hip_copy_probe(args.copy_probe, 10, ctypes.byref(gbps)))
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "launches": 11, "achieved_GBps": gbps.value}))
        return

    n, m = args.neurons, args.synapses
    t_gen = time.perf_counter()
    spec = gen.random_izhikevich(n, m, dmax=20, seed=2, stdp=True)
    t_build = time.perf_counter()
    sim = build_network(nemo_amd, spec, stdp_function())
    t_ready = time.perf_counter()
    print("built %d neurons, %d synapses (%d plastic) in %.1f s (generation %.1f s)"
          % (n, sim.synapse_count, sim.plastic_count, t_ready - t_build, t_build - t_gen), file=sys.stderr, flush=True)

    sim.run(args.warmup, stdp_period=1000, reward=1.0)
    print("warmup done", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    counts, fired, events = sim.run(args.steps, stdp_period=1000, reward=1.0)
    wall = time.perf_counter() - t0
    print("timed %d steps in %.3f s" % (args.steps, wall), file=sys.stderr, flush=True)

    kern, work = sim.profile(args.profile_steps)
    steps_p = work["steps"]
    per_launch = {
        "propagate": (BYTES_PER_EVENT * work["events"] + BYTES_PER_STDP_SYNAPSE * work["stdp_synapses"]) / steps_p,
        "update": BYTES_PER_IZH_NEURON * n,
    }
    dominant = max(per_launch, key=lambda k: kern[k])
    achieved = per_launch[dominant] / (kern[dominant] * 1e-6) / 1e9
    step_bytes = sum(per_launch.values())

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
            "traffic": None,
            "algorithmic_bytes_per_launch": per_launch[dominant],
        },
        "step_roofline_frac": (step_bytes / (wall / args.steps)) / 1e9 / HBM_PEAK_GBS,
    }
    if not args.no_cpu_baseline:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cb = cpu_baseline(spec, args.cpu_seconds, 100000)
        result["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        result["speedup_vs_cpu"] = result["value"] / cb["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
