"""Benchmark: Brian2's clock-driven step on one MI355X through the hip target.

Workload (the 1M-neuron configuration): CUBA LIF, N = 1,000,000 (800k exc /
200k inh), fixed in-degree 1000 drawn without replacement (1e9 synapses,
int32 indices, fp32 weights), per-synapse delays U{1..50} steps, dt = 0.1 ms,
fp64 state (bit-exact verification mode).  One "step" is one simulation time
step of the whole network (state update, threshold, spike queue, propagation).

Prints one JSON line (metric, value, roofline, cpu_baseline, ...) on stdout;
progress goes to stderr.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n=20000, K=1000, warm=500, steps=5000):
    """The oracle (plain C restatement, 1 thread) on a bounded sample of the same
    workload: the same model, in-degree, delays and weights at N = n neurons."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np

    import oracle_py as O

    meta = dict(N=n, NE=n * 4 // 5, K=K, taum=0.02, taue=0.005, taui=0.01, El=-0.049, Vt=-0.05, Vr=-0.06,
                refractory=0.005, dt=1e-4, dmin=1, dmax=50, seed=1234, we_base=1.62e-3 * 80.0 / K,
                wi_base=-9e-3 * 80.0 / K * 0.4)
    G, paths = O.build_cuba_delay(meta)
    for p in paths:
        p.counts = np.zeros(len(p.pre), np.int32)
    O.run([G], paths, meta["dt"], warm, mon_group=None)
    before = sum(int(p.counts.sum()) for p in paths)
    t0 = time.perf_counter()
    O.run([G], paths, meta["dt"], steps, start_step=warm, mon_group=None)
    wall = time.perf_counter() - t0
    events = sum(int(p.counts.sum()) for p in paths) - before
    return dict(value=events / wall, unit="events/s", cores=1, kind="port",
                sample=f"oracle (C restatement, 1 thread): same model at N={n}, in-degree {K}, delays 1-50, "
                       f"{steps} steps after {warm} warm-up steps; {events} events in {wall:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--K", type=int, default=1000)
    ap.add_argument("--graph-steps", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--float", dest="float_t", default="double", choices=["double", "float"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(see DESIGN.md for how the network would shard)", file=sys.stderr)
        sys.exit(2)

    import torch

    from brian2_amd import synth

    log(f"device: {torch.cuda.get_device_name(0)}")
    t0 = time.perf_counter()
    c = synth.cuba_synw_network(N=args.N, K=args.K, float_t=args.float_t, graph_steps=args.graph_steps,
                                bucket_cap=None)
    torch.cuda.synchronize()
    log(f"built network: {c['n_syn']} synapses in {time.perf_counter() - t0:.1f} s")
    net = c["net"]
    t0 = time.perf_counter()
    net.run(args.warmup)
    torch.cuda.synchronize()
    log(f"warm-up {args.warmup} steps (includes kernel compilation, graph capture): {time.perf_counter() - t0:.2f} s")

    ev0 = net.n_events
    spk0 = c["mon"].mon_n.item()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.run(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    events = net.n_events - ev0
    spikes = c["mon"].mon_n.item() - spk0
    ms_per_step = wall / args.steps * 1e3
    rate = spikes / args.N / (args.steps * 1e-4)
    log(f"timed {args.steps} steps: {wall:.3f} s, {ms_per_step * 1e3:.1f} us/step, {events} events, "
        f"{spikes} spikes ({rate:.1f} Hz)")

    # per-kernel durations (HIP events around each launch on the plan's stream)
    ev_p0 = {s.name: s.n_events for s in net.paths}
    spk_p0 = c["mon"].mon_n.item()
    prof = net.profile(args.profile_steps)
    ev_p = {s.name: s.n_events - ev_p0[s.name] for s in net.paths}
    spk_p = c["mon"].mon_n.item() - spk_p0
    P = args.profile_steps
    fbytes = 8 if args.float_t == "double" else 4
    b_neuron = 3 * fbytes * 2 + 8 + 2       # v, ge, gi read+write, lastspike read, not_refractory r/w
    b_event = 4 + 4 + 1                     # post index, fp32 weight, delay code
    kernels = []
    for kind, name, ms in prof:
        avg_ms = ms / P
        if kind == "step":
            alg = args.N * b_neuron + spk_p / P * 4
        else:
            alg = ev_p[name] / P * b_event
        kernels.append(dict(kernel=f"{kind}_{name}", avg_us=avg_ms * 1e3, alg_bytes=alg,
                            gbs=alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0))
    for k in kernels:
        log(f"  {k['kernel']}: {k['avg_us']:.1f} us/launch, {k['alg_bytes'] / 1e6:.1f} MB algorithmic, "
            f"{k['gbs']:.0f} GB/s")
    dom = max(kernels, key=lambda k: k["avg_us"])
    step_bytes = args.N * b_neuron + events / args.steps * b_event + spikes / args.steps * 4
    step_frac = step_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9)
    log(f"whole step: {step_bytes / 1e6:.1f} MB algorithmic per step -> {step_frac * 100:.1f}% of HBM peak")

    traffic = None
    tfile = os.path.join(REPO, "profiles", "traffic_cuba1m.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(dom["kernel"])

    out = {
        "metric": "synaptic events/s",
        "value": events / wall,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.float_t == "double" else "f32",
        "data": "synthetic",
        "config": {"workload": "CUBA LIF 1M neurons, fixed in-degree 1000 (1e9 synapses), delays 1-50 steps, "
                               "fp32 weights, dt 0.1 ms (the 1M-neuron configuration)",
                   "N": args.N, "in_degree": args.K, "n_synapses": c["n_syn"], "sim_steps_timed": args.steps,
                   "mean_rate_hz": rate, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom["kernel"],
                     "avg_launch_us": dom["avg_us"], "step_frac": step_frac},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a bounded sample) ...")
        out["cpu_baseline"] = cpu_baseline()
        log(f"  {out['cpu_baseline']['value']:.3e} events/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
