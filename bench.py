"""Benchmark: Brian2's clock-driven step on one MI355X through the hip target.

Workload (the 1M-neuron configuration): CUBA LIF, N = 1,000,000 (800k exc /
200k inh), fixed in-degree 1000 drawn without replacement (1e9 synapses,
int32 indices, fp32 weights), per-synapse delays U{1..50} steps, dt = 0.1 ms,
fp64 state (bit-exact verification mode), 500 ms simulated.  One "step" is
one simulation time step of the whole network (state update, threshold, spike
queue, propagation).

Measurement: after W warm-up steps the network state is stored on the device,
K steps are timed with hipGraph replay (the value), the state is restored and
the same K steps are run again one launch at a time with HIP events around
each launch on the plan's stream (per-kernel durations and the roofline; the
simulation is deterministic, so both passes process identical events).

Second leg: the batched sweep (1024 CUBA networks, N ~ U{500..8000}, swept
excitatory weight): networks·simulated-seconds per second of the persistent
batched kernel vs the plain per-network path on a sample of the same networks,
with the sample's spikes checked identical.

Prints one JSON line on stdout; progress goes to stderr.

Third leg: the differentiable LIF network (256 trials of 200 ms): forward and
exact adjoint backward, trials x simulated seconds per second.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-fuse] [--no-sweep] [--no-adjoint] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
DT = 1e-4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n=20000, K=1000, warm=500, steps=5000):
    """The oracle (plain C restatement, 1 thread) on a bounded sample of the same
    workload: the same model, in-degree, delays and weights at N = n neurons."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np

    import oracle_py as O

    meta = dict(N=n, NE=n * 4 // 5, K=K, taum=0.02, taue=0.005, taui=0.01, El=-0.049, Vt=-0.05, Vr=-0.06,
                refractory=0.005, dt=DT, dmin=1, dmax=50, seed=1234, we_base=1.62e-3 * 80.0 / K,
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


def kernel_bytes(kind, n, events, spikes, fbytes):
    """Algorithmic bytes of one launch: neuron state read + write (v, ge, gi),
    lastspike read, not_refractory read + write; per spike the ring entry, the
    monitor record and the lastspike write; per event the post index, the fp32
    weight and the delay code.  A fused launch (group step and the gathers into
    the group) moves both."""
    step = n * (3 * fbytes * 2 + 8 + 2) + spikes * (4 + 8 + 8)
    if kind == "step":
        return step
    gather = events * (4 + 4 + 1)
    return step + gather if kind == "fused" else gather


def sweep_bench(n_nets=1024, steps=10000, sample=16, sample_steps=1000, float_t="double"):
    """Batched sweep vs the plain per-network path (nets·sim-s/s)."""
    import numpy as np
    import torch

    from brian2_amd import batched, synth

    nets = synth.sweep_networks(n_nets=n_nets)
    B = batched.BatchedCUBA(nets, float_t=float_t)
    t0 = time.perf_counter()
    B.build()
    best, res = B.autotune(trial_steps=200)
    log(f"sweep: {n_nets} networks ({sum(n.N for n in nets)} neurons, {B.n_syn} synapses) built and tuned in "
        f"{time.perf_counter() - t0:.1f} s: block {best[0]}, {best[1]} steps per launch")
    steps -= steps % B.steps_per_launch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.run(steps)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    rate_b = n_nets * steps * DT / wall_b
    log(f"sweep: batched {steps} steps in {wall_b:.2f} s -> {rate_b:.1f} nets*sim-s/s")
    idx = np.unique(np.linspace(0, n_nets - 1, sample).astype(int))
    wall_p, sim_p, same = 0.0, 0.0, True
    for k in idx:
        net = nets[k]
        c = synth.cuba_network(N=net.N, NE=net.NE, seed=net.seed, we=net.namespace["we"], float_t=float_t)
        c["net"].run(10)  # kernel compilation and graph capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c["net"].run(sample_steps)
        torch.cuda.synchronize()
        wall_p += time.perf_counter() - t0
        sim_p += sample_steps * DT
        i, s = B.spikes(k)
        sel = s < 10 + sample_steps
        o = np.lexsort((i[sel], s[sel]))
        same &= bool(np.array_equal(i[sel][o], c["mon"].i) and np.array_equal(s[sel][o], c["mon"].step))
        del c
    rate_p = sim_p / wall_p
    log(f"sweep: plain path {len(idx)} networks x {sample_steps} steps in {wall_p:.2f} s -> {rate_p:.2f} "
        f"nets*sim-s/s; spikes identical: {same}")
    return {"metric": "nets*sim-s/s", "value": rate_b, "unit": "networks x simulated s / s",
            "plain_value": rate_p, "vs_plain": rate_b / rate_p, "spikes_identical": same,
            "dtype": "f64" if float_t == "double" else "f32", "n_nets": n_nets, "steps": steps,
            "block": best[0], "steps_per_launch": best[1],
            "plain_sample": f"{len(idx)} of the {n_nets} networks (evenly spaced), {sample_steps} steps each, "
                            f"one Network per network, hipGraph replay of 100 steps"}


def adjoint_bench(trials=256):
    """Differentiable LIF network: forward + exact adjoint backward over all trials."""
    import torch

    from brian2_amd.adjoint import LIFNetwork

    t0 = time.perf_counter()
    net = LIFNetwork(trials=trials)
    loss, saved = net.forward(net.w_in0, net.w_rec0, net.c["tau_m"], net.c["tau_s"])
    net.backward(saved)
    torch.cuda.synchronize()
    log(f"adjoint: {trials} trials built and warmed in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    loss, saved = net.forward(net.w_in0, net.w_rec0, net.c["tau_m"], net.c["tau_s"])
    torch.cuda.synchronize()
    t_f = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = net.backward(saved)
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    sim = trials * net.T * net.dt
    log(f"adjoint: forward {t_f * 1e3:.1f} ms, backward {t_b * 1e3:.1f} ms for {trials} trials x {net.T} steps; "
        f"loss {loss:.6g}, dL/dtau_m {g[2]:.6g}, dL/dtau_s {g[3]:.6g}")
    return {"metric": "trials*sim-s/s (forward + adjoint backward)", "value": sim / (t_f + t_b),
            "unit": "trials x simulated s / s", "trials": trials, "steps": net.T, "forward_ms": t_f * 1e3,
            "backward_ms": t_b * 1e3, "ckpt_every": net.C, "loss": loss, "dtype": "f64 state, f32 weights"}


def copy_calibration(nbytes=256 << 20, reps=3):
    """A plain device-to-device copy of known size (hipMemcpyAsync), so that a
    counter run (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) can scale the counters
    of the simulation kernels by what they read for exactly nbytes."""
    import ctypes

    import torch

    from brian2_amd import _lib

    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(reps):
        _lib.check(_lib.lib.hb_memcpy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()), nbytes, 3,
                                      stream), "copy")
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} device copies of {nbytes} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--K", type=int, default=1000)
    ap.add_argument("--graph-steps", type=int, default=100)
    ap.add_argument("--float", dest="float_t", default="double", choices=["double", "float"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fuse", action="store_true",
                    help="the group step and each pathway's gather as separate launches (gathers on side streams)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="start with device copies of known size (for counter runs)")
    ap.add_argument("--sweep-steps", type=int, default=10000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path (a network would shard by "
              "neuron range, exchanging each step's spike lists)", file=sys.stderr)
        sys.exit(2)

    import torch

    from brian2_amd import synth

    log(f"device: {torch.cuda.get_device_name(0)}")
    if args.copy_calibration:
        copy_calibration()
    t0 = time.perf_counter()
    c = synth.cuba_synw_network(N=args.N, K=args.K, float_t=args.float_t, graph_steps=args.graph_steps,
                                network_options={"fuse": not args.no_fuse})
    torch.cuda.synchronize()
    log(f"built network: {c['n_syn']} synapses in {time.perf_counter() - t0:.1f} s")
    net, mon = c["net"], c["mon"]
    t0 = time.perf_counter()
    net.run(args.warmup)
    torch.cuda.synchronize()
    log(f"warm-up {args.warmup} steps (includes kernel compilation, graph capture): {time.perf_counter() - t0:.2f} s")
    net.store()

    ev0 = net.n_events
    spk0 = int(mon.mon_n.item())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.run(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    events = net.n_events - ev0
    spikes = int(mon.mon_n.item()) - spk0
    ms_per_step = wall / args.steps * 1e3
    rate = spikes / args.N / (args.steps * DT)
    log(f"timed {args.steps} steps: {wall:.3f} s, {ms_per_step * 1e3:.1f} us/step, {events} events, "
        f"{spikes} spikes ({rate:.1f} Hz)")

    # the same steps again, one launch at a time, HIP events around each launch
    net.restore()
    ev_p0 = {s.name: s.n_events for s in net.paths}
    prof = net.profile(args.steps)
    net.run(0)
    ev_p = {s.name: s.n_events - ev_p0[s.name] for s in net.paths}
    same = sum(ev_p.values()) == events and int(mon.mon_n.item()) - spk0 == spikes
    log(f"profiled pass: {sum(ev_p.values())} events, identical to the timed pass: {same}")
    fbytes = 8 if args.float_t == "double" else 4
    kernels = []
    for kind, name, ms in prof:
        avg_us = ms / args.steps * 1e3
        # a fused launch runs the group's step and the gathers of every pathway into it
        ev = sum(ev_p[s.name] for s in net.paths if s.target.name == name) if kind == "fused" else ev_p.get(name, 0)
        alg = kernel_bytes(kind, args.N, ev / args.steps, spikes / args.steps, fbytes)
        kernels.append(dict(kernel=f"{kind}_{name}", avg_us=avg_us, alg_bytes=alg,
                            gbs=alg / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0))
    for k in kernels:
        log(f"  {k['kernel']}: {k['avg_us']:.1f} us/launch, {k['alg_bytes'] / 1e6:.1f} MB algorithmic, "
            f"{k['gbs']:.0f} GB/s")
    dom = max(kernels, key=lambda k: k["avg_us"])
    step_bytes = kernel_bytes("step", args.N, 0, spikes / args.steps, fbytes) + \
        kernel_bytes("gather", 0, events / args.steps, 0, fbytes)
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
                               "fp32 weights, dt 0.1 ms, 500 ms simulated (the 1M-neuron configuration)",
                   "N": args.N, "in_degree": args.K, "n_synapses": c["n_syn"], "sim_steps_timed": args.steps,
                   "mean_rate_hz": rate, "spikes": spikes, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom["kernel"],
                     "avg_launch_us": dom["avg_us"], "step_frac": step_frac},
        "kernels": kernels,
        "profiled_pass_identical": same,
    }
    del c, net, mon
    if not args.no_sweep:
        import gc

        gc.collect()
        torch.cuda.empty_cache()
        out["batched"] = sweep_bench(steps=args.sweep_steps)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a bounded sample) ...")
        out["cpu_baseline"] = cpu_baseline()
        log(f"  {out['cpu_baseline']['value']:.3e} events/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
