"""Benchmark: Brian2's clock-driven step on one MI355X through the hip target.

Workload (the 1M-neuron configuration): CUBA LIF, N = 1,000,000 (800k exc /
200k inh), fixed in-degree 1000 drawn without replacement (1e9 synapses,
int32 indices, fp32 weights), per-synapse delays U{1..50} steps, dt = 0.1 ms,
fp64 state (bit-exact verification mode).

One bench step = 100 simulation time steps (10 ms simulated), replayed as one
hipGraph of 100 step launches.  The network alternates between bursts and
quiet epochs of a few ms each, so one simulation step is not a representative
unit of work; 100 of them are.  Before the timed region the network settles
for --settle simulation steps (default 2000, far beyond the 50-step delay
window and the start-up transient) plus W warm-up bench steps, all untimed.
Then exactly K bench steps are timed (K x 100 simulation steps), bracketed by
torch.cuda.synchronize().

Roofline: the step kernel of this workload is one launch per simulation step
(`fused_neurongroup`: the group's tiles and both pathways' gathers).  Its
average duration is the device time of the timed region, taken from HIP
events recorded on the plan's stream around the graph launches (host round
trips between chunks excluded), divided by the launches; its algorithmic
bytes per launch are N x 58 B + 20 B per spike + 9 B per synaptic event
(fp64 state read and written, lastspike read, not_refractory read and
written; per spike the ring entry, the monitor record and lastspike; per
event the target index, the fp32 weight and the delay code), from the events
and spikes counted in that region.  A second pass replays the same steps
from the stored state one launch at a time with HIP events around each launch
(per-kernel figures; identical events and spikes are checked).

Further legs: the batched sweep (1024 CUBA networks, fp64 and fp32) against
the plain per-network path; the differentiable LIF network (forward +
adjoint backward); the fp32 mode's error against fp64 per configuration; the
CPU baseline (the C restatement, one thread and OpenMP) on a bounded sample.

Prints one JSON line on stdout; progress goes to stderr.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--settle S] [--no-fuse] [--float float]
                    [--no-sweep] [--no-adjoint] [--no-fp32-error] [--no-cpu-baseline] [--no-profile-pass]
"""
import argparse
import hashlib
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


def cpu_baseline(n=100000, K=1000, warm=1000, steps=1000):
    """The C restatement of the step (the oracle) on a bounded sample of the same
    workload: the same model, in-degree, delays and weights at N = n neurons (a
    tenth of the configuration), built and settled once, then the same steps timed
    on one thread and on every core of the host (OpenMP), each from the settled
    state with the settled spikes in flight."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np

    import oracle_py as O

    meta = dict(N=n, NE=n * 4 // 5, K=K, taum=0.02, taue=0.005, taui=0.01, El=-0.049, Vt=-0.05, Vr=-0.06,
                refractory=0.005, dt=DT, dmin=1, dmax=50, seed=1234, we_base=1.62e-3 * 80.0 / K,
                wi_base=-9e-3 * 80.0 / K * 0.4)
    ncpu = O.cpu_threads()
    G, paths = O.build_cuba_delay(meta)
    for p in paths:
        p.counts = np.zeros(len(p.pre), np.int32)
    O.run([G], paths, meta["dt"], warm, mon_group=None, threads=ncpu)
    settled = (G.state.copy(), G.lastspike.copy(), G.not_refractory.copy())
    out = {}
    for threads in (1, ncpu):
        G.state[:], G.lastspike[:], G.not_refractory[:] = settled
        qs = O.OracleQueues(len(paths))  # fresh queues: the first 50 steps miss the settle's spikes in flight
        before = sum(int(p.counts.sum()) for p in paths)
        t0 = time.perf_counter()
        O.run([G], paths, meta["dt"], steps, start_step=warm, mon_group=None, threads=threads, queues=qs)
        wall = time.perf_counter() - t0
        events = sum(int(p.counts.sum()) for p in paths) - before
        out[threads] = (events / wall, events, wall, G.state.copy())
        log(f"  cpu baseline, {threads} thread(s): {events} events in {wall:.2f} s -> {events / wall:.3e} events/s")
    same = bool(np.array_equal(out[1][3], out[ncpu][3]))
    v1, vn = out[1][0], out[ncpu][0]
    return dict(value=vn, unit="events/s", cores=ncpu, kind="port",
                single_thread=v1, threads_used=ncpu, host_cpu_count=os.cpu_count(),
                sample=f"plain CPU path (restatement): the oracle (C, -O2, no FMA contraction) on the same model at "
                       f"N={n}, in-degree {K}, delays 1-50, {steps} steps after {warm} settle steps; "
                       f"{out[ncpu][1]} events in {out[ncpu][2]:.1f} s on {ncpu} OpenMP threads, "
                       f"{out[1][1]} in {out[1][2]:.1f} s on one; final state identical: {same}")


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


def sweep_bench(n_nets=1024, steps=10000, sample=16, sample_steps=1000, float_t="double", ref_counts=None):
    """Batched sweep vs the plain per-network path (nets·sim-s/s)."""
    import numpy as np
    import torch

    from brian2_amd import batched, synth

    nets = synth.sweep_networks(n_nets=n_nets)
    B = batched.BatchedCUBA(nets, float_t=float_t)
    t0 = time.perf_counter()
    B.build()
    best, res = B.autotune(trial_steps=200)
    log(f"sweep ({float_t}): {n_nets} networks ({sum(n.N for n in nets)} neurons, {B.n_syn} synapses) built and "
        f"tuned in {time.perf_counter() - t0:.1f} s: block {best[0]}, {best[1]} steps per launch")
    steps -= steps % B.steps_per_launch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.run(steps)
    torch.cuda.synchronize()
    wall_b = time.perf_counter() - t0
    rate_b = n_nets * steps * DT / wall_b
    log(f"sweep ({float_t}): batched {steps} steps in {wall_b:.2f} s -> {rate_b:.1f} nets*sim-s/s")
    counts = B.spk_n.cpu().numpy().astype(np.int64)
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
    log(f"sweep ({float_t}): plain path {len(idx)} networks x {sample_steps} steps in {wall_p:.2f} s -> "
        f"{rate_p:.2f} nets*sim-s/s; spikes identical: {same}")
    out = {"metric": "nets*sim-s/s", "value": rate_b, "unit": "networks x simulated s / s",
           "plain_value": rate_p, "vs_plain": rate_b / rate_p, "spikes_identical": same,
           "dtype": "f64" if float_t == "double" else "f32", "n_nets": n_nets, "steps": steps,
           "n_neurons": int(sum(n.N for n in nets)), "n_synapses": int(B.n_syn),
           "block": best[0], "steps_per_launch": best[1],
           "plain_sample": f"{len(idx)} of the {n_nets} networks (evenly spaced), {sample_steps} steps each, "
                           f"one Network per network, hipGraph replay of 100 steps"}
    if ref_counts is not None:
        d = counts - ref_counts
        out["fp32_error"] = {"vs": "the fp64 sweep (same networks, same steps)",
                             "spikes_f64": int(ref_counts.sum()), "spikes_f32": int(counts.sum()),
                             "total_rel_delta": float(d.sum() / max(1, ref_counts.sum())),
                             "mean_abs_rel_delta_per_net": float(np.mean(np.abs(d) / np.maximum(1, ref_counts))),
                             "nets_with_equal_counts": int((d == 0).sum())}
    del B
    return out, counts


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


def fp32_error_pair(make, n_steps, chunk, variables, fine_steps=2000):
    """Runs one configuration in fp64 and in fp32 side by side.  Step by step
    until the spike lists first differ (at most fine_steps steps), comparing
    every state variable after each step; then in chunks to the end.  Returns
    the spike counts of both, the first step at which the spike lists differ
    (None: never), and per state variable the largest |x32 - x64| over the steps
    before that one, absolute and relative to max |x64|."""
    import numpy as np

    c64, c32 = make("double"), make("float")
    err = {v: [0.0, 0.0] for v in variables}
    first_div = None
    done = 0
    compared = 0
    while done < n_steps:
        fine = first_div is None and done < fine_steps
        k = 1 if fine else min(chunk, n_steps - done)
        c64["net"].run(k)
        c32["net"].run(k)
        done += k
        if not fine:
            continue
        n64, n32 = int(c64["mon"].mon_n.item()), int(c32["mon"].mon_n.item())
        if n64 != n32 or (n64 and not (np.array_equal(c64["mon"].i, c32["mon"].i))):
            first_div = done - 1
            continue
        compared += 1
        for v in variables:
            a, b = c64["G"].get(v), c32["G"].get(v).astype(np.float64)
            scale = float(np.max(np.abs(a))) or 1.0
            e = float(np.max(np.abs(a - b)))
            err[v][0] = max(err[v][0], e)
            err[v][1] = max(err[v][1], e / scale)
    n64, n32 = len(c64["mon"].i), len(c32["mon"].i)
    del c64, c32
    return {"steps": n_steps, "spikes_f64": n64, "spikes_f32": n32,
            "spike_count_rel_delta": (n32 - n64) / max(1, n64),
            "first_divergent_step": first_div, "steps_compared_before_divergence": compared,
            "max_abs_err_before_divergence": {v: err[v][0] for v in variables},
            "max_rel_err_before_divergence": {v: err[v][1] for v in variables}}


def fp32_errors():
    """The fp32 mode's error against fp64 (the verification mode) per configuration."""
    import gc

    import torch

    from brian2_amd import synth

    out = {}
    t0 = time.perf_counter()
    out["cuba4000"] = fp32_error_pair(lambda ft: synth.cuba_network(N=4000, float_t=ft), 10000, 100,
                                      ["v", "ge", "gi"])
    out["cobahh4000"] = fp32_error_pair(lambda ft: synth.cobahh_network(N=4000, float_t=ft), 10000, 100,
                                        ["v", "m", "n", "h", "ge", "gi"])
    out["cuba_1m_at_100k"] = fp32_error_pair(
        lambda ft: synth.cuba_synw_network(N=100_000, K=1000, float_t=ft), 5000, 250, ["v", "ge", "gi"])
    out["cuba_1m_at_100k"]["note"] = ("the 1M-neuron configuration's model, in-degree, delays and weights at "
                                      "N = 100,000")
    gc.collect()
    torch.cuda.empty_cache()
    log(f"fp32 error record: {time.perf_counter() - t0:.1f} s")
    for k, v in out.items():
        log(f"  {k}: spikes {v['spikes_f64']} (f64) vs {v['spikes_f32']} (f32), first divergent step "
            f"{v['first_divergent_step']}, max rel err before it {v['max_rel_err_before_divergence']}")
    return out


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


def measured_traffic(kernel, sha1, workload):
    """HBM bytes per launch of `kernel` from the committed counter runs, only when
    they measured this very kernel source (sha1) on this workload; else None."""
    tfile = os.path.join(REPO, "profiles", "traffic_cuba1m.json")
    if not os.path.exists(tfile):
        return None
    with open(tfile) as f:
        t = json.load(f)
    entries = t.get(kernel) or []
    for e in entries if isinstance(entries, list) else [entries]:
        if e.get("kernel_sha1") == sha1 and e.get("workload") == workload:
            return {k: e[k] for k in ("fetch_bytes", "write_bytes", "bytes", "dispatches", "kernel_sha1") if k in e}
    return None


def _rocprof_rows(d, suffix):
    import csv
    import glob

    rows = []
    for fn in sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)):
        with open(fn, newline="") as f:
            rows += list(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no *{suffix} under {d}")
    return rows


def summarize_trace(d, last=0):
    """Markdown table of a `rocprofv3 --kernel-trace --output-format csv` run: per
    kernel the dispatches, total / average / median / min / max duration, over
    every dispatch and, when last > 0, over each kernel's last `last` dispatches
    (the timed steps of a bench run with --no-profile-pass are the step kernel's
    last dispatches); for the kernel with the most GPU time, the idle gaps between
    its launches in that window."""
    import numpy as np

    rows = _rocprof_rows(d, "kernel_trace.csv")
    by = {}
    for r in rows:
        by.setdefault(r["Kernel_Name"], []).append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    for k in by:
        by[k].sort()
    total = sum(e - s for v in by.values() for s, e in v)
    order = sorted(by, key=lambda k: -sum(e - s for s, e in by[k]))
    out = []
    for title, sel in (("every dispatch", lambda v: v), (f"last {last} dispatches of each kernel",
                                                         lambda v: v[-last:])):
        if title.startswith("last") and last <= 0:
            continue
        out += [f"{title}:", "", "| kernel | calls | total ms | avg us | median us | min us | max us | % of GPU time |",
                "|---|---|---|---|---|---|---|---|"]
        for k in order:
            dur = np.array([e - s for s, e in sel(by[k])], np.float64) / 1e3
            name = k if len(k) <= 60 else k[:57] + "..."
            out.append(f"| `{name}` | {len(dur)} | {dur.sum() / 1e3:.3f} | {dur.mean():.2f} | {np.median(dur):.2f} | "
                       f"{dur.min():.2f} | {dur.max():.2f} | {dur.sum() / 1e3 / (total / 1e6) * 100:.1f} |")
        out.append("")
    top = by[order[0]][-last:] if last > 0 else by[order[0]]
    st = np.array([s for s, _ in top], np.int64)
    en = np.array([e for _, e in top], np.int64)
    if len(top) > 1:
        gap = (st[1:] - en[:-1]) / 1e3
        out.append(f"`{order[0][:60]}` ({len(top)} dispatches): start-to-start {np.mean(np.diff(st)) / 1e3:.2f} us on "
                   f"average; idle between launches {gap.mean():.2f} us on average, {np.median(gap):.2f} us median; "
                   f"first start to last end {(en[-1] - st[0]) / 1e6:.3f} ms")
    return "\n".join(out)


def summarize_counters(d, counter, last=0, copy_bytes=256 << 20):
    """Per-dispatch average of one counter (`rocprofv3 --pmc <counter> --output-format
    csv`) per kernel (over each kernel's last `last` dispatches when last > 0), in
    KiB and in bytes scaled by what the same counter reads for a device copy of
    copy_bytes (the largest copy-kernel dispatch of the run: bench.py
    --copy-calibration's copies)."""
    import numpy as np

    rows = _rocprof_rows(d, "counter_collection.csv")
    by = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        did = int(r.get("Dispatch_Id") or r.get("Correlation_Id"))
        per = by.setdefault(r["Kernel_Name"], {})
        per[did] = per.get(did, 0.0) + float(r["Counter_Value"])
    copy = [k for k in by if k.startswith("__amd_rocclr_copyBuffer")]
    if not copy:
        raise SystemExit("no copy kernel in the counter run (run the bench with --copy-calibration)")
    copy_kib = max(max(by[k].values()) for k in copy)
    factor = copy_bytes / 1024.0 / copy_kib
    out = {"_calibration": {"counter": counter, "copy_bytes": copy_bytes, "copy_kib_per_dispatch": copy_kib,
                            "factor": factor}}
    for k, v in by.items():
        vals = [v[i] for i in sorted(v)]
        if last > 0:
            vals = vals[-last:]
        kib = float(np.mean(vals))
        out[k[:80]] = {"dispatches": len(vals), "kib_per_dispatch": kib, "bytes_per_dispatch": kib * 1024.0 * factor}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed bench steps (each 100 simulation steps)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed bench steps after the settle")
    ap.add_argument("--settle", type=int, default=2000, help="untimed simulation steps before the warm-up")
    ap.add_argument("--sim-steps-per-step", type=int, default=100)
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--K", type=int, default=1000)
    ap.add_argument("--float", dest="float_t", default="double", choices=["double", "float"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fuse", action="store_true",
                    help="the group step and each pathway's gather as separate launches (gathers on side streams)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-fp32-error", action="store_true")
    ap.add_argument("--no-profile-pass", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="start with device copies of known size (for counter runs)")
    ap.add_argument("--sweep-steps", type=int, default=10000)
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print a markdown summary of a rocprofv3 kernel trace (csv) and exit")
    ap.add_argument("--summarize-counters", nargs=2, metavar=("COUNTER", "DIR"),
                    help="print per-dispatch averages of a rocprofv3 counter run (csv) as JSON and exit")
    ap.add_argument("--last", type=int, default=0,
                    help="with --summarize-*: also (trace) or only (counters) each kernel's last LAST dispatches")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace, args.last))
        return
    if args.summarize_counters:
        print(json.dumps(summarize_counters(args.summarize_counters[1], args.summarize_counters[0], args.last),
                         indent=1))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path (a network would shard by "
              "neuron range, exchanging each step's spike lists)", file=sys.stderr)
        sys.exit(2)

    import torch

    from brian2_amd import synth

    SPS = args.sim_steps_per_step
    log(f"device: {torch.cuda.get_device_name(0)}")
    if args.copy_calibration:
        copy_calibration()
    t0 = time.perf_counter()
    c = synth.cuba_synw_network(N=args.N, K=args.K, float_t=args.float_t, graph_steps=SPS,
                                network_options={"fuse": not args.no_fuse})
    torch.cuda.synchronize()
    log(f"built network: {c['n_syn']} synapses in {time.perf_counter() - t0:.1f} s")
    net, mon = c["net"], c["mon"]
    t0 = time.perf_counter()
    net.run(args.settle + args.warmup * SPS)
    torch.cuda.synchronize()
    log(f"settle {args.settle} + warm-up {args.warmup} x {SPS} simulation steps (includes kernel compilation, graph "
        f"capture): {time.perf_counter() - t0:.2f} s")
    net.store()
    step0 = net.t_step

    K = args.steps
    n_sim = K * SPS
    ev0 = net.n_events
    spk0 = int(mon.mon_n.item())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.run(n_sim, timing=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = net.last_run_gpu_ms
    events = net.n_events - ev0
    spikes = int(mon.mon_n.item()) - spk0
    rate = spikes / args.N / (n_sim * DT)
    log(f"timed {K} x {SPS} simulation steps: {wall:.3f} s wall, {gpu_ms:.2f} ms device time of the step launches, "
        f"{wall / n_sim * 1e6:.1f} us/simulation step, {events} events ({events / n_sim:.3e} per simulation step), "
        f"{spikes} spikes ({rate:.1f} Hz)")

    fbytes = 8 if args.float_t == "double" else 4
    G = c["G"]
    fused = getattr(G, "fused", None) is not None
    kernels = []
    same = None
    if not args.no_profile_pass:
        # the same steps again, one launch at a time, HIP events around each launch
        net.restore()
        ev_p0 = {s.name: s.n_events for s in net.paths}
        prof = net.profile(n_sim)
        net.run(0)
        ev_p = {s.name: s.n_events - ev_p0[s.name] for s in net.paths}
        same = sum(ev_p.values()) == events and int(mon.mon_n.item()) - spk0 == spikes
        log(f"profiled pass: {sum(ev_p.values())} events, identical to the timed pass: {same}")
        for kind, name, ms in prof:
            avg_us = ms / n_sim * 1e3
            ev = sum(ev_p[s.name] for s in net.paths if s.target.name == name) if kind == "fused" else ev_p.get(name, 0)
            alg = kernel_bytes(kind, args.N, ev / n_sim, spikes / n_sim, fbytes)
            kernels.append(dict(kernel=f"{kind}_{name}", avg_us=avg_us, alg_bytes=alg,
                                gbs=alg / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0))
        for k in kernels:
            log(f"  {k['kernel']}: {k['avg_us']:.1f} us/launch (one launch at a time), {k['alg_bytes'] / 1e6:.1f} MB "
                f"algorithmic, {k['gbs']:.0f} GB/s")
    step_bytes = kernel_bytes("step", args.N, 0, spikes / n_sim, fbytes) + \
        kernel_bytes("gather", 0, events / n_sim, 0, fbytes)
    if fused:
        # one launch per simulation step: the timed region's device time per launch
        dom_name = f"fused_{G.name}"
        avg_us = gpu_ms * 1e3 / n_sim
        alg = step_bytes
        src_sha1 = hashlib.sha1(G.fused.source.encode()).hexdigest()
        avg_src = "HIP events on the plan's stream around the timed region's graph launches / launches"
    else:
        dom = max(kernels, key=lambda k: k["avg_us"]) if kernels else None
        dom_name = dom["kernel"] if dom else None
        avg_us = dom["avg_us"] if dom else float("nan")
        alg = dom["alg_bytes"] if dom else float("nan")
        src_sha1 = None
        avg_src = "HIP events around each launch of the profiled pass"
    gbs = alg / (avg_us * 1e-6) / 1e9
    step_frac = step_bytes / (gpu_ms * 1e-3 / n_sim) / (HBM_PEAK_GBS * 1e9)
    log(f"roofline: {dom_name} {avg_us:.2f} us/launch, {alg / 1e6:.2f} MB algorithmic -> {gbs:.0f} GB/s "
        f"({gbs / HBM_PEAK_GBS * 100:.1f}% of 8 TB/s)")
    workload = f"cuba_synw N={args.N} K={args.K} {args.float_t} settle={args.settle + args.warmup * SPS} steps={n_sim}"

    out = {
        "metric": "synaptic events/s",
        "value": events / wall,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.float_t == "double" else "f32",
        "data": "synthetic",
        "config": {"workload": f"CUBA LIF {args.N} neurons, fixed in-degree {args.K} ({c['n_syn']} synapses), delays "
                               f"1-50 steps, fp32 weights, dt 0.1 ms (the 1M-neuron configuration); timed: {n_sim} "
                               f"simulation steps ({n_sim * DT * 1e3:.0f} ms simulated) after {step0} untimed ones",
                   "step": f"{SPS} simulation steps ({SPS * DT * 1e3:.0f} ms simulated), one hipGraph replay",
                   "N": args.N, "in_degree": args.K, "n_synapses": c["n_syn"], "settle_steps": args.settle,
                   "warmup_sim_steps": args.warmup * SPS, "sim_steps_timed": n_sim,
                   "events_per_sim_step": events / n_sim, "mean_rate_hz": rate, "spikes": spikes,
                   "us_per_sim_step": wall / n_sim * 1e6, "launches": "fused" if fused else "separate",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": measured_traffic(dom_name, src_sha1, workload),
                     "kernel": dom_name, "avg_launch_us": avg_us, "alg_bytes_per_launch": alg,
                     "avg_launch_source": avg_src, "kernel_sha1": src_sha1, "step_frac": step_frac,
                     "workload_key": workload},
        "kernels_profiled_pass": kernels,
        "profiled_pass_identical": same,
    }
    del c, net, mon, G
    import gc

    gc.collect()
    torch.cuda.empty_cache()
    if not args.no_sweep:
        out["batched"], ref = sweep_bench(steps=args.sweep_steps)
        gc.collect()
        torch.cuda.empty_cache()
        out["batched"]["fp32"], _ = sweep_bench(steps=args.sweep_steps, float_t="float", ref_counts=ref)
        gc.collect()
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
    if not args.no_fp32_error:
        out["fp32_error"] = fp32_errors()
        if "batched" in out:
            out["fp32_error"]["sweep_1024"] = out["batched"]["fp32"].get("fp32_error")
    if not args.no_cpu_baseline:
        log("cpu baseline (the restatement on a bounded sample) ...")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
