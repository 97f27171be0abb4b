"""ctypes wrapper of the oracle (oracle/libhbref.so) for the tests.

Test infrastructure only: the oracle is the checker, never the thing measured.
"""
import ast
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "libhbref.so")

i32, i64, u64, dbl, u8 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint8
P = ctypes.POINTER

HBR_CUBA_EXACT, HBR_COBAHH_EE, HBR_LIF_EXACT, HBR_SPIKEGEN, HBR_PASSIVE, HBR_POISSON = 1, 2, 3, 4, 5, 6
S_POISSON_GROUP = 17  # generator stream of PoissonGroup draws


class Group(ctypes.Structure):
    _fields_ = [
        ("model", i32), ("N", i32), ("state", P(dbl)), ("lastspike", P(dbl)),
        ("not_refractory", P(u8)), ("params", P(dbl)), ("n_gen", i64), ("gen_index", P(i32)),
        ("gen_timebin", P(i32)), ("gen_lastindex", i64), ("poisson_seed", u64), ("poisson_stream", u64),
        ("poisson_rdt", P(dbl)),
    ]


class Pathway(ctypes.Structure):
    _fields_ = [
        ("source_group", i32), ("target_group", i32), ("source_start", i32), ("source_end", i32),
        ("n_syn", i64), ("pre", P(i32)), ("post", P(i32)), ("w", P(dbl)), ("w_scalar", dbl),
        ("delay", P(dbl)), ("delay_scalar", dbl), ("target_var", i32), ("counts", P(i32)),
    ]


class Sim(ctypes.Structure):
    _fields_ = [
        ("dt", dbl), ("start_step", i64), ("n_steps", i64), ("n_groups", i32), ("groups", P(Group)),
        ("n_paths", i32), ("paths", P(Pathway)), ("mon_group", i32), ("spk_i", P(i32)),
        ("spk_step", P(i64)), ("spk_cap", i64), ("n_spk", i64), ("rec_group", i32), ("n_rec", i32),
        ("n_rec_vars", i32), ("rec_idx", P(i32)), ("rec_vars", P(i32)), ("rec_out", P(dbl)),
        ("threads", i32), ("queues", P(ctypes.c_void_p)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        L = ctypes.CDLL(LIB_PATH)
        L.hbr_rand64.restype = u64
        L.hbr_rand64.argtypes = [u64, u64, u64]
        L.hbr_uniform.restype = dbl
        L.hbr_uniform.argtypes = [u64, u64, u64]
        L.hbr_uniform_fill.argtypes = [u64, u64, u64, i64, P(dbl)]
        L.hbr_uniform_int.restype = ctypes.c_uint32
        L.hbr_uniform_int.argtypes = [u64, ctypes.c_uint32]
        L.hbr_bernoulli.restype = i64
        L.hbr_bernoulli.argtypes = [u64, u64, i32, i32, i32, dbl, P(i32), P(i32), i64]
        L.hbr_fixed_indegree.restype = ctypes.c_int
        L.hbr_fixed_indegree.argtypes = [u64, u64, i32, i32, i32, P(i32)]
        L.hbr_pair_int.argtypes = [u64, u64, i32, P(i32), P(i32), i64, i32, i32, P(i32)]
        L.hbr_order_by_pre.argtypes = [P(i32), i32, i32, i32, P(i32), P(i32)]
        L.hbr_pair_uniform_f32.argtypes = [u64, u64, i32, P(i32), P(i32), i64, dbl, dbl, P(ctypes.c_float)]
        L.hbr_queue_create.restype = ctypes.c_void_p
        L.hbr_queue_create.argtypes = [i32, i32]
        L.hbr_queue_destroy.argtypes = [ctypes.c_void_p]
        L.hbr_queue_prepare.restype = ctypes.c_int
        L.hbr_queue_prepare.argtypes = [ctypes.c_void_p, P(dbl), i64, P(i32), i64, dbl]
        L.hbr_queue_push.argtypes = [ctypes.c_void_p, P(i32), i32]
        L.hbr_queue_peek.restype = i64
        L.hbr_queue_peek.argtypes = [ctypes.c_void_p, P(i32), i64]
        L.hbr_queue_advance.argtypes = [ctypes.c_void_p]
        L.hbr_queue_size.restype = i64
        L.hbr_queue_size.argtypes = [ctypes.c_void_p]
        L.hbr_queue_offset.restype = i32
        L.hbr_queue_offset.argtypes = [ctypes.c_void_p]
        L.hbr_run.restype = ctypes.c_int
        L.hbr_run.argtypes = [P(Sim)]
        L.hbr_last_error.restype = ctypes.c_char_p
        for f in ("hbr_exp", "hbr_expm1", "hbr_log"):
            getattr(L, f).restype = dbl
            getattr(L, f).argtypes = [dbl]
        L.hbr_pow.restype = dbl
        L.hbr_pow.argtypes = [dbl, dbl]
        _lib = L
    return _lib


def ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(P(t))


# ---------------------------------------------------------------------------
# generator helpers (the oracle's restatement of the input generator)
S_CONN, S_V0, S_DELAY, S_W, S_INDEG, S_NORMAL = 1, 2, 3, 4, 5, 6
S_RATE, S_POISSON, S_CONN_IN, S_CONN_REC, S_W_IN, S_W_REC, S_V0_LIF, S_TARGET = 9, 10, 11, 12, 13, 14, 15, 16


def uniform(seed, stream, n, start=0):
    out = np.empty(int(n))
    lib().hbr_uniform_fill(seed, stream, start, int(n), ptr(out, dbl))
    return out


def bernoulli(seed, i0, i1, n_post, p, stream=S_CONN):
    L = lib()
    n = L.hbr_bernoulli(seed, stream, i0, i1, n_post, p, None, None, 0)
    pre = np.empty(n, np.int32)
    post = np.empty(n, np.int32)
    L.hbr_bernoulli(seed, stream, i0, i1, n_post, p, ptr(pre, i32), ptr(post, i32), n)
    return pre, post


def fixed_indegree(seed, n_pre, n_post, K, stream=S_INDEG):
    out = np.empty(n_post * K, np.int32)
    rc = lib().hbr_fixed_indegree(seed, stream, n_pre, n_post, K, ptr(out, i32))
    assert rc == 0, lib().hbr_last_error()
    return out.reshape(n_post, K)


def pair_int(seed, stream, n_post, pre, post, lo, hi):
    out = np.empty(len(pre), np.int32)
    lib().hbr_pair_int(seed, stream, n_post, ptr(pre, i32), ptr(post, i32), len(pre), lo, hi, ptr(out, i32))
    return out


def pair_uniform_f32(seed, stream, n_post, pre, post, lo, hi):
    out = np.empty(len(pre), np.float32)
    lib().hbr_pair_uniform_f32(seed, stream, n_post, ptr(pre, i32), ptr(post, i32), len(pre), lo, hi,
                               ptr(out, ctypes.c_float))
    return out


# ---------------------------------------------------------------------------
class Queue:
    """The oracle's CSpikeQueue restatement."""

    def __init__(self, start, end):
        self.L = lib()
        self.q = self.L.hbr_queue_create(start, end)

    def prepare(self, delays, sources, dt):
        delays = np.ascontiguousarray(delays, np.float64)
        sources = np.ascontiguousarray(sources, np.int32)
        rc = self.L.hbr_queue_prepare(self.q, ptr(delays, dbl), len(delays), ptr(sources, i32), len(sources), dt)
        assert rc == 0, self.L.hbr_last_error()

    def push(self, spikes):
        spikes = np.ascontiguousarray(spikes, np.int32)
        self.L.hbr_queue_push(self.q, ptr(spikes, i32), len(spikes))

    def peek(self):
        n = self.L.hbr_queue_peek(self.q, None, 0)
        out = np.empty(max(n, 1), np.int32)
        self.L.hbr_queue_peek(self.q, ptr(out, i32), n)
        return out[:n]

    def advance(self):
        self.L.hbr_queue_advance(self.q)

    def size(self):
        return int(self.L.hbr_queue_size(self.q))

    def __del__(self):
        if getattr(self, "q", None):
            self.L.hbr_queue_destroy(self.q)


# ---------------------------------------------------------------------------
class OracleGroup:
    def __init__(self, model, N, state, params, lastspike=None, not_refractory=None, gen=None, refractory=True,
                 poisson=None):
        self.model = model
        self.N = N
        self.state = np.ascontiguousarray(state, np.float64)
        self.params = np.ascontiguousarray(params, np.float64)
        self.lastspike = (np.full(N, -1e4) if lastspike is None else np.array(lastspike, np.float64))
        self.not_refractory = (np.ones(N, np.uint8) if not_refractory is None
                               else np.array(not_refractory, np.uint8))
        if not refractory:  # no lastspike / not_refractory variables (refractory=False)
            self.lastspike = self.not_refractory = None
        self.gen = gen  # (index, timebin) for spike generators
        self.gen_lastindex = 0
        self.poisson = poisson  # (seed, rates * dt) for Poisson groups

    def struct(self):
        g = Group()
        g.model, g.N = self.model, self.N
        g.state = ptr(self.state, dbl)
        g.lastspike = ptr(self.lastspike, dbl)
        g.not_refractory = ptr(self.not_refractory, u8)
        g.params = ptr(self.params, dbl)
        if self.gen is not None:
            self._gi = np.ascontiguousarray(self.gen[0], np.int32)
            self._gt = np.ascontiguousarray(self.gen[1], np.int32)
            g.n_gen = len(self._gi)
            g.gen_index = ptr(self._gi, i32)
            g.gen_timebin = ptr(self._gt, i32)
        g.gen_lastindex = self.gen_lastindex
        if self.poisson is not None:
            self._rdt = np.ascontiguousarray(self.poisson[1], np.float64)
            g.poisson_seed, g.poisson_stream = self.poisson[0], S_POISSON_GROUP
            g.poisson_rdt = ptr(self._rdt, dbl)
        return g


class OraclePathway:
    def __init__(self, source, target, pre, post, target_var, w=None, w_scalar=0.0, delay=None,
                 delay_scalar=0.0, source_range=None, count=False):
        self.source, self.target = source, target
        self.pre = np.ascontiguousarray(pre, np.int32)
        self.post = np.ascontiguousarray(post, np.int32)
        self.w = None if w is None else np.ascontiguousarray(w, np.float64)
        self.w_scalar = w_scalar
        self.delay = None if delay is None else np.ascontiguousarray(delay, np.float64)
        self.delay_scalar = delay_scalar
        self.target_var = target_var
        self.source_range = source_range
        self.counts = np.zeros(len(self.pre), np.int32) if count else None


def cpu_threads():
    """Threads for the oracle's OpenMP leg: OMP_NUM_THREADS when set, else the cores
    this process may run on."""
    n = os.environ.get("OMP_NUM_THREADS")
    return int(n) if n else len(os.sched_getaffinity(0))


class OracleQueues:
    """The pathways' spike queues kept across oracle runs (a Synapses object keeps its
    SpikeQueue from one run to the next, brian2/synapses/synapses.py:421-437)."""

    def __init__(self, n_paths):
        self.arr = (ctypes.c_void_p * max(1, n_paths))()

    def __del__(self):
        L = lib()
        for q in self.arr:
            if q:
                L.hbr_queue_destroy(q)


def run(groups, paths, dt, n_steps, start_step=0, mon_group=0, rec_group=0, rec_idx=(), rec_vars=(),
        spk_cap=None, threads=1, queues=None):
    """Runs the oracle; returns (spike_i, spike_step, rec [steps, vars, idx]).
    queues: an OracleQueues kept across calls (spikes in flight survive the end of a
    run; each call prepares the queues with its dt), else fresh queues per call."""
    L = lib()
    gs = (Group * len(groups))(*[g.struct() for g in groups])
    ps = (Pathway * max(len(paths), 1))()
    for k, p in enumerate(paths):
        s = ps[k]
        s.source_group, s.target_group = groups.index(p.source), groups.index(p.target)
        start, end = p.source_range or (0, p.source.N)
        s.source_start, s.source_end = start, end
        s.n_syn = len(p.pre)
        s.pre, s.post = ptr(p.pre, i32), ptr(p.post, i32)
        s.w = ptr(p.w, dbl)
        s.w_scalar = p.w_scalar
        s.delay = ptr(p.delay, dbl)
        s.delay_scalar = p.delay_scalar
        s.target_var = p.target_var
        s.counts = ptr(p.counts, i32)
    if spk_cap is None:
        spk_cap = max(1, sum(g.N for g in groups)) * n_steps if n_steps * groups[0].N < 5e7 else int(5e7)
    spk_i = np.empty(spk_cap, np.int32)
    spk_step = np.empty(spk_cap, np.int64)
    rec_idx = np.ascontiguousarray(rec_idx, np.int32)
    rec_vars = np.ascontiguousarray(rec_vars, np.int32)
    rec_out = np.empty((n_steps, len(rec_vars), len(rec_idx))) if len(rec_idx) else None
    sim = Sim()
    sim.dt, sim.start_step, sim.n_steps = dt, start_step, n_steps
    sim.n_groups, sim.groups = len(groups), gs
    sim.n_paths, sim.paths = len(paths), ps
    sim.mon_group = -1 if mon_group is None else mon_group
    sim.spk_i, sim.spk_step, sim.spk_cap = ptr(spk_i, i32), ptr(spk_step, i64), spk_cap
    sim.rec_group, sim.n_rec, sim.n_rec_vars = rec_group, len(rec_idx), len(rec_vars)
    sim.rec_idx, sim.rec_vars = ptr(rec_idx, i32), ptr(rec_vars, i32)
    sim.rec_out = ptr(rec_out, dbl)
    sim.threads = int(threads)
    sim.queues = ctypes.cast(queues.arr, P(ctypes.c_void_p)) if queues is not None else None
    rc = L.hbr_run(ctypes.byref(sim))
    if rc != 0:
        raise RuntimeError(L.hbr_last_error().decode())
    for k, g in enumerate(groups):
        g.gen_lastindex = gs[k].gen_lastindex
    n = sim.n_spk
    assert n <= spk_cap, "spike buffer too small"
    return spk_i[:n].copy(), spk_step[:n].copy(), rec_out


# ---------------------------------------------------------------------------
# configurations (inputs rebuilt from the generator, identical to the goldens)
def cuba_params(meta):
    return [meta["taum"], meta["taue"], meta["taui"], meta["El"], meta["Vt"], meta["Vr"], meta["refractory"]]


def build_cuba4000(meta):
    N, NE, seed = meta["N"], meta["NE"], meta["seed"]
    u = uniform(seed, S_V0, N)
    v0 = meta["Vr"] + u * (meta["Vt"] - meta["Vr"])
    state = np.stack([v0, np.zeros(N), np.zeros(N)])
    G = OracleGroup(HBR_CUBA_EXACT, N, state, cuba_params(meta))
    pre_e, post_e = bernoulli(seed, 0, NE, N, meta["p"])
    pre_i, post_i = bernoulli(seed, NE, N, N, meta["p"])
    # Brian's schedule sorts pathways by name: synapses_1_pre (Ci) before synapses_pre (Ce)
    Ci = OraclePathway(G, G, pre_i, post_i, 2, w_scalar=meta["wi"])
    Ce = OraclePathway(G, G, pre_e, post_e, 1, w_scalar=meta["we"])
    return G, [Ci, Ce]


def polar_normal(seed, n, stream=S_NORMAL):
    L = lib()
    out = np.empty(n)
    import math

    for i in range(n):
        a = 0
        while True:
            u1 = 2.0 * L.hbr_uniform(seed, stream, i * 64 + 2 * a) - 1.0
            u2 = 2.0 * L.hbr_uniform(seed, stream, i * 64 + 2 * a + 1) - 1.0
            s = u1 * u1 + u2 * u2
            a += 1
            if 0.0 < s < 1.0:
                break
        out[i] = u1 * math.sqrt((-2.0 * L.hbr_log(s)) / s)
    return out


def cobahh_params(m):
    return [m["Cm"], m["gl"], m["El"], m["EK"], m["ENa"], m["g_na"], m["g_kd"], m["VT"], m["taue"], m["taui"],
            m["Ee"], m["Ei"], m["mV"], m["ms"], m["refractory"]]


def build_cobahh(meta):
    N, NE, seed = meta["N"], meta["NE"], meta["seed"]
    z = polar_normal(seed, 3 * N)
    v0 = meta["El"] + z[:N] * 0.005
    ge0 = (z[N:2 * N] * 1.5 + 4) * 10.0 * 1e-9
    gi0 = (z[2 * N:] * 12 + 20) * 10.0 * 1e-9
    state = np.stack([v0, np.zeros(N), np.zeros(N), np.zeros(N), ge0, gi0])
    G = OracleGroup(HBR_COBAHH_EE, N, state, cobahh_params(meta))
    pre_e, post_e = bernoulli(seed, 0, NE, N, meta["p"])
    pre_i, post_i = bernoulli(seed, NE, N, N, meta["p"])
    Ce = OraclePathway(G, G, pre_e, post_e, 4, w_scalar=meta["we"], source_range=(0, NE))
    Ci = OraclePathway(G, G, pre_i, post_i, 5, w_scalar=meta["wi"], source_range=(NE, N))
    return G, [Ci, Ce]


def order_by_pre(pre_of_post):
    """Synapses of a fixed in-degree table in (source, target) order."""
    n_post, K = pre_of_post.shape
    n_pre = int(pre_of_post.max()) + 1 if pre_of_post.size else 0
    pre = np.empty(n_post * K, np.int32)
    post = np.empty(n_post * K, np.int32)
    tab = np.ascontiguousarray(pre_of_post, np.int32)
    lib().hbr_order_by_pre(ptr(tab, i32), n_pre, n_post, K, ptr(pre, i32), ptr(post, i32))
    return pre, post


def cuba_delay_synapses(meta):
    N, K, seed = meta["N"], meta["K"], meta["seed"]
    pre, post = order_by_pre(fixed_indegree(seed, N, N, K))
    d = pair_int(seed, S_DELAY, N, pre, post, meta["dmin"], meta["dmax"])
    w = pair_uniform_f32(seed, S_W, N, pre, post, 0.5, 1.5)
    exc = pre < meta["NE"]
    w32 = (w.astype(np.float64) * np.where(exc, meta["we_base"], meta["wi_base"])).astype(np.float32)
    return pre, post, d, w32, exc


def build_cuba_delay(meta):
    N, seed, dt = meta["N"], meta["seed"], meta["dt"]
    u = uniform(seed, S_V0, N)
    v0 = meta["Vr"] + u * (meta["Vt"] - meta["Vr"])
    state = np.stack([v0, np.zeros(N), np.zeros(N)])
    G = OracleGroup(HBR_CUBA_EXACT, N, state, cuba_params(meta))
    pre, post, d, w32, exc = cuba_delay_synapses(meta)
    Ce = OraclePathway(G, G, pre[exc], post[exc], 1, w=w32[exc].astype(np.float64), delay=d[exc] * dt)
    Ci = OraclePathway(G, G, pre[~exc], post[~exc], 2, w=w32[~exc].astype(np.float64), delay=d[~exc] * dt)
    return G, [Ci, Ce]


def calc_timestep(t, dt):
    """Clock._calc_timestep (brian2/core/clocks.py:350-371), epsilon_dt = 1e-4."""
    new_i = int(np.int64(np.round(t / dt)))
    if new_i * dt == t or abs(new_i * dt - t) / dt <= 1e-4:
        return new_i
    return int(np.int64(np.ceil(t / dt)))


def dt_change_runs(meta):
    """The runs of the dt-change configuration as Brian's clock takes them: per run
    (dt, first step, steps).  dt = (k * 0.1) * ms and each run's end time t_prev + dur
    * ms as the script computes them (Network.run keeps t_end, network.py:1254-1263;
    Clock.set_interval starts at calc_timestep(t, dt), clocks.py:402-429)."""
    out, t = [], 0.0
    for k, dur in meta["runs"]:
        dt = (k * 0.1) * 0.001
        t_end = t + dur * 0.001
        a, b = calc_timestep(t, dt), calc_timestep(t_end, dt)
        out.append((dt, a, b - a))
        t = t_end
    return out


def run_dt_changes(meta):
    """The oracle over the dt-change configuration's runs, queues kept across runs.
    Returns (spike indices, spike times, group)."""
    G, paths = build_cuba_delay(meta)
    qs = OracleQueues(len(paths))
    ii, tt = [], []
    for dt, a, n in dt_change_runs(meta):
        i, st, _ = run([G], paths, dt, n, start_step=a, queues=qs)
        ii.append(i)
        tt.append(st * dt)
    return np.concatenate(ii), np.concatenate(tt), G


LIF = dict(N=1000, NE=800, n_in=200, tau_m=0.02, tau_s=0.005, theta=0.01, bias=0.0105, p=0.1,
           rate_lo=5.0, rate_hi=50.0, w_scale=0.5e-3, n_steps=2000, n_readout=10)


def lif_inputs(trials, seed=1234, c=LIF):
    """The differentiable LIF configuration's inputs from the generator: input
    rates r_i = 5 + 45 U Hz; Poisson trains (input i spikes at step k of trial b
    iff U(S_POISSON, (b*n_in + i)*n_steps + k) < r_i dt); Bernoulli(0.1) input
    and recurrent connectivity (CSR order: source, then target); weights
    sigma*z (z standard normal, polar method; sigma = 0.5 mV/sqrt(expected
    fan-in)) rounded to float32; v0 = theta * U per (trial, neuron); I0 = 0."""
    dt = 1e-4
    n_in, N, T = c["n_in"], c["N"], c["n_steps"]
    rates = c["rate_lo"] + (c["rate_hi"] - c["rate_lo"]) * uniform(seed, S_RATE, n_in)
    pre_in, post_in = bernoulli(seed, 0, n_in, N, c["p"], stream=S_CONN_IN)
    pre_rec, post_rec = bernoulli(seed, 0, N, N, c["p"], stream=S_CONN_REC)
    sigma = c["w_scale"] / np.sqrt(c["p"] * (n_in + N))
    w_in = (sigma * polar_normal(seed, len(pre_in), S_W_IN)).astype(np.float32)
    w_rec = (sigma * polar_normal(seed, len(pre_rec), S_W_REC)).astype(np.float32)
    trains = []
    for b in range(trials):
        u = uniform(seed, S_POISSON, n_in * T, start=b * n_in * T).reshape(n_in, T)
        ii, kk = np.nonzero(u < (rates * dt)[:, None])
        o = np.lexsort((ii, kk))
        trains.append((ii[o].astype(np.int32), kk[o].astype(np.int32)))
    v0 = c["theta"] * uniform(seed, S_V0_LIF, trials * N).reshape(trials, N)
    targets = 0.02 + 0.16 * uniform(seed, S_TARGET, c["n_readout"])
    return dict(rates=rates, pre_in=pre_in, post_in=post_in, pre_rec=pre_rec, post_rec=post_rec, w_in=w_in,
                w_rec=w_rec, trains=trains, v0=v0, targets=targets, sigma=sigma)


def build_lif(inp, trial, c=LIF):
    """One trial of the LIF network on the oracle; pathways in Brian's order
    (syn_in_pre, then syn_rec_pre)."""
    N = c["N"]
    src = OracleGroup(HBR_SPIKEGEN, c["n_in"], np.zeros((1, c["n_in"])), [0.0], gen=inp["trains"][trial])
    G = OracleGroup(HBR_LIF_EXACT, N, np.stack([inp["v0"][trial], np.zeros(N)]),
                    [c["tau_m"], c["tau_s"], c["theta"], c["bias"]], refractory=False)
    P_in = OraclePathway(src, G, inp["pre_in"], inp["post_in"], 1, w=inp["w_in"].astype(np.float64))
    P_rec = OraclePathway(G, G, inp["pre_rec"], inp["post_rec"], 1, w=inp["w_rec"].astype(np.float64))
    return src, G, [P_in, P_rec]


def load_golden(name):
    z = np.load(os.path.join(REPO, "tests", "golden", f"golden_{name}.npz"))
    meta = ast.literal_eval(str(z["meta"]))  # the fixture's metadata is a dict literal
    return z, meta


# ---------------------------------------------------------------------------
# STDP (examples/synapses/STDP.py, after Song, Miller & Abbott 2000): Poisson
# inputs (SpikeGeneratorGroup trains from the generator) -> conductance-based
# neurons, all-to-all synapses with additive STDP (event-driven traces)
STDP = dict(n_in=1000, n_out=20, n_steps=10000, rate=15.0, taum=10e-3, taupre=20e-3, taupost=20e-3, Ee=0.0,
            vt=-54e-3, vr=-60e-3, El=-74e-3, taue=5e-3, gmax=0.01, dApre=0.01)
S_STDP_W, S_STDP_V0, S_STDP_IN = 18, 19, 20


def stdp_inputs(seed=1234, c=STDP):
    """Inputs of the STDP configuration: trains (input i spikes at step k iff
    U(S_STDP_IN, i*n_steps + k) < rate*dt, sorted by (step, index)); synapses
    all-to-all in (pre, post) order; w = gmax U(S_STDP_W, k) per synapse k;
    v0 = vr + (vt - vr) U(S_STDP_V0, j)."""
    dt = 1e-4
    n_in, n_out, T = c["n_in"], c["n_out"], c["n_steps"]
    u = uniform(seed, S_STDP_IN, n_in * T).reshape(n_in, T)
    ii, kk = np.nonzero(u < c["rate"] * dt)
    o = np.lexsort((ii, kk))
    pre = np.repeat(np.arange(n_in, dtype=np.int32), n_out)
    post = np.tile(np.arange(n_out, dtype=np.int32), n_in)
    w = c["gmax"] * uniform(seed, S_STDP_W, len(pre))
    v0 = c["vr"] + (c["vt"] - c["vr"]) * uniform(seed, S_STDP_V0, n_out)
    return dict(idx=ii[o].astype(np.int32), steps=kk[o].astype(np.int32), pre=pre, post=post, w=w, v0=v0)


# ---------------------------------------------------------------------------
# gap junctions (examples/synapses/gapjunctions.py): a summed variable
# Igap_post = w * (v_pre - v_post) over all-to-all synapses (self-pairs included)
GAP = dict(n=300, n_steps=5000, v0=1.05, tau=10e-3, w=0.02)
S_GAP_W, S_GAP_V = 21, 22


def gap_inputs(seed=1234, c=GAP):
    """Synapses in (pre, post) order; w = w (0.5 + U(S_GAP_W, k)); v = U(S_GAP_V, i)."""
    n = c["n"]
    pre = np.repeat(np.arange(n, dtype=np.int32), n)
    post = np.tile(np.arange(n, dtype=np.int32), n)
    w = c["w"] * (0.5 + uniform(seed, S_GAP_W, n * n))
    v = uniform(seed, S_GAP_V, n)
    return dict(pre=pre, post=post, w=w, v=v)
