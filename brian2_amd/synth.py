"""Synthetic inputs of the benchmark configurations, generated from (config, seed).

Every draw is a pure function of (seed, stream, counter) under SplitMix64
(csrc/hb_rng.h), so the same network can be rebuilt on any machine, on the host
(numpy, small networks) or on the device (1e9-synapse networks).

Streams: 1 Bernoulli connectivity (ctr = i*n_post + j), 2 initial v (ctr = i),
3 delays and 4 weights (keyed by the pair (i, j)), 5 fixed in-degree sampling,
6 standard normals (Marsaglia polar method).
"""
import ctypes
import math

import numpy as np

from . import _lib
from .runtime import (
    Model,
    Network,
    NeuronGroup,
    SpikeMonitor,
    StateMonitor,
    Synapses,
    dev,
    ptr,
    zeros,
)

L = _lib.lib
S_CONN, S_V0, S_DELAY, S_W, S_INDEG, S_NORMAL = 1, 2, 3, 4, 5, 6
SEED = 1234
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def rand64(seed, stream, ctr):
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) ^ _mix64(np.uint64(stream) + GAMMA))
        return _mix64(key + (np.asarray(ctr, np.uint64) + np.uint64(1)) * GAMMA)


def uniform(seed, stream, ctr):
    return (rand64(seed, stream, ctr) >> np.uint64(11)).astype(np.float64) * 2.0**-53


def uniform_int(h, n):
    return (((np.asarray(h, np.uint64) >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)).astype(np.int64)


def polar_normal(seed, n, stream=S_NORMAL):
    """Standard normals by Marsaglia's polar method: normal i uses the pairs
    (seed, stream, i*64 + 2a), (..., i*64 + 2a + 1) for attempts a = 0, 1, ...
    until 0 < s < 1 (vectorised over i; log is the shared hb_log)."""
    out = np.empty(n)
    pending = np.arange(n, dtype=np.uint64)
    a = 0
    while len(pending):
        u1 = 2.0 * uniform(seed, stream, pending * np.uint64(64) + np.uint64(2 * a)) - 1.0
        u2 = 2.0 * uniform(seed, stream, pending * np.uint64(64) + np.uint64(2 * a + 1)) - 1.0
        s = u1 * u1 + u2 * u2
        ok = (s > 0.0) & (s < 1.0)
        so = s[ok]
        lg = np.array([L.hb_host_log(float(x)) for x in so])
        out[pending[ok].astype(np.int64)] = u1[ok] * np.sqrt((-2.0 * lg) / so)
        pending = pending[~ok]
        a += 1
    return out


def bernoulli_host(seed, i0, i1, n_post, p, stream=S_CONN):
    pres, posts = [], []
    j = np.arange(n_post, dtype=np.uint64)
    for i in range(i0, i1):
        u = uniform(seed, stream, np.uint64(i) * np.uint64(n_post) + j)
        sel = np.nonzero(u < p)[0]
        pres.append(np.full(len(sel), i, np.int32))
        posts.append(sel.astype(np.int32))
    return np.concatenate(pres), np.concatenate(posts)


# ---------------------------------------------------------------------------
# device-side connectivity
# ---------------------------------------------------------------------------
def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def device_bernoulli_csr(seed, src0, src1, n_post, p):
    """CSR by source (targets ascending) of Bernoulli(p) pairs; one delay segment."""
    rows = src1 - src0
    cnt = zeros(rows + 1, "int")
    s = _stream()
    _lib.check(L.hb_gen_bernoulli_rows(seed, S_CONN, src0, src1, n_post, p, ptr(cnt), None, s), "bernoulli count")
    _lib.check(L.hb_scan_u32(ptr(cnt), rows, s), "scan")
    total = int(cnt[rows].item())
    post = zeros(total, "int")
    _lib.check(L.hb_gen_bernoulli_rows(seed, S_CONN, src0, src1, n_post, p, ptr(cnt), ptr(post), s), "bernoulli fill")
    return cnt, post, total


def device_indegree_csr(seed, n_pre, n_post, K, src0, src1, dmin, dmax, w_base, pre_of_post=None):
    """CSR by (source, delay) of a fixed in-degree network with pair delays/weights."""
    import torch

    s = _stream()
    if pre_of_post is None:
        pre_of_post = zeros(n_post * K, "int")
        status = zeros(1, "int")
        _lib.check(L.hb_gen_fixed_indegree(seed, S_INDEG, n_pre, n_post, K, ptr(pre_of_post), ptr(status), s),
                   "fixed in-degree")
        if int(status.item()):
            raise _lib.HipBrianError("fixed in-degree sampling did not converge")
    n_seg = dmax - dmin + 1
    nsegs = (src1 - src0) * n_seg
    seg_ptr = zeros(nsegs + 1, "int")
    cursor = zeros(nsegs + 1, "int")
    # upper bound of the pathway's synapses: count of pre_of_post entries in [src0, src1)
    total = int(((pre_of_post >= src0) & (pre_of_post < src1)).sum().item())
    post = zeros(total, "int")
    w = torch.zeros(max(total, 1), dtype=torch.float32, device="cuda")
    _lib.check(L.hb_csr_from_indegree(ptr(pre_of_post), n_post, K, src0, src1, seed, S_DELAY, dmin, n_seg, S_W,
                                      w_base, 0.5, 1.5, ptr(seg_ptr), ptr(cursor), ptr(post), ptr(w), s),
               "csr from in-degree")
    del cursor
    return seg_ptr, post, w, total, pre_of_post


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
CUBA = dict(taum=0.02, taue=0.005, taui=0.01, El=-0.049, Vt=-0.05, Vr=-0.06)


def cuba_v0(seed, N):
    u = uniform(seed, S_V0, np.arange(N, dtype=np.uint64))
    return CUBA["Vr"] + u * (CUBA["Vt"] - CUBA["Vr"])


def cuba_network(N=4000, NE=None, p=0.02, seed=SEED, float_t="double", we=None, wi=None, monitor=True,
                 record=None, bucket_cap=None, graph_steps=100, device_conn=True, network_options=None):
    """CUBA (examples/CUBA.py), connectivity Bernoulli(p) over all pairs, delay 0."""
    NE = N * 4 // 5 if NE is None else NE
    m = Model.load("cuba")
    ns = {}
    if we is not None:
        ns["we"] = we
    if wi is not None:
        ns["wi"] = wi
    G = NeuronGroup(N, m, float_t=float_t)
    G.v = cuba_v0(seed, N)
    Ce = Synapses(G, G, m.pathways["exc"], name="synapses")
    Ci = Synapses(G, G, m.pathways["inh"], name="synapses_1")
    for S in (Ce, Ci):
        S.namespace = dict(ns)
    if device_conn:
        for S, (a, b) in ((Ce, (0, NE)), (Ci, (NE, N))):
            cnt, post, total = device_bernoulli_csr(seed, a, b, N, p)
            # the CSR covers the whole source group: rows outside [a, b) are empty
            seg = zeros(N + 1, "int")
            seg[a:b + 1] = cnt[: b - a + 1]
            seg[b + 1:] = total
            S.connect_device_csr(seg, post, {}, 0, 1, total)
    else:
        for S, (a, b) in ((Ce, (0, NE)), (Ci, (NE, N))):
            pre, post = bernoulli_host(seed, a, b, N, p)
            S.connect(pre, post)
    objs = [G, Ce, Ci]
    mon = SpikeMonitor(G) if monitor else None
    if mon is not None:
        objs.append(mon)
    st = None
    if record is not None:
        st = StateMonitor(G, ["v", "ge", "gi"], record=record)
        objs.append(st)
    net = Network(*objs, dt=1e-4, **dict(dict(bucket_cap=bucket_cap, graph_steps=graph_steps), **(network_options or {})))
    return dict(net=net, G=G, Ce=Ce, Ci=Ci, mon=mon, state=st)


def cuba_synw_network(N=5000, K=1000, dmin=1, dmax=50, seed=SEED, ginh=0.4, float_t="double", monitor=True,
                      record=None, bucket_cap=None, graph_steps=100, network_options=None, count_events=False):
    """CUBA with fixed in-degree K, per-synapse fp32 weights and delays (the 1M-neuron workload).
    count_events: per-synapse delivered-event counters on both pathways."""
    NE = N * 4 // 5
    m = Model.load("cuba_synw")
    G = NeuronGroup(N, m, float_t=float_t)
    G.v = cuba_v0(seed, N)
    we_base = 1.62e-3 * (80.0 / K)
    wi_base = -9e-3 * (80.0 / K) * ginh
    Ce = Synapses(G[:NE], G, m.pathways["exc"], name="synapses", storage={"w": np.float32},
                  count_events=count_events)
    Ci = Synapses(G[NE:], G, m.pathways["inh"], name="synapses_1", storage={"w": np.float32},
                  count_events=count_events)
    pop = None
    n_syn = 0
    for S, (a, b), wb in ((Ce, (0, NE), we_base), (Ci, (NE, N), wi_base)):
        seg, post, w, total, pop = device_indegree_csr(seed, N, N, K, a, b, dmin, dmax, wb, pre_of_post=pop)
        S.connect_device_csr(seg, post, {"w": w}, dmin, dmax - dmin + 1, total)
        n_syn += total
    del pop
    objs = [G, Ce, Ci]
    mon = SpikeMonitor(G) if monitor else None
    if mon is not None:
        objs.append(mon)
    st = None
    if record is not None:
        st = StateMonitor(G, ["v", "ge", "gi"], record=record)
        objs.append(st)
    net = Network(*objs, dt=1e-4, **dict(dict(bucket_cap=bucket_cap, graph_steps=graph_steps), **(network_options or {})))
    return dict(net=net, G=G, Ce=Ce, Ci=Ci, mon=mon, state=st, n_syn=n_syn, we_base=we_base, wi_base=wi_base)


COBAHH = dict(El=-0.06, taue=0.005, taui=0.01)


def cobahh_network(N=4000, p=0.02, seed=SEED, float_t="double", monitor=True, record=None, bucket_cap=None,
                   graph_steps=100, network_options=None):
    """COBAHH (examples/COBAHH.py): Pe = P[:0.8N] -> ge += we, Pi = P[0.8N:] -> gi += wi."""
    NE = N * 4 // 5
    m = Model.load("cobahh")
    G = NeuronGroup(N, m, float_t=float_t)
    z = polar_normal(seed, 3 * N)
    G.v = COBAHH["El"] + z[:N] * 0.005
    G.ge = (z[N:2 * N] * 1.5 + 4) * 10.0 * 1e-9
    G.gi = (z[2 * N:] * 12 + 20) * 10.0 * 1e-9
    Ce = Synapses(G[:NE], G, m.pathways["exc"], name="synapses")
    Ci = Synapses(G[NE:], G, m.pathways["inh"], name="synapses_1")
    for S, (a, b) in ((Ce, (0, NE)), (Ci, (NE, N))):
        cnt, post, total = device_bernoulli_csr(seed, a, b, N, p)
        S.connect_device_csr(cnt, post, {}, 0, 1, total)
    objs = [G, Ce, Ci]
    mon = SpikeMonitor(G) if monitor else None
    if mon is not None:
        objs.append(mon)
    st = None
    if record is not None:
        st = StateMonitor(G, ["v", "m", "n", "h", "ge", "gi"], record=record)
        objs.append(st)
    net = Network(*objs, dt=1e-4, **dict(dict(bucket_cap=bucket_cap, graph_steps=graph_steps), **(network_options or {})))
    return dict(net=net, G=G, Ce=Ce, Ci=Ci, mon=mon, state=st)


# ---------------------------------------------------------------------------
# batched sweep: 1024 independent CUBA networks of uneven size
# ---------------------------------------------------------------------------
S_SWEEP_N, S_SWEEP_S = 7, 8


def sweep_networks(n_nets=1024, seed=SEED, nmin=500, nmax=8000, smin=0.5, smax=2.0, model="cuba"):
    """N_k = nmin + U{0..nmax-nmin}, 80/20 split, we_k = we * s_k with s_k
    log-uniform in [smin, smax], connectivity seed seed + k."""
    from .batched import SweepNetwork

    m = Model.load(model)
    we = m.pathways["exc"].variables["we"]["value"]
    lo, hi = L.hb_host_log(smin), L.hb_host_log(smax)
    nets = []
    for k in range(n_nets):
        N = nmin + int(uniform_int(rand64(seed, S_SWEEP_N, k), nmax - nmin + 1))
        u = float(uniform(seed, S_SWEEP_S, k))
        s = L.hb_host_exp(lo + u * (hi - lo))
        nets.append(SweepNetwork(N, N * 4 // 5, seed + k, {"we": we * s}))
    return nets
