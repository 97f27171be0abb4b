"""The 1M-neuron configuration's model (CUBA, fixed in-degree 1000, per-synapse fp32
weights, delays 1-50 steps) at sizes where only the device's large-network code
paths run, against the oracle.

* N = 100,000 (1e8 synapses), 1000 steps: spikes and final state bit-exact with
  the oracle, in the 1M network's kernel shape (tiles of 1024 neurons, the group's
  tiles and both gathers fused into one launch, two workgroups per CU) and with
  separate launches over 1563 tiles of 64 neurons (gathers over more tiles than
  a workgroup has threads, multi-pass folds, per-target index lists).
* N = 1,000,000 (1e9 synapses), 300 steps, checked through properties that do not
  need the oracle to run the whole network:
  - every step's delivered events equal the sum, over the (spike, delay) pairs
    due that step, of the CSR segment lengths (computed on the host from the
    recorded spikes and the CSR offsets);
  - 1000 sampled targets, re-simulated by the oracle from the device's recorded
    spikes (a spike generator replaying them into the targets' own synapses,
    read back from the device CSR), end with bit-identical state and spikes.
"""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _meta(N, n_steps):
    z, meta = O.load_golden("cuba_delay")  # the configuration's constants (fixture of N = 5000)
    m = dict(meta)
    m.update(N=N, NE=N * 4 // 5, n_steps=n_steps)
    return m


@pytest.fixture(scope="module")
def oracle_100k():
    meta = _meta(100_000, 1000)
    print(f"oracle: N={meta['N']}, K={meta['K']}, {meta['n_steps']} steps", flush=True)
    G, paths = O.build_cuba_delay(meta)
    i, s, _ = O.run([G], paths, meta["dt"], meta["n_steps"], threads=O.cpu_threads(), spk_cap=20_000_000)
    return meta, i, s, G.state.copy(), G.lastspike.copy()


@pytest.mark.parametrize("shape", [{}, {"tile": 64, "fuse": False, "gather_pairs": 128}],
                         ids=["fused-tile1024", "separate-tile64"])
def test_cuba_100k_bitexact_vs_oracle(oracle_100k, shape):
    from brian2_amd import synth

    meta, i, s, state, lastspike = oracle_100k
    print(f"device: N={meta['N']}, launch shape {shape}", flush=True)
    c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                ginh=meta["ginh"], network_options=shape)
    c["net"].run(meta["n_steps"])
    G = c["G"]
    if shape:
        assert G.fused is None and G.n_tiles > 1024
    else:
        assert G.fused is not None and G.tile == 1024 and G.waves == 8
    assert len(i) > 10_000
    assert np.array_equal(c["mon"].i, i) and np.array_equal(c["mon"].step, s)
    for k, name in enumerate(("v", "ge", "gi")):
        assert np.array_equal(G.get(name), state[k]), name
    assert np.array_equal(G.get("lastspike"), lastspike)


def _csr_host(S):
    seg = S.csr["seg_ptr"].cpu().numpy().view(np.uint32).astype(np.int64)
    return seg


def _incoming(S, targets, N):
    """The synapses of `targets` (ascending) in S's device CSR: source, local target
    index, delay in steps, fp32 weight."""
    import torch

    n = S.csr["n_syn"]
    post = S.csr["post"][:n]
    mark = torch.zeros(N, dtype=torch.bool, device="cuda")
    mark[torch.from_numpy(targets).long().cuda()] = True
    chunk = 1 << 27
    pos = torch.cat([torch.nonzero(mark[post[a:a + chunk].long()]).flatten() + a for a in range(0, n, chunk)])
    seg = S.csr["seg_ptr"].long() & 0xFFFFFFFF
    row = torch.searchsorted(seg, pos, right=True) - 1
    n_seg = S.csr["n_seg"]
    src = (S.src_range[0] + row // n_seg).cpu().numpy().astype(np.int32)
    d = (S.csr["dmin"] + row % n_seg).cpu().numpy().astype(np.int64)
    w = S.csr["syn"]["w"][pos].cpu().numpy().astype(np.float64)
    p = np.searchsorted(targets, post[pos].cpu().numpy()).astype(np.int32)
    o = np.lexsort((p, src))  # (source, target): the order Synapses.connect(i=, j=) receives them in
    return src[o], p[o], d[o], w[o]


def test_cuba_1m_properties():
    import torch

    from brian2_amd import synth

    meta = _meta(1_000_000, 300)
    N, dt = meta["N"], meta["dt"]
    print("device: N=1e6, K=1000 (1e9 synapses), 300 steps: 150 one at a time, then 150 from graphs", flush=True)
    c = synth.cuba_synw_network(N=N, K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                ginh=meta["ginh"])
    net, G = c["net"], c["G"]
    paths = {"synapses": c["Ce"], "synapses_1": c["Ci"]}
    per_step = []
    for _ in range(150):
        e0 = net.n_events
        net.run(1)
        per_step.append(net.n_events - e0)
    e0 = net.n_events
    net.run(150)
    rest = net.n_events - e0
    assert G.fused is not None and G.tile == 1024
    i, s = c["mon"].i, c["mon"].step
    assert len(i) > 1000
    # (1) events per step from the recorded spikes and the CSR offsets
    expect = np.zeros(300 + meta["dmax"] + 1, np.int64)
    for S in paths.values():
        seg = _csr_host(S)
        n_seg = S.csr["n_seg"]
        a, b = S.src_range
        sel = (i >= a) & (i < b)
        rows = (i[sel].astype(np.int64) - a)[:, None] * n_seg + np.arange(n_seg)[None, :]
        lens = seg[rows + 1] - seg[rows]
        due = s[sel][:, None] + (S.csr["dmin"] + np.arange(n_seg))[None, :]
        np.add.at(expect, due.ravel(), lens.ravel())
    assert np.array_equal(np.array(per_step), expect[:150])
    assert rest == int(expect[150:300].sum())
    assert int(expect[:300].sum()) > 10_000_000
    # (2) sampled targets re-simulated by the oracle from the recorded spikes
    rng = np.random.default_rng(7)
    targets = np.sort(rng.choice(N, 1000, replace=False)).astype(np.int32)
    src_g = O.OracleGroup(O.HBR_SPIKEGEN, N, np.zeros((1, N)), [0.0], gen=(i.astype(np.int32), s.astype(np.int32)))
    v0 = synth.cuba_v0(meta["seed"], N)[targets]
    tg = O.OracleGroup(O.HBR_CUBA_EXACT, len(targets), np.stack([v0, np.zeros(len(targets)), np.zeros(len(targets))]),
                       O.cuba_params(meta))
    opaths = []
    for name, var, rng_ in (("synapses_1", 2, (meta["NE"], N)), ("synapses", 1, (0, meta["NE"]))):
        src, p, d, w = _incoming(paths[name], targets, N)
        opaths.append(O.OraclePathway(src_g, tg, src, p, var, w=w, delay=d * dt, source_range=rng_))
    ti, ts, _ = O.run([src_g, tg], opaths, dt, 300, mon_group=1, spk_cap=1_000_000)
    for k, name in enumerate(("v", "ge", "gi")):
        assert np.array_equal(G.get(name)[targets], tg.state[k]), name
    mine = np.isin(i, targets)
    assert np.array_equal(np.searchsorted(targets, i[mine]), ti) and np.array_equal(s[mine], ts)
    del c, net, G
    torch.cuda.empty_cache()
