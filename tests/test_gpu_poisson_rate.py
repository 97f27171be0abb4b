"""PoissonGroup and PopulationRateMonitor on the device.

PoissonGroup draws come from the counter-based generator, so the device must
reproduce the oracle's spike trains exactly (the oracle's draws are checked
statistically against rates * dt in test_oracle.py; Brian itself draws with numpy,
so agreement with Brian is statistical only).  Population rates are checked
against the spike lists of Brian's cython runtime (golden CUBA 4000) and the
oracle, count for count, and against ratemonitor.pyx's rate expression.
"""
import numpy as np
import pytest

import oracle_py as O
from test_gpu_kats import onpre_spec

pytestmark = pytest.mark.gpu

DT = 1e-4


def _rates(n, seed=5):
    return 5.0 + 195.0 * O.uniform(seed, 3, n)  # 5 .. 200 Hz


def _oracle_poisson(N, rates, seed, n_steps, start_step=0):
    g = O.OracleGroup(O.HBR_POISSON, N, np.zeros(N), [], refractory=False, poisson=(seed, rates * DT))
    return O.run([g], [], DT, n_steps, start_step=start_step, mon_group=0)[:2]


def test_poisson_spikes_equal_oracle():
    from brian2_amd import Network, PoissonGroup, PopulationRateMonitor, SpikeMonitor

    N, steps, seed = 3000, 2000, 77
    rates = _rates(N)
    P = PoissonGroup(N, rates, seed=seed)
    mon = SpikeMonitor(P)
    rm = PopulationRateMonitor(P)
    net = Network(P, mon, rm, dt=DT)
    net.run(steps)
    i_o, s_o = _oracle_poisson(N, rates, seed, steps)
    assert len(i_o) > 1000
    assert np.array_equal(mon.i, i_o) and np.array_equal(mon.step, s_o)
    assert np.array_equal(mon.count, np.bincount(i_o, minlength=N))
    assert np.array_equal(rm.count, np.bincount(s_o, minlength=steps))
    assert np.array_equal(rm.step, np.arange(steps))
    assert np.array_equal(rm.rate(DT), np.bincount(s_o, minlength=steps).astype(np.float64) / DT / N)


def test_poisson_subgroup_monitors_and_pieces():
    """Monitors of a subgroup; the same trains when the run is split and after store/restore."""
    from brian2_amd import Network, PoissonGroup, PopulationRateMonitor, SpikeMonitor

    N, seed = 1000, 3
    rates = _rates(N, seed=9)
    P = PoissonGroup(N, rates, seed=seed)
    mon = SpikeMonitor(P[100:400])
    rm = PopulationRateMonitor(P[250:900])
    net = Network(P, mon, rm, dt=DT)
    net.run(300)
    net.store()
    net.run(450)
    i_a, s_a, c_a = mon.i.copy(), mon.step.copy(), rm.count.copy()
    net.restore()
    for k in (1, 199, 250):
        net.run(k)
    assert np.array_equal(mon.i, i_a) and np.array_equal(mon.step, s_a) and np.array_equal(rm.count, c_a)
    i_o, s_o = _oracle_poisson(N, rates, seed, 750)
    sel = (i_o >= 100) & (i_o < 400)
    assert np.array_equal(mon.i, i_o[sel]) and np.array_equal(mon.step, s_o[sel])
    sel = (i_o >= 250) & (i_o < 900)
    assert np.array_equal(rm.count, np.bincount(s_o[sel], minlength=750))


def test_poisson_drives_synapses_like_the_oracle():
    """PoissonGroup -> Synapses (weights, delays 0..4 steps) -> 'v : 1' targets, bit-exact."""
    from brian2_amd import Network, NeuronGroup, PoissonGroup, StateMonitor, Synapses

    n_src, n_tgt, steps, seed = 2000, 300, 600, 11
    rates = _rates(n_src, seed=21)
    rng = np.random.default_rng(4)
    n_syn = 40000
    pre = rng.integers(0, n_src, n_syn).astype(np.int32)
    post = rng.integers(0, n_tgt, n_syn).astype(np.int32)
    w = rng.normal(0.0, 1.0, n_syn)
    d = rng.integers(0, 5, n_syn).astype(np.float64) * DT
    P = PoissonGroup(n_src, rates, seed=seed)
    T = NeuronGroup(n_tgt, variables={"v": "float64"})
    S = Synapses(P, T, onpre_spec("w", ("w",)))
    S.connect(pre, post)
    S.set("w", w)
    S.set_delay(d)
    st = StateMonitor(T, ["v"], record=np.arange(n_tgt), when="end")
    Network(P, T, S, st, dt=DT).run(steps)
    src = O.OracleGroup(O.HBR_POISSON, n_src, np.zeros(n_src), [], refractory=False, poisson=(seed, rates * DT))
    tgt = O.OracleGroup(O.HBR_PASSIVE, n_tgt, np.zeros(n_tgt), [], refractory=False)
    path = O.OraclePathway(src, tgt, pre, post, 0, w=w, delay=d)
    O.run([src, tgt], [path], DT, steps, mon_group=None)
    assert np.array_equal(st.values()[0][:, -1], tgt.state)


def test_rate_monitor_cuba4000_equals_brian_spikes():
    """PopulationRateMonitor of CUBA 4000 (whole group and a subgroup) against the spikes
    Brian's cython runtime recorded (1 s)."""
    from brian2_amd import PopulationRateMonitor, synth

    z, meta = O.load_golden("cuba4000")
    c = synth.cuba_network(N=meta["N"], NE=meta["NE"], p=meta["p"], seed=meta["seed"])
    G = c["G"]
    # the network accepts one rate monitor per group: run twice, whole group then a subgroup
    rm = PopulationRateMonitor(G)
    c["net"].add(rm)
    c["net"].run(10000)
    cnt = np.bincount(z["step"], minlength=10000)
    assert np.array_equal(rm.count, cnt)
    assert np.array_equal(rm.rate(meta["dt"]), cnt.astype(np.float64) / meta["dt"] / meta["N"])
    c2 = synth.cuba_network(N=meta["N"], NE=meta["NE"], p=meta["p"], seed=meta["seed"])
    rs = PopulationRateMonitor(c2["G"][1000:3500])
    c2["net"].add(rs)
    c2["net"].run(4000)
    sel = (z["i"] >= 1000) & (z["i"] < 3500) & (z["step"] < 4000)
    assert np.array_equal(rs.count, np.bincount(z["step"][sel], minlength=4000))
    sm = rs.smooth_rate(meta["dt"], "flat", width=5 * meta["dt"])
    raw = rs.rate(meta["dt"])
    assert np.allclose(sm, np.convolve(raw, np.ones(5) / 5, mode="same"), rtol=0, atol=1e-9)


def test_rate_monitor_of_a_spike_generator():
    from brian2_amd import Network, PopulationRateMonitor, SpikeGeneratorGroup

    idx = np.array([0, 3, 3, 7, 1, 2, 9], np.int32)
    tb = np.array([0, 0, 5, 5, 5, 8, 12], np.int32)
    G = SpikeGeneratorGroup(10, idx, timebins=tb)
    rm = PopulationRateMonitor(G[1:8])
    net = Network(G, rm, dt=DT)
    net.run(6)
    net.run(10)
    expect = np.zeros(16, np.int64)
    for i, t in zip(idx, tb):
        if 1 <= i < 8:
            expect[t] += 1
    assert np.array_equal(rm.count, expect) and np.array_equal(rm.step, np.arange(16))


@pytest.mark.parametrize("shape", [{}, {"tile": 1024}, {"tile": 1024, "fuse": False}],
                         ids=["default", "tile1024-fused", "tile1024-separate"])
def test_poisson_two_groups_three_pathways_like_the_oracle(shape):
    """PoissonGroup -> two pathways (delays 1-4 and 2-7 steps) into one group and a third
    (1-3) into another: with tiles of 1024 each group and the gathers into it run as one
    launch per step (two of them, each reading the source's spike ring); final state
    bit-exact against the oracle's sequential fold in pathway order."""
    from brian2_amd import Network, NeuronGroup, PoissonGroup, StateMonitor, Synapses

    n_src, n1, n2, steps, seed = 3000, 300, 200, 500, 13
    rates = _rates(n_src, seed=31)
    rng = np.random.default_rng(8)
    P = PoissonGroup(n_src, rates, seed=seed)
    T1 = NeuronGroup(n1, variables={"v": "float64"}, name="t1")
    T2 = NeuronGroup(n2, variables={"v": "float64"}, name="t2")
    src = O.OracleGroup(O.HBR_POISSON, n_src, np.zeros(n_src), [], refractory=False, poisson=(seed, rates * DT))
    o1 = O.OracleGroup(O.HBR_PASSIVE, n1, np.zeros(n1), [], refractory=False)
    o2 = O.OracleGroup(O.HBR_PASSIVE, n2, np.zeros(n2), [], refractory=False)
    objs, opaths = [P, T1, T2], []
    for name, tgt, otgt, n_syn, d0, d1 in (("a", T1, o1, 30000, 1, 4), ("b", T1, o1, 20000, 2, 7),
                                           ("c", T2, o2, 25000, 1, 3)):
        pre = rng.integers(0, n_src, n_syn).astype(np.int32)
        post = rng.integers(0, tgt.N, n_syn).astype(np.int32)
        w = rng.normal(0.0, 1.0, n_syn)
        d = rng.integers(d0, d1 + 1, n_syn).astype(np.float64) * DT
        S = Synapses(P, tgt, onpre_spec("w", ("w",)), name=name)
        S.connect(pre, post)
        S.set("w", w)
        S.set_delay(d)
        objs.append(S)
        opaths.append(O.OraclePathway(src, otgt, pre, post, 0, w=w, delay=d))
    st1 = StateMonitor(T1, ["v"], record=np.arange(n1), when="end")
    st2 = StateMonitor(T2, ["v"], record=np.arange(n2), when="end")
    print(f"poisson -> 2 groups, 3 pathways, launch shape {shape}", flush=True)
    net = Network(*objs, st1, st2, dt=DT, **shape)
    net.run(steps)
    if shape.get("tile") == 1024 and shape.get("fuse", True):
        assert T1.fused is not None and T2.fused is not None
    O.run([src, o1, o2], opaths, DT, steps, mon_group=None)
    assert np.array_equal(st1.values()[0][:, -1], o1.state)
    assert np.array_equal(st2.values()[0][:, -1], o2.state)
