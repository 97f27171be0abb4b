"""A change of dt between runs with spikes in flight, on the device.

The reference re-bins the spikes queued in each pathway when dt changes
(CSpikeQueue::prepare, brian2/synapses/spikequeue.h:58-74) and the clock moves to
the same time at the new dt (core/clocks.py:373-429).  Checked against:

* the reference's own known answer (brian2/tests/test_synapses.py:1310-1337,
  test_changed_dt_spikes_in_queue): one spike fanned out with delays j ms into
  ten neurons that fire one step after each input, dt = 0.5, 1, 0.1 ms;
* Brian's cython runtime (golden_cuba_dt_change.npz): the CUBA-with-delays model
  in four runs at dt = 0.1, 0.05, 0.1 and 0.2 ms, spike indices and times and
  the final state bit for bit, under three launch shapes.
"""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _spec(template, owner, vector, variables, indices=None):
    from brian2_amd.kernels import CodeSpec

    return CodeSpec({"name": f"{owner}_{template}_codeobject", "template": template, "owner": owner,
                     "blocks": {"None": {"scalar": [], "vector": vector, "conditional_write": {}}},
                     "variables": variables, "indices": indices or {n: "_idx" for n in variables}})


def test_changed_dt_spikes_in_queue_kat():
    """test_synapses.py:1310-1337: G1 spikes once at t = 0; synapse j has delay j ms and
    adds 1.1 to v of neuron j, which then fires one step later (v > 1, reset v = 0)."""
    from brian2_amd import Model, Network, NeuronGroup, SpikeGeneratorGroup, SpikeMonitor, Synapses

    v = {"kind": "ArrayVariable", "dtype": "float64", "owner": "neurongroup_1"}
    th = _spec("threshold", "neurongroup_1",
               [{"var": "_cond", "op": ":=", "expr": "v > 1", "dtype": "bool", "constant": True}],
               {"_cond": {"kind": "AuxiliaryVariable", "dtype": "bool"}, "v": v})
    rs = _spec("reset", "neurongroup_1", [{"var": "v", "op": "=", "expr": "0", "dtype": "float64"}], {"v": v})
    on_pre = _spec("synapses", "synapses",
                   [{"var": "v", "op": "+=", "expr": "1.1", "dtype": "float64"}],
                   {"v": v, "_postsynaptic_idx": {"kind": "DynamicArrayVariable", "dtype": "int32",
                                                  "owner": "synapses"}},
                   {"v": "_postsynaptic_idx", "_postsynaptic_idx": "_idx"})
    G1 = SpikeGeneratorGroup(1, [0], times=[0.0], dt=0.5e-3)
    G2 = NeuronGroup(10, Model({"group": {"threshold": _raw(th), "reset": _raw(rs)}}),
                     name="neurongroup_1", variables={"v": "float64"})
    S = Synapses(G1, G2, on_pre)
    S.connect(np.zeros(10, np.int64), np.arange(10))
    S.set_delay(np.arange(10) * 1e-3)  # delay = 'j*ms'
    mon = SpikeMonitor(G2)
    net = Network(G1, G2, S, mon, dt=0.5e-3)
    net.run(10)                # 5 ms at dt = 0.5 ms
    net.run(3, dt=1e-3)        # 3 ms at dt = 1 ms
    assert net.t_step == 8
    net.run(20, dt=0.1e-3)     # 2 ms at dt = 0.1 ms
    assert net.t_step == 100
    expected = np.array([0.5, 1.5, 2.5, 3.5, 4.5, 6, 7, 8, 8.1, 9.1]) * 1e-3
    assert np.allclose(mon.t_, expected, rtol=0, atol=1e-12), mon.t_
    assert np.array_equal(mon.i, np.arange(10))


def _raw(spec):
    """A CodeSpec back to its dict form (Model takes dicts)."""
    return {"name": spec.name, "template": spec.template, "owner": spec.owner,
            "blocks": {"None": {"scalar": spec.scalar, "vector": spec.vector,
                                "conditional_write": spec.conditional_write}},
            "variables": spec.variables, "indices": spec.indices}


@pytest.mark.parametrize("shape", [{"tile": 1024}, {"fuse": False}, {"tile": 64, "gather_pairs": 32}],
                         ids=["fused-tile1024", "separate", "tile64"])
def test_cuba_dt_changes_bitexact_vs_brian(shape):
    from brian2_amd import Model, Network, NeuronGroup, SpikeMonitor, Synapses, synth

    z, meta = O.load_golden("cuba_dt_change")
    print(f"cuba_dt_change, N={meta['N']}, launch shape {shape}", flush=True)
    N, dt0 = meta["N"], meta["dt"]
    m = Model.load("cuba_synw")
    G = NeuronGroup(N, m)
    G.v = synth.cuba_v0(meta["seed"], N)
    pre, post, d, w32, exc = O.cuba_delay_synapses(meta)
    Ce = Synapses(G, G, m.pathways["exc"], name="synapses", storage={"w": np.float32}, count_events=True)
    Ci = Synapses(G, G, m.pathways["inh"], name="synapses_1", storage={"w": np.float32}, count_events=True)
    for S, sel in ((Ce, exc), (Ci, ~exc)):
        S.connect(pre[sel], post[sel])
        S.set("w", w32[sel].astype(np.float64))
        S.set_delay(d[sel] * dt0)
    mon = SpikeMonitor(G)
    net = Network(G, Ce, Ci, mon, dt=dt0, **shape)
    for dt, first, n in O.dt_change_runs(meta):
        net.run(n, dt=dt)
        assert net.t_step == first + n
    if "tile" in shape and shape["tile"] == 1024:
        assert G.fused is not None
    assert np.array_equal(mon.i, z["i"])
    assert np.array_equal(mon.t_, z["t"])
    for k, name in enumerate(("v", "ge", "gi")):
        assert np.array_equal(G.get(name), z["final"][k]), name
    assert np.array_equal(G.get("lastspike"), z["lastspike"])
    # per-synapse deliveries (carried ids included) equal the oracle's
    Go, paths = O.build_cuba_delay(meta)
    for p in paths:
        p.counts = np.zeros(len(p.pre), np.int32)
    qs = O.OracleQueues(len(paths))
    for dt, first, n in O.dt_change_runs(meta):
        O.run([Go], paths, dt, n, start_step=first, queues=qs)
    Oi, Oe = paths
    for S, sel, P in ((Ce, exc, Oe), (Ci, ~exc, Oi)):
        assert np.array_equal(S.event_counts(), P.counts), S.name
        assert S.n_events == int(P.counts.sum())
