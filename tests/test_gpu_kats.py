"""The reference's transmission known answers (tests/kats.py) on the device."""
import numpy as np
import pytest

import kats

pytestmark = pytest.mark.gpu


def onpre_spec(expr, syn_vars=(), target_dtype="float64"):
    """Brian's statements for on_pre='v += <expr>' (one target update)."""
    from brian2_amd.kernels import CodeSpec

    variables = {"v": {"kind": "ArrayVariable", "dtype": target_dtype, "owner": "target"},
                 "_postsynaptic_idx": {"kind": "DynamicArrayVariable", "dtype": "int32", "owner": "synapses"}}
    indices = {"v": "_postsynaptic_idx", "_postsynaptic_idx": "_idx"}
    for n in syn_vars:
        variables[n] = {"kind": "DynamicArrayVariable", "dtype": "float64", "owner": "synapses"}
        indices[n] = "_idx"
    return CodeSpec({"name": "synapses_pre_codeobject", "template": "synapses", "owner": "synapses",
                     "blocks": {"None": {"scalar": [], "vector": [{"var": "v", "op": "+=", "expr": expr,
                                                                   "dtype": target_dtype}],
                                         "conditional_write": {}}},
                     "variables": variables, "indices": indices})


def device_kat(case):
    from brian2_amd import Network, NeuronGroup, SpikeGeneratorGroup, StateMonitor, Synapses

    idx, steps = case["spikes"]
    src = SpikeGeneratorGroup(case["n_src"], idx, timebins=steps)
    tgt = NeuronGroup(case["n_tgt"], variables={"v": case["dtype"]})
    if case["w"] is None:
        S = Synapses(src, tgt, onpre_spec("1", target_dtype=case["dtype"]))
    else:
        S = Synapses(src, tgt, onpre_spec("w", ("w",)))
    S.connect(case["i"], case["j"])
    if case["w"] is not None:
        S.set("w", case["w"])
    d = case["delay"]
    S.set_delay(float(d) * kats.DT if np.ndim(d) == 0 else np.asarray(d, np.float64) * kats.DT)
    mon = StateMonitor(tgt, ["v"], record=np.arange(case["n_tgt"]), when="end")
    net = Network(src, tgt, S, mon, dt=kats.DT)
    net.run(case["steps"])
    return mon.values()[0]


@pytest.mark.parametrize("name", sorted(kats.CASES))
def test_device_transmission_kats(name):
    case = kats.CASES[name]
    assert np.array_equal(device_kat(case), kats.expected_matrix(case))


def test_device_transmission_kats_in_two_runs():
    """The same answers when the run is split (queue and ring persist across runs)."""
    from brian2_amd import Network, NeuronGroup, SpikeGeneratorGroup, StateMonitor, Synapses

    case = kats.CASES["all_to_one_heterogeneous_delays"]
    idx, steps = case["spikes"]
    src = SpikeGeneratorGroup(case["n_src"], idx, timebins=steps)
    tgt = NeuronGroup(1, variables={"v": "float64"})
    S = Synapses(src, tgt, onpre_spec("w", ("w",)))
    S.connect(case["i"], case["j"])
    S.set("w", case["w"])
    S.set_delay(np.asarray(case["delay"], np.float64) * kats.DT)
    mon = StateMonitor(tgt, ["v"], record=[0], when="end")
    net = Network(src, tgt, S, mon, dt=kats.DT)
    net.run(1)
    net.run(3)
    assert np.array_equal(mon.values()[0, 0], [3, 12, 33, 48])


def test_device_subgroup_propagation_kat():
    """test_subgroup.py:718-729 (test_synaptic_propagation): G1 (10 neurons, v > 1, reset
    v = 0) with the odd neurons at v = 1.1; Synapses(G1[1:6], G2[10:], on_pre='v += 1'),
    connect(i == j); after one step and the propagation, G2.v is 1 at 10, 12, 14."""
    from brian2_amd import Model, Network, NeuronGroup, Synapses
    from brian2_amd.kernels import CodeSpec

    v1 = {"kind": "ArrayVariable", "dtype": "float64", "owner": "g1"}
    th = {"name": "g1_th", "template": "threshold", "owner": "g1",
          "blocks": {"None": {"scalar": [], "conditional_write": {},
                              "vector": [{"var": "_cond", "op": ":=", "expr": "v > 1", "dtype": "bool"}]}},
          "variables": {"_cond": {"kind": "AuxiliaryVariable", "dtype": "bool"}, "v": v1},
          "indices": {"_cond": "_idx", "v": "_idx"}}
    rs = {"name": "g1_rs", "template": "reset", "owner": "g1",
          "blocks": {"None": {"scalar": [], "conditional_write": {},
                              "vector": [{"var": "v", "op": "=", "expr": "0", "dtype": "float64"}]}},
          "variables": {"v": v1}, "indices": {"v": "_idx"}}
    G1 = NeuronGroup(10, Model({"group": {"threshold": th, "reset": rs}}), name="g1", variables={"v": "float64"})
    v0 = np.zeros(10)
    v0[1::2] = 1.1  # G1.v['i%2==1'] = 1.1
    G1.v = v0
    G2 = NeuronGroup(20, variables={"v": "float64"})
    S = Synapses(G1[1:6], G2[10:], onpre_spec("1"))
    S.connect(np.arange(5), np.arange(5))  # i == j within the subgroups
    net = Network(G1, G2, S, dt=kats.DT)
    net.run(1)
    expected = np.zeros(20)
    expected[[10, 12, 14]] = 1
    assert np.array_equal(G2.get("v"), expected)
    assert isinstance(S.spec, CodeSpec)
