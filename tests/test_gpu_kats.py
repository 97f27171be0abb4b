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
