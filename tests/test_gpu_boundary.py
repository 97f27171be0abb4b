"""A Brian script through the `hip` target's boundary: the network description
brian2 handed over (tests/golden/boundary_*, written by make_boundary.py with
brian2_amd.brian_target registered) -> frontend.from_specs -> run on the GPU ->
frontend.export_state, by Brian's array names, against Brian's own cython run
of the same script (golden_*.npz), bit for bit."""
import ast
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    with open(os.path.join(GOLDEN, f"boundary_{name}.json")) as f:
        d = json.load(f)
    z = np.load(os.path.join(GOLDEN, f"boundary_{name}.npz"))
    return d, {k[len("array:"):]: z[k] for k in z.files if k.startswith("array:")}


def _golden(name):
    z = np.load(os.path.join(GOLDEN, f"golden_{name}.npz"))
    return z, ast.literal_eval(str(z["meta"]))


@pytest.mark.parametrize("pieces", [[10000], [1, 4999, 5000]], ids=["one-run", "three-runs"])
def test_cuba4000_through_the_boundary_equals_brian(pieces):
    from brian2_amd import frontend

    print(f"cuba4000 via from_specs, runs of {pieces} steps", flush=True)
    d, arrays = _fixture("cuba4000")
    net = frontend.from_specs(d["desc"], arrays, d["code_objects"])
    for k in pieces:
        net.run(k)
    out = frontend.export_state(net)
    z, meta = _golden("cuba4000")
    dt = meta["dt"]
    assert np.array_equal(out["_array_spikemonitor_i"], z["i"])
    assert np.array_equal(out["_array_spikemonitor_t"], z["step"] * dt)
    assert np.array_equal(out["_array_spikemonitor_count"], np.bincount(z["i"], minlength=4000))
    assert out["_array_spikemonitor_N"][0] == len(z["i"])
    for k, name in enumerate(("v", "ge", "gi")):
        assert np.array_equal(out[f"_array_neurongroup_{name}"], z["final"][k]), name
        # StateMonitor (when='start') rows: (step, record index)
        assert np.array_equal(out[f"_array_statemonitor_{name}"][:2000].T, z["trace"][k]), name
    assert np.array_equal(out["_array_neurongroup_lastspike"], z["lastspike"])
    assert np.array_equal(out["_array_statemonitor_t"], np.arange(10000) * dt)
    assert out["_array_defaultclock_timestep"][0] == 10000


def test_stdp_through_the_boundary_equals_brian():
    from brian2_amd import frontend

    print("stdp via from_specs", flush=True)
    d, arrays = _fixture("stdp")
    net = frontend.from_specs(d["desc"], arrays, d["code_objects"])
    net.run(d["n_steps"])
    out = frontend.export_state(net)
    z, meta = _golden("stdp")
    assert np.array_equal(out["_array_spikemonitor_i"], z["i"])
    assert np.array_equal(out["_array_spikemonitor_t"], z["step"] * meta["dt"])
    assert np.array_equal(out["_array_neurons_v"], z["final"][0])
    assert np.array_equal(out["_array_neurons_ge"], z["final"][1])
    for k, name in enumerate(("w", "Apre", "Apost", "lastupdate")):
        assert np.array_equal(out[f"_array_stdp_{name}"], z["syn"][k]), name


def test_state_set_between_runs_is_uploaded():
    """frontend.update_state (what HipDevice does before a further run): a changed
    state variable reaches the device; an unchanged description changes nothing."""
    from brian2_amd import frontend

    d, arrays = _fixture("cuba4000")
    a = frontend.from_specs(d["desc"], arrays, d["code_objects"])
    b = frontend.from_specs(d["desc"], arrays, d["code_objects"])
    a.run(500)
    b.run(500)
    sa = frontend.export_state(a)
    frontend.update_state(b, d["desc"], dict(arrays, **{k: v for k, v in sa.items() if k in arrays}))
    a.run(500)
    b.run(500)
    assert np.array_equal(frontend.export_state(a)["_array_neurongroup_v"],
                          frontend.export_state(b)["_array_neurongroup_v"])
    v = frontend.export_state(b)["_array_neurongroup_v"].copy()
    v[:100] = -0.06
    frontend.update_state(b, d["desc"], {"_array_neurongroup_v": v})
    assert np.array_equal(b.brian["groups"]["neurongroup"].get("v"), v)
