"""The `hip` target's Brian side (brian2_amd.brian_target), checked against what
brian2 itself produced when tests/golden/make_boundary.py ran it (no brian2
import here):

* the statements the hip code generator recorded are the ones Brian's cython
  target emitted for the same models (spec_cuba.json, spec_stdp.json);
* the network description handed to frontend.from_specs names Brian's arrays
  (_array_<owner>_<var>, devices/device.py:500-515) and holds Brian's inputs
  (the oracle generator's connectivity, the golden's initial state);
* the copy back into Brian: after finish_run, Brian's own accessors (G.v_,
  mon.i, mon.t_, mon.count, st.v_, S.w ...) returned exactly the values
  handed over, with monitors resized to the run's records.
"""
import ast
import json
import os

import numpy as np

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    with open(os.path.join(GOLDEN, f"boundary_{name}.json")) as f:
        d = json.load(f)
    z = np.load(os.path.join(GOLDEN, f"boundary_{name}.npz"))
    arrays = {k[len("array:"):]: z[k] for k in z.files if k.startswith("array:")}
    readback = {k[len("readback_"):]: z[k] for k in z.files if k.startswith("readback_")}
    return d, arrays, readback


def _golden(name):
    z = np.load(os.path.join(GOLDEN, f"golden_{name}.npz"))
    return z, ast.literal_eval(str(z["meta"]))


def _spec(name):
    with open(os.path.join(GOLDEN, f"spec_{name}.json")) as f:
        return {c["name"]: c for c in json.load(f)}


def _blocks(c):
    return {b: (v["scalar"], v["vector"], v["conditional_write"]) for b, v in c["blocks"].items()}


def test_recorded_statements_equal_the_cython_targets():
    """Per template, the same statement blocks (names differ: the cython capture
    ran the model in a second scope, e.g. neurongroup_1)."""
    for fixture, spec in (("cuba4000", "cuba"), ("stdp", "stdp")):
        d, _, _ = _fixture(fixture)
        want = _spec(spec)
        for template in ("stateupdate", "threshold", "reset", "synapses"):
            got = sorted(json.dumps(_blocks(c), sort_keys=True) for c in d["code_objects"]
                         if c["template"] == template)
            ref = sorted(json.dumps(_blocks(c), sort_keys=True) for c in want.values() if c["template"] == template)
            assert got and got == ref, (fixture, template)


def test_description_of_cuba4000():
    d, arrays, _ = _fixture("cuba4000")
    desc = d["desc"]
    z, meta = _golden("cuba4000")
    assert desc["dt"] == meta["dt"] and desc["t_step"] == 0 and d["n_steps"] == 10000
    (g,) = desc["groups"]
    assert (g["kind"], g["N"]) == ("NeuronGroup", 4000)
    assert set(g["variables"]) == {"v", "ge", "gi", "lastspike", "not_refractory"}
    G, paths = O.build_cuba4000(meta)
    assert np.array_equal(arrays["_array_neurongroup_v"], G.state[0])
    assert np.all(arrays["_array_neurongroup_ge"] == 0) and np.all(arrays["_array_neurongroup_lastspike"] == -1e4)
    syn = {s["name"]: s for s in desc["synapses"]}
    assert set(syn) == {"synapses", "synapses_1"}
    for s in syn.values():
        assert (s["source"], s["target"], s["source_start"], s["source_stop"]) == ("neurongroup", "neurongroup", 0, 4000)
        (p,) = s["pathways"]
        assert p["prepost"] == "pre" and p["order"] == -1
        assert np.all(arrays[p["delay"]] == 0)
    # the connectivity Brian holds is the oracle generator's, in synapse order
    pre_e, post_e = O.bernoulli(meta["seed"], 0, meta["NE"], meta["N"], meta["p"])
    assert np.array_equal(arrays["_array_synapses__synaptic_pre"], pre_e)
    assert np.array_equal(arrays["_array_synapses__synaptic_post"], post_e)
    mons = {m["kind"]: m for m in desc["monitors"]}
    assert mons["StateMonitor"]["variables"] == ["v", "ge", "gi"] and mons["StateMonitor"]["when"] == "start"
    assert np.array_equal(arrays[mons["StateMonitor"]["record"]], [0, 1, 2, 3])
    assert (mons["SpikeMonitor"]["start"], mons["SpikeMonitor"]["stop"]) == (0, 4000)


def test_description_of_stdp():
    d, arrays, _ = _fixture("stdp")
    desc = d["desc"]
    kinds = {g["name"]: g["kind"] for g in desc["groups"]}
    assert kinds == {"neurons": "NeuronGroup", "inputs": "SpikeGeneratorGroup"}
    (s,) = desc["synapses"]
    assert [p["prepost"] for p in s["pathways"]] == ["pre", "post"]
    assert [p["order"] for p in s["pathways"]] == [-1, 1]
    assert set(s["variables"]) == {"w", "Apre", "Apost", "lastupdate"}
    inp = O.stdp_inputs(1234)
    assert np.array_equal(arrays["_array_stdp__synaptic_pre"], inp["pre"])
    assert np.array_equal(arrays["_array_stdp__synaptic_post"], inp["post"])
    assert np.array_equal(arrays["_array_stdp_w"], inp["w"])
    assert np.array_equal(arrays["_array_inputs_neuron_index"], inp["idx"])


def test_copy_back_into_brian_cuba4000():
    _, _, rb = _fixture("cuba4000")
    z, meta = _golden("cuba4000")
    dt, n = meta["dt"], 10000
    for k, name in enumerate(("v", "ge", "gi")):
        assert np.array_equal(rb[name], z["final"][k])
    assert np.array_equal(rb["lastspike"], z["lastspike"])
    assert np.array_equal(rb["not_refractory"], np.arange(4000) % 3 != 0)
    assert np.array_equal(rb["mon_i"], z["i"]) and np.array_equal(rb["mon_t"], z["step"] * dt)
    assert np.array_equal(rb["mon_count"], np.bincount(z["i"], minlength=4000))
    assert rb["mon_num_spikes"][0] == len(z["i"])
    trace = (np.arange(3 * n * 4, dtype=np.float64).reshape(3, n, 4) * 0.25 - 7.0) * 1e-3
    # Brian's StateMonitor accessors give (record index, step)
    for k, name in enumerate(("st_v", "st_ge", "st_gi")):
        assert np.array_equal(rb[name], trace[k].T)
    assert np.array_equal(rb["st_t"], np.arange(n) * dt)
    assert rb["t"][0] == n * dt


def test_copy_back_into_brian_stdp():
    _, _, rb = _fixture("stdp")
    z, meta = _golden("stdp")
    assert np.array_equal(rb["v"], z["final"][0]) and np.array_equal(rb["ge"], z["final"][1])
    for k, name in enumerate(("w", "Apre", "Apost", "lastupdate")):
        assert np.array_equal(rb[name], z["syn"][k]), name
    assert np.array_equal(rb["mon_i"], z["i"]) and np.array_equal(rb["mon_t"], z["step"] * meta["dt"])
