"""Writes the boundary fixtures: what the `hip` target's Brian side hands to
brian2_amd.frontend.from_specs for a Brian script, captured from brian2 itself
(this container only; nothing here travels).

    make -C oracle && make -C oracle ref
    python tests/golden/make_boundary.py

For each network (the inputs of golden_cuba4000.npz and golden_stdp.npz, built
with Brian's own objects exactly as make_golden.py builds them), with the hip
target registered (brian2_amd.brian_target.register) and selected
(set_device("hip"), prefs.codegen.target = "hip"):

  1. brian_target.prepare_run: Network.run's set-up (clock interval,
     before_run, which creates the hip code objects), then describe() ->
     boundary_<name>.json (desc + the recorded code objects) and
     boundary_<name>.npz (the arrays, by Brian's array names).  The GPU test
     (tests/test_gpu_boundary.py) feeds these to from_specs and compares the run
     with Brian's own cython run (the golden).
  2. brian_target.finish_run with a state shaped like export_state's (the
     golden's final state and spikes, a recorded trace): Brian's own accessors
     (G.v_, mon.i, mon.t_, mon.count, st.v_, S.w_ ...) then read the values back;
     they are stored as readback_* in the npz, and the CPU test checks that they
     are the values handed over (the copy-back mapping and resizing).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import _brian_env as be  # noqa: E402
import make_golden as mg  # noqa: E402

sys.path.insert(0, be.REPO)


def _setup():
    b2 = be.setup_brian("cython")
    from brian2 import prefs, set_device

    import brian2_amd.brian_target as bt

    bt.register()
    set_device("hip")
    prefs.codegen.target = "hip"
    return b2, bt


def _save(name, desc, arrays, code_objects, readback, n_steps):
    with open(os.path.join(HERE, f"boundary_{name}.json"), "w") as f:
        json.dump({"desc": desc, "code_objects": code_objects, "n_steps": n_steps}, f, indent=1, sort_keys=True)
        f.write("\n")
    out = {f"array:{k}": np.asarray(v) for k, v in arrays.items()}
    out.update({f"readback_{k}": np.asarray(v) for k, v in readback.items()})
    np.savez_compressed(os.path.join(HERE, f"boundary_{name}.npz"), **out)
    print(f"boundary_{name}: {len(code_objects)} code objects, {len(arrays)} arrays, {len(readback)} readbacks")


def cuba4000(bt, lib):
    from brian2 import NeuronGroup, Network, SpikeMonitor, StateMonitor, Synapses, defaultclock, get_device, ms, mV
    from brian2 import start_scope, volt

    start_scope()
    N, NE = 4000, 3200
    taum, taue, taui = 20 * ms, 5 * ms, 10 * ms
    Vt, Vr, El = -50 * mV, -60 * mV, -49 * mV
    we = (60 * 0.27 / 10) * mV
    wi = (-20 * 4.5 / 10) * mV
    defaultclock.dt = 0.1 * ms
    eqs = """
    dv/dt  = (ge+gi-(v-El))/taum : volt (unless refractory)
    dge/dt = -ge/taue : volt
    dgi/dt = -gi/taui : volt
    """
    P = NeuronGroup(N, eqs, threshold="v>Vt", reset="v = Vr", refractory=5 * ms, method="exact")
    u = mg.uniform(lib, mg.SEED, mg.S_V0, N)
    P.v = (float(Vr) + u * (float(Vt) - float(Vr))) * volt
    P.ge = 0 * mV
    P.gi = 0 * mV
    pre_e, post_e = mg.bernoulli(lib, mg.SEED, 0, NE, N, 0.02)
    pre_i, post_i = mg.bernoulli(lib, mg.SEED, NE, N, N, 0.02)
    Ce = Synapses(P, P, on_pre="ge += we")
    Ci = Synapses(P, P, on_pre="gi += wi")
    Ce.connect(i=pre_e, j=post_e)
    Ci.connect(i=pre_i, j=post_i)
    mon = SpikeMonitor(P)
    st = StateMonitor(P, ["v", "ge", "gi"], record=[0, 1, 2, 3])
    net = Network(P, Ce, Ci, mon, st)
    n_steps = 10000
    device = get_device()
    ns = dict(taum=taum, taue=taue, taui=taui, Vt=Vt, Vr=Vr, El=El, we=we, wi=wi)
    desc, arrays, cos, clock, k, t_end = bt.prepare_run(device, net, n_steps * defaultclock.dt, namespace=ns)
    assert k == n_steps
    # hand back a state shaped like export_state's: the golden's final state and spikes
    z, meta = _golden("cuba4000")
    trace = synthetic_trace(n_steps)
    state = {"_array_neurongroup_v": z["final"][0], "_array_neurongroup_ge": z["final"][1],
             "_array_neurongroup_gi": z["final"][2], "_array_neurongroup_lastspike": z["lastspike"],
             "_array_neurongroup_not_refractory": np.arange(N) % 3 != 0,
             "_array_spikemonitor_i": z["i"], "_array_spikemonitor_t": z["step"] * meta["dt"],
             "_array_spikemonitor_count": np.bincount(z["i"], minlength=N).astype(np.int32),
             "_array_spikemonitor_N": np.array([len(z["i"])], np.int32),
             "_array_statemonitor_t": np.arange(n_steps) * meta["dt"],
             "_array_statemonitor_v": trace[0], "_array_statemonitor_ge": trace[1],
             "_array_statemonitor_gi": trace[2], "_array_statemonitor_N": np.array([n_steps], np.int32),
             "_array_defaultclock_timestep": np.array([n_steps], np.int64),
             "_array_defaultclock_t": np.array([n_steps * meta["dt"]])}
    bt.finish_run(device, net, state, clock, t_end)
    readback = {"v": P.v_[:], "ge": P.ge_[:], "gi": P.gi_[:], "lastspike": P.lastspike_[:],
                "not_refractory": P.not_refractory[:], "mon_i": mon.i[:], "mon_t": mon.t_[:],
                "mon_count": mon.count[:], "mon_num_spikes": np.array([mon.num_spikes]), "st_t": st.t_[:],
                "st_v": st.v_[:], "st_ge": st.ge_[:], "st_gi": st.gi_[:], "t": np.array([float(net.t_)])}
    _save("cuba4000", desc, arrays, cos, readback, n_steps)


def stdp(bt, lib):
    from brian2 import NeuronGroup, Network, SpikeGeneratorGroup, SpikeMonitor, Synapses, defaultclock, get_device
    from brian2 import second, start_scope, volt

    import oracle_py

    start_scope()
    c = oracle_py.STDP
    defaultclock.dt = 1e-4 * second
    dt = 1e-4
    inp = oracle_py.stdp_inputs(mg.SEED)
    dApost = -c["dApre"] * c["taupre"] / c["taupost"] * 1.05 * c["gmax"]
    ns = dict(taum=c["taum"] * second, taupre=c["taupre"] * second, taupost=c["taupost"] * second,
              Ee=c["Ee"] * volt, vt=c["vt"] * volt, vr=c["vr"] * volt, El=c["El"] * volt, taue=c["taue"] * second,
              gmax=c["gmax"], dApre=c["dApre"] * c["gmax"], dApost=dApost)
    src = SpikeGeneratorGroup(c["n_in"], inp["idx"], inp["steps"] * dt * second, name="inputs")
    G = NeuronGroup(c["n_out"], """dv/dt = (ge * (Ee-v) + El - v) / taum : volt
dge/dt = -ge / taue : 1""", threshold="v>vt", reset="v = vr", method="euler", namespace=ns, name="neurons")
    G.v = inp["v0"] * volt
    S = Synapses(src, G, """w : 1
dApre/dt = -Apre / taupre : 1 (event-driven)
dApost/dt = -Apost / taupost : 1 (event-driven)""", on_pre="""ge += w
Apre += dApre
w = clip(w + Apost, 0, gmax)""", on_post="""Apost += dApost
w = clip(w + Apre, 0, gmax)""", namespace=ns, name="stdp")
    S.connect(i=inp["pre"], j=inp["post"])
    S.w = inp["w"]
    mon = SpikeMonitor(G)
    net = Network(src, G, S, mon)
    n_steps = int(c["n_steps"])
    device = get_device()
    desc, arrays, cos, clock, k, t_end = bt.prepare_run(device, net, n_steps * defaultclock.dt, namespace={})
    assert k == n_steps
    z, meta = _golden("stdp")
    n_syn = len(inp["pre"])
    state = {"_array_neurons_v": z["final"][0], "_array_neurons_ge": z["final"][1],
             "_array_stdp_w": z["syn"][0], "_array_stdp_Apre": z["syn"][1], "_array_stdp_Apost": z["syn"][2],
             "_array_stdp_lastupdate": z["syn"][3],
             "_array_spikemonitor_i": z["i"], "_array_spikemonitor_t": z["step"] * dt,
             "_array_spikemonitor_count": np.bincount(z["i"], minlength=c["n_out"]).astype(np.int32),
             "_array_spikemonitor_N": np.array([len(z["i"])], np.int32)}
    for n in ("lastspike", "not_refractory"):
        if f"_array_neurons_{n}" in arrays:
            state[f"_array_neurons_{n}"] = arrays[f"_array_neurons_{n}"]
    bt.finish_run(device, net, state, clock, t_end)
    readback = {"v": G.v_[:], "ge": G.ge[:], "w": S.w[:], "Apre": S.Apre[:], "Apost": S.Apost[:],
                "lastupdate": S.lastupdate_[:], "mon_i": mon.i[:], "mon_t": mon.t_[:], "mon_count": mon.count[:],
                "n_syn": np.array([n_syn]), "t": np.array([float(net.t_)])}
    _save("stdp", desc, arrays, cos, readback, n_steps)


def synthetic_trace(n_steps):
    """A recorded trace (variable, step, record index) handed back in the copy-back
    check: distinct values that compress well."""
    return (np.arange(3 * n_steps * 4, dtype=np.float64).reshape(3, n_steps, 4) * 0.25 - 7.0) * 1e-3


def _golden(name):
    import ast

    z = np.load(os.path.join(HERE, f"golden_{name}.npz"))
    return z, ast.literal_eval(str(z["meta"]))


def main():
    lib = mg.oracle()
    b2, bt = _setup()
    cuba4000(bt, lib)
    stdp(bt, lib)


if __name__ == "__main__":
    main()
