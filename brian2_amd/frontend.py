"""The Brian-independent half of the `hip` device: a network described the way
Brian's runtime holds it, built on the device, and its state handed back.

``from_specs(desc, arrays, code_objects)`` takes

* ``code_objects``: the hot-path code objects of a network as Brian's code
  generator produced them for the hip target -- per code object its name,
  template, owner, the optimised scalar and vector statements of each block,
  its variables (with the device array names) and indices (the form a
  ``HipCodeObject`` records, brian2/codegen/codeobject.py:65-84);
* ``arrays``: the values of the variables by Brian's array names,
  ``_array_<owner>_<var>`` (brian2/devices/device.py:500-515): group state,
  ``_synaptic_pre`` / ``_synaptic_post`` (absolute indices into the underlying
  groups, synapses.py:2000-2006), synaptic variables, pathway delays in
  seconds, monitor record indices, spike generator lists, Poisson rates;
* ``desc``: the object table that the code objects do not carry -- the clock
  (dt and time step), each group's kind and size, each Synapses object's
  source and target with their subgroup ranges and its pathways (schedule
  order, delay array, code objects), the monitors and their sources.

and returns a ``runtime.Network`` that runs the step on the device.
``export_state(net)`` returns what a run changed, again by Brian's array names
(group state, plastic synaptic variables, monitor records and counts, the
clock), which is what the Brian-side device copies back into its arrays.

brian2_amd.brian_target produces ``desc``/``arrays``/``code_objects`` from a
live brian2 Network and calls these two functions; nothing here imports
brian2.
"""
import numpy as np

from .kernels import CodeSpec
from .runtime import (
    Model,
    Network,
    NeuronGroup,
    PoissonGroup,
    PopulationRateMonitor,
    SpikeGeneratorGroup,
    SpikeMonitor,
    StateMonitor,
    Subgroup,
    Synapses,
)

GROUP_TEMPLATES = ("stateupdate", "threshold", "reset")


def _arr(arrays, name, what):
    if name not in arrays:
        raise KeyError(f"{what}: array {name!r} is missing")
    return np.asarray(arrays[name])


def _view(group, start, stop):
    return group if (start, stop) == (0, group.N) else Subgroup(group, start, stop)


def from_specs(desc, arrays, code_objects, float_t="double", **network_options):
    """A device Network from a network described by Brian's objects (see the module
    docstring).  network_options go to runtime.Network (launch shapes only)."""
    cos = {c["name"]: c for c in code_objects}
    by_owner = {}
    for c in code_objects:
        by_owner.setdefault(c["owner"], {})[c["template"]] = c
    dt = float(desc["dt"])
    groups, objs, names = {}, [], {}
    for g in desc["groups"]:
        kind, name, N = g["kind"], g["name"], int(g["N"])
        if kind == "NeuronGroup":
            specs = {t: c for t, c in by_owner.get(name, {}).items() if t in GROUP_TEMPLATES}
            model = Model({"model": name, "group": specs})
            variables = {v: dt_ for v, dt_ in g.get("variables", {}).items()}
            G = NeuronGroup(N, model, float_t=float_t, name=name, variables=variables)
            for v in G.vars:
                an = f"_array_{name}_{v}"
                if an in arrays:
                    G.set(v, np.asarray(arrays[an]))
        elif kind == "SpikeGeneratorGroup":
            idx = _arr(arrays, g["indices"], name)
            times = _arr(arrays, g["times"], name)
            period = float(g.get("period", 0.0))
            G = SpikeGeneratorGroup(N, idx, times=times, dt=dt, period_bins=int(round(period / dt)) if period else 0,
                                    name=name)
        elif kind == "PoissonGroup":
            G = PoissonGroup(N, _arr(arrays, g["rates"], name), seed=int(g.get("seed", 1234)), name=name)
        else:
            raise NotImplementedError(f"group kind {kind!r}")
        groups[name] = G
        objs.append(G)
    syns = {}
    for s in desc.get("synapses", []):
        name = s["name"]
        src = _view(groups[s["source"]], int(s["source_start"]), int(s["source_stop"]))
        tgt = _view(groups[s["target"]], int(s["target_start"]), int(s["target_stop"]))
        pw = {p["prepost"]: p for p in s.get("pathways", [])}
        if set(pw) - {"pre", "post"}:
            raise NotImplementedError(f"{name}: pathways other than on_pre / on_post")
        pre_spec = CodeSpec(cos[pw["pre"]["code_object"]]) if "pre" in pw else None
        post_spec = CodeSpec(cos[pw["post"]["code_object"]]) if "post" in pw else None
        post_delay = 0.0
        if "post" in pw:
            d = np.asarray(_arr(arrays, pw["post"]["delay"], name), np.float64).reshape(-1)
            # Brian keeps a per-synapse delay array unless a scalar delay was given
            # (synapses.py:258-313); the on_post pathway here takes one delay for all synapses
            if d.size > 1 and np.any(d != d[0]):
                raise NotImplementedError(f"{name}: per-synapse delays of an on_post pathway")
            post_delay = float(d[0]) if d.size else 0.0
        summed = {x["target_var"]: CodeSpec(cos[x["code_object"]]) for x in s.get("summed", [])}
        storage = {k: np.float32 for k in s.get("float32_storage", [])}
        S = Synapses(src, tgt, pre_spec, name=name, on_post=post_spec, post_delay=post_delay, summed=summed,
                     order=int(pw["pre"].get("order", -1)) if "pre" in pw else 0, storage=storage,
                     count_events=bool(s.get("count_events", False)))
        pre = _arr(arrays, s["pre"], name).astype(np.int64)
        post = _arr(arrays, s["post"], name).astype(np.int64)
        S.connect(pre - int(s["source_start"]), post - int(s["target_start"]))
        for v, an in s.get("variables", {}).items():
            if v in S.syn_vars:
                S.set(v, _arr(arrays, an, name))
        if "pre" in pw:
            d = np.asarray(_arr(arrays, pw["pre"]["delay"], name), np.float64).reshape(-1)
            S.set_delay(float(d[0]) if d.size == 1 else d)
        syns[name] = S
        objs.append(S)
    mons = {}
    for m in desc.get("monitors", []):
        kind, name = m["kind"], m["name"]
        G = groups[m["source"]]
        if kind == "SpikeMonitor":
            M = SpikeMonitor(_view(G, int(m.get("start", 0)), int(m.get("stop", G.N))), name=name)
        elif kind == "StateMonitor":
            rec = _arr(arrays, m["record"], name) if isinstance(m["record"], str) else np.asarray(m["record"])
            M = StateMonitor(G, list(m["variables"]), record=rec, name=name, when=m.get("when", "start"))
        elif kind == "PopulationRateMonitor":
            M = PopulationRateMonitor(_view(G, int(m.get("start", 0)), int(m.get("stop", G.N))), name=name)
        else:
            raise NotImplementedError(f"monitor kind {kind!r}")
        mons[name] = M
        objs.append(M)
    net = Network(*objs, dt=dt, **network_options)
    net.t_step = int(desc.get("t_step", 0))
    net.brian = {"groups": groups, "synapses": syns, "monitors": mons, "desc": desc,
                 "clock": desc.get("clock", "defaultclock")}
    return net


def export_state(net):
    """What a run changed, by Brian's array names: every state variable of the
    neuron groups, the synaptic variables the synapses' code writes, the monitors'
    records (SpikeMonitor i, t, count, N; StateMonitor t, one [steps, record] array
    per variable, N; PopulationRateMonitor t, rate, N) and the clock's timestep
    and t."""
    b = net.brian
    out = {}
    for name, G in b["groups"].items():
        if isinstance(G, NeuronGroup):
            for v in G.vars:
                out[f"_array_{name}_{v}"] = G.get(v)
    for name, S in b["synapses"].items():
        if S.csr is not None:
            for v in sorted(S.plastic):
                out[f"_array_{name}_{v}"] = S.get(v)
    dt = net.dt
    for name, M in b["monitors"].items():
        if isinstance(M, SpikeMonitor):
            i, step = M.i, M.step
            # spikemonitor.pyx records the index relative to the monitored (sub)group
            out[f"_array_{name}_i"] = (i - M.range[0]).astype(np.int32)
            out[f"_array_{name}_t"] = M.t_
            out[f"_array_{name}_count"] = M.count.astype(np.int32) if M.count_dev is not None else \
                np.zeros(M.range[1] - M.range[0], np.int32)
            out[f"_array_{name}_N"] = np.array([len(i)], np.int32)
        elif isinstance(M, StateMonitor):
            vals = M.values()  # [vars, record, steps]
            steps = vals.shape[2]
            out[f"_array_{name}_t"] = M.t_
            for k, v in enumerate(M.variables):
                out[f"_array_{name}_{v}"] = vals[k].T.copy()
            out[f"_array_{name}_N"] = np.array([steps], np.int32)
        elif isinstance(M, PopulationRateMonitor):
            out[f"_array_{name}_t"] = M.t_
            out[f"_array_{name}_rate"] = M.rate()
            out[f"_array_{name}_N"] = np.array([len(M.step)], np.int32)
    clk = b["clock"]
    out[f"_array_{clk}_timestep"] = np.array([net.t_step], np.int64)
    out[f"_array_{clk}_t"] = np.array([net.t_step * dt], np.float64)
    return out


def update_state(net, desc, arrays):
    """Before a further run of a device network built by from_specs: uploads the
    state that the caller may have changed between runs (every neuron group
    variable and every synaptic variable, by Brian's array names)."""
    b = net.brian
    for g in desc["groups"]:
        G = b["groups"][g["name"]]
        if isinstance(G, NeuronGroup):
            for v in G.vars:
                an = f"_array_{g['name']}_{v}"
                if an in arrays:
                    G.set(v, np.asarray(arrays[an]))
    for s in desc.get("synapses", []):
        S = b["synapses"][s["name"]]
        for v, an in s.get("variables", {}).items():
            if v in S.syn_vars and an in arrays:
                S.set(v, np.asarray(arrays[an]))
