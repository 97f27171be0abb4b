"""The Brian side of the `hip` target: a code generation target and a device
that run a Brian network's step loop through brian2_amd.

    import brian2_amd.brian_target as bt
    bt.register()                      # codegen_targets.add / all_devices["hip"]
    set_device("hip")                  # or prefs.codegen.target = "hip" + set_device
    ...unchanged script: NeuronGroup, Synapses, SpikeMonitor, run(...)...

Pieces (each cites the reference interface it plugs into):

* ``HipCodeGenerator`` (a ``CPPCodeGenerator``, brian2/codegen/generators/
  cpp_generator.py:200): Brian's own ``translate`` makes and optimises the
  statements (translation.py:181-300, optimisation.py:63-167); the generator
  records them, per block, where ``translate_statement_sequence``
  (cpp_generator.py:381-424) receives them.  Those records are what the
  device's kernels are generated from (brian2_amd.codegen).
* ``HipCodeObject`` (codeobject.py:65-133): ``compile_block`` keeps the record,
  ``run_block`` runs nothing (the whole loop runs in ``HipDevice.network_run``).
  Templates: the run-loop templates only (stateupdate, threshold, reset,
  synapses, synapses_push_spikes, spikemonitor, statemonitor, ratemonitor,
  spikegenerator, summed_variable), with the metadata (USES_VARIABLES,
  ITERATE_ALL, ...) of the runtime's cython templates, so the variables a code
  object resolves are the ones the cython runtime resolves.  Everything else
  (connect, string-expression assignments) keeps the fallback targets
  (``codegen.synapse_connect_target``, ``codegen.string_expression_target``).
* ``HipDevice`` (a ``RuntimeDevice``, devices/device.py:470-615) overriding
  ``network_run`` (core/network.py:1047 via core/base.py:331-356): the same
  set-up as ``Network.run`` (clock interval, namespace, ``before_run``), then
  ``describe`` -> ``frontend.from_specs`` -> ``run`` on the GPU ->
  ``frontend.export_state`` -> ``copy_back`` into Brian's arrays, the clock
  set to the end of the run, ``after_run``.  The device network persists
  between runs of one Network (spikes in flight stay in the device ring, as in
  the reference's SpikeQueue); state that the script changed between runs is
  uploaded again before the next one.

brian2 is imported only inside the functions here; nothing else in the package
imports this module.
"""
import numpy as np

from . import frontend

RUN_TEMPLATES = (
    "stateupdate",
    "threshold",
    "reset",
    "synapses",
    "synapses_push_spikes",
    "spikemonitor",
    "statemonitor",
    "ratemonitor",
    "spikegenerator",
    "summed_variable",
)

# statements recorded by HipCodeGenerator.translate_statement_sequence, by code object name
_RECORDED = {}
_REGISTERED = {}


def _dtype_name(dt):
    if dt is bool or dt is np.bool_:
        return "bool"
    if dt is int:
        return "int64"
    if dt is float:
        return "float64"
    try:
        return np.dtype(dt).name
    except TypeError:
        return str(dt)


def _statement(s):
    """One brian2.codegen.statements.Statement as a dict (var, op, expr, dtype, the
    boolean-split forms of optimisation.py:116-145)."""
    d = {"var": s.var, "op": s.op, "expr": str(s.expr), "dtype": _dtype_name(s.dtype),
         "constant": bool(s.constant), "subexpression": bool(s.subexpression), "scalar": bool(s.scalar)}
    if s.used_boolean_variables:
        d["bool_vars"] = list(s.used_boolean_variables)
        d["bool_exprs"] = [[[[v, bool(val)] for v, val in key], str(e)]
                           for key, e in s.boolean_simplified_expressions.items()]
    return d


def _variable(var, device):
    from brian2.core.functions import Function
    from brian2.core.variables import ArrayVariable, AuxiliaryVariable, Constant, DynamicArrayVariable, Subexpression

    if isinstance(var, Function):
        return {"kind": "Function"}
    e = {"kind": type(var).__name__}
    if hasattr(var, "dtype"):
        e["dtype"] = _dtype_name(var.dtype)
    if isinstance(var, Constant):
        v = var.value
        if isinstance(v, (bool, np.bool_)):
            v = bool(v)
        elif isinstance(v, (int, np.integer)):
            v = int(v)
        elif isinstance(v, (float, np.floating)):
            v = float(v)
        e["value"] = v
    elif isinstance(var, ArrayVariable):
        e.update(size=int(var.size), scalar=bool(var.scalar),
                 owner=var.owner.name if var.owner is not None else None,
                 array_name=device.get_array_name(var), read_only=bool(var.read_only),
                 constant=bool(var.constant), dynamic=isinstance(var, DynamicArrayVariable))
    elif isinstance(var, Subexpression):
        e["expr"] = str(var.expr)
    elif isinstance(var, AuxiliaryVariable):
        e["scalar"] = bool(var.scalar)
    return e


def _record(gen, sc_statements, ve_statements):
    """The captured form of one code object: per block the scalar and vector
    statements with the arrays they read and write, the variables they use and
    their indices (the form brian2_amd.kernels.CodeSpec reads)."""
    from brian2.devices.device import get_device
    from brian2.utils.stringtools import get_identifiers

    blocks, used = {}, set()
    for bname in sc_statements:
        sc, ve = sc_statements[bname], ve_statements[bname]
        sread, swrite, sidx, _ = gen.arrays_helper(sc)
        vread, vwrite, vidx, vcw = gen.arrays_helper(ve)
        blocks[str(bname)] = {"scalar": [_statement(s) for s in sc], "vector": [_statement(s) for s in ve],
                              "scalar_read": sorted(sread), "scalar_write": sorted(swrite),
                              "vector_read": sorted(vread), "vector_write": sorted(vwrite),
                              "vector_indices": sorted(vidx), "conditional_write": dict(vcw)}
        for s in list(sc) + list(ve):
            used.add(s.var)
            used |= get_identifiers(str(s.expr))
            for e in s.boolean_simplified_expressions.values() if s.used_boolean_variables else ():
                used |= get_identifiers(str(e))
        used |= set(sread) | set(vread) | set(vidx)
    device = get_device()
    variables = {n: _variable(v, device) for n, v in gen.variables.items() if n in used}
    return {"name": gen.name, "template": gen.template_name,
            "owner": gen.owner.name if gen.owner is not None else None, "blocks": blocks,
            "variables": variables,
            "indices": {n: gen.variable_indices[n] for n in variables if n in gen.variable_indices}}


def register():
    """Registers the `hip` code generation target and device with brian2 (imported
    here).  Returns (HipCodeObject, HipDevice instance).  Idempotent."""
    if _REGISTERED:
        return _REGISTERED["codeobj"], _REGISTERED["device"]
    from brian2.codegen.codeobject import CodeObject
    from brian2.codegen.generators.cpp_generator import CPPCodeGenerator
    from brian2.codegen.runtime.cython_rt.cython_rt import CythonCodeObject
    from brian2.codegen.targets import codegen_targets
    from brian2.devices.device import RuntimeDevice, all_devices

    class HipCodeGenerator(CPPCodeGenerator):
        """Brian's C++ generator, recording the optimised statements it receives."""

        class_name = "hip"

        def translate_statement_sequence(self, sc_statements, ve_statements):
            if self.template_name in RUN_TEMPLATES:
                _RECORDED[self.name] = _record(self, sc_statements, ve_statements)
            return super().translate_statement_sequence(sc_statements, ve_statements)

    class _Template:
        """A run-loop template: the cython runtime template's metadata; rendering
        keeps the generated snippets (the kernels are built per network later)."""

        def __init__(self, name, meta):
            self.name = name
            self.template_source = f"hip:{name}"
            self.variables = getattr(meta, "variables", None)
            self.iterate_all = getattr(meta, "iterate_all", None)
            self.allows_scalar_write = getattr(meta, "allows_scalar_write", False)
            self.writes_read_only = getattr(meta, "writes_read_only", None)

        def __call__(self, scalar_code, vector_code, **kwds):
            return {"scalar": scalar_code, "vector": vector_code}

    def host_target():
        """The target of code objects outside the run loop (connect, string-expression
        assignments, ...): Brian's own 'auto' choice, cython or numpy
        (devices/device.py:45-83)."""
        from brian2.devices.device import auto_target

        return auto_target()

    class HipTemplater:
        def __getattr__(self, name):
            if name.startswith("_"):
                raise AttributeError(name)
            if name not in RUN_TEMPLATES:
                return getattr(host_target().templater, name)
            return _Template(name, getattr(CythonCodeObject.templater, name))

    class HipCodeObject(CodeObject):
        """A code object of the hip target (codeobject.py:65-133)."""

        class_name = "hip"
        generator_class = HipCodeGenerator
        templater = HipTemplater()

        def __init__(self, owner, code, variables, variable_indices, template_name, template_source,
                     compiler_kwds, name="codeobject*"):
            super().__init__(owner, code, variables, variable_indices, template_name, template_source,
                             compiler_kwds, name=name)
            self.hip = _RECORDED.pop(self.name, None)
            if self.hip is None:
                raise RuntimeError(f"no statements were recorded for code object {self.name!r}")

        @classmethod
        def is_available(cls):
            try:
                import torch

                from . import _lib  # noqa: F401

                return bool(torch.cuda.is_available())
            except Exception:
                return False

        def compile_block(self, block):
            return self.hip  # kernels are generated per network (HipDevice.network_run)

        def run_block(self, block):
            if block == "run":
                raise NotImplementedError(f"{self.name}: hip code objects run inside HipDevice.network_run")
            return None

    class HipDevice(RuntimeDevice):
        """Runtime arrays on the host (RuntimeDevice), the run loop on the GPU."""

        def __init__(self):
            super().__init__()
            self._networks = {}  # id(brian Network) -> (weakref, brian2_amd Network, structure key)
            self.hip_seed = 1234
            self.network_options = {}

        def code_object(self, owner, name, abstract_code, variables, template_name, variable_indices,
                        codeobj_class=None, **kwds):
            # templates outside the run loop run on the host target
            if self.code_object_class(codeobj_class) is HipCodeObject and template_name not in RUN_TEMPLATES:
                codeobj_class = host_target()
            return super().code_object(owner, name, abstract_code, variables, template_name, variable_indices,
                                       codeobj_class=codeobj_class, **kwds)

        def seed(self, seed=None):
            super().seed(seed)
            self.hip_seed = int(seed) if seed is not None else 1234

        def network_run(self, net, duration, report=None, report_period=None, namespace=None, profile=None,
                        level=0):
            return _network_run(self, net, duration, namespace=namespace, level=level + 1)

    codegen_targets.add(HipCodeObject)
    dev = HipDevice()
    all_devices["hip"] = dev
    _REGISTERED.update(codeobj=HipCodeObject, device=dev, generator=HipCodeGenerator)
    return HipCodeObject, dev


# ---------------------------------------------------------------------------
def _objects(net):
    """The network's top-level objects by kind (Network.objects, network.py:310)."""
    from brian2.groups.neurongroup import NeuronGroup
    from brian2.input.poissongroup import PoissonGroup
    from brian2.input.spikegeneratorgroup import SpikeGeneratorGroup
    from brian2.monitors.ratemonitor import PopulationRateMonitor
    from brian2.monitors.spikemonitor import SpikeMonitor
    from brian2.monitors.statemonitor import StateMonitor
    from brian2.synapses.synapses import Synapses

    out = {"groups": [], "synapses": [], "monitors": []}
    for obj in net.sorted_objects:
        if isinstance(obj, (NeuronGroup, SpikeGeneratorGroup, PoissonGroup)):
            out["groups"].append(obj)
        elif isinstance(obj, Synapses):
            out["synapses"].append(obj)
        elif isinstance(obj, (SpikeMonitor, StateMonitor, PopulationRateMonitor)):
            out["monitors"].append(obj)
    kinds = (NeuronGroup, SpikeGeneratorGroup, PoissonGroup, Synapses, SpikeMonitor, StateMonitor,
             PopulationRateMonitor)
    for obj in net.objects:
        if not isinstance(obj, kinds):
            raise NotImplementedError(f"the hip target does not run {type(obj).__name__} objects ({obj.name})")
    return out


def _range(g):
    """(underlying group, start, stop) of a group or Subgroup (groups/subgroup.py:10-60)."""
    from brian2.groups.subgroup import Subgroup

    if isinstance(g, Subgroup):
        return g.source, int(g.start), int(g.stop)
    return g, 0, int(len(g))


def describe(net, device, seed=1234):
    """The network as frontend.from_specs takes it: (desc, arrays, code_objects).
    Called after ``net.before_run`` so that every run-loop code object exists."""
    from brian2.core.variables import ArrayVariable
    from brian2.groups.neurongroup import NeuronGroup
    from brian2.input.poissongroup import PoissonGroup
    from brian2.input.spikegeneratorgroup import SpikeGeneratorGroup
    from brian2.monitors.ratemonitor import PopulationRateMonitor
    from brian2.monitors.spikemonitor import SpikeMonitor
    from brian2.monitors.statemonitor import StateMonitor

    objs = _objects(net)
    clocks = {obj.clock for obj in net.sorted_objects}
    if len(clocks) != 1:
        raise NotImplementedError("the hip target runs networks with one clock")
    clock = next(iter(clocks))
    arrays, code_objects = {}, []
    for obj in net.sorted_objects:
        for co in getattr(obj, "_code_objects", []):
            if getattr(co, "hip", None) is not None:
                code_objects.append(co.hip)

    def put(var):
        name = device.get_array_name(var)
        arrays[name] = np.array(var.get_value(), copy=True)
        return name

    desc = {"dt": float(clock.dt_), "t_step": int(clock.variables["timestep"].get_value()[0]),
            "clock": clock.name, "groups": [], "synapses": [], "monitors": []}
    for G in objs["groups"]:
        g = {"name": G.name, "N": int(len(G))}
        if isinstance(G, NeuronGroup):
            g["kind"] = "NeuronGroup"
            g["variables"] = {}
            for n, v in G.variables.items():
                if isinstance(v, ArrayVariable) and v.owner is not None and v.owner.name == G.name and \
                        not v.scalar and v.size == len(G) and n not in ("i", "N") and not n.startswith("_"):
                    g["variables"][n] = _dtype_name(v.dtype)
                    put(v)
        elif isinstance(G, SpikeGeneratorGroup):
            g.update(kind="SpikeGeneratorGroup", indices=put(G.variables["neuron_index"]),
                     times=put(G.variables["spike_time"]), period=float(G.variables["period"].get_value().item()))
        elif isinstance(G, PoissonGroup):
            if "rates" not in G.variables or not isinstance(G.variables["rates"], ArrayVariable):
                raise NotImplementedError(f"{G.name}: rates given as an expression")
            rv = G.variables["rates"]
            name = device.get_array_name(rv)
            arrays[name] = np.broadcast_to(np.asarray(rv.get_value(), np.float64), (len(G),)).copy()
            g.update(kind="PoissonGroup", rates=name, seed=int(seed))
        desc["groups"].append(g)
    for S in objs["synapses"]:
        src, s0, s1 = _range(S.source)
        tgt, t0, t1 = _range(S.target)
        s = {"name": S.name, "source": src.name, "source_start": s0, "source_stop": s1, "target": tgt.name,
             "target_start": t0, "target_stop": t1, "pre": put(S.variables["_synaptic_pre"]),
             "post": put(S.variables["_synaptic_post"]), "pathways": [], "summed": [], "variables": {}}
        for p in S._pathways:
            if p.event != "spike":
                raise NotImplementedError(f"{S.name}: pathways on events other than spikes")
            s["pathways"].append({"prepost": p.prepost, "code_object": p.codeobj.name, "delay": put(p._delays),
                                  "order": int(p.order)})
        for vname, upd in getattr(S, "summed_updaters", {}).items():
            s["summed"].append({"target_var": upd.target_var.name, "code_object": upd.codeobj.name})
        for n, v in S.variables.items():
            if isinstance(v, ArrayVariable) and v.owner is not None and v.owner.name == S.name and \
                    not n.startswith("_") and not v.scalar and n not in ("i", "j", "N", "N_incoming",
                                                                         "N_outgoing", "delay"):
                s["variables"][n] = put(v)
        desc["synapses"].append(s)
    for M in objs["monitors"]:
        G, a, b = _range(M.source)
        m = {"name": M.name, "source": G.name, "start": a, "stop": b}
        if isinstance(M, SpikeMonitor):
            if M.event != "spike":
                raise NotImplementedError(f"{M.name}: EventMonitor of events other than spikes")
            m["kind"] = "SpikeMonitor"
        elif isinstance(M, StateMonitor):
            if a != 0:
                raise NotImplementedError(f"{M.name}: StateMonitor of a subgroup")
            m.update(kind="StateMonitor", variables=list(M.record_variables), record=put(M.variables["_indices"]),
                     when=M.when)
        elif isinstance(M, PopulationRateMonitor):
            m["kind"] = "PopulationRateMonitor"
        desc["monitors"].append(m)
    return desc, arrays, code_objects


def copy_back(net, device, state):
    """Writes export_state's arrays (Brian's array names) into the network's
    variables: group state and plastic synaptic variables in place, monitor
    records resized to their totals (devices/device.py:500-555)."""
    from brian2.core.variables import ArrayVariable, DynamicArrayVariable

    by_name = {}
    for obj in net.sorted_objects:
        for v in getattr(obj, "variables", {}).values():
            if isinstance(v, ArrayVariable) and v.owner is not None:
                by_name.setdefault(device.get_array_name(v), v)
    clocks = {obj.clock.name for obj in net.sorted_objects}
    missing = []
    for name, values in state.items():
        v = by_name.get(name)
        if v is None:
            if not any(name.startswith(f"_array_{c}_") for c in clocks):  # the clock is set by the caller
                missing.append(name)
            continue
        values = np.asarray(values)
        if isinstance(v, DynamicArrayVariable) and not v.scalar and tuple(np.shape(v.get_value())) != values.shape:
            v.resize(values.shape if values.ndim > 1 else len(values))
        if v.scalar:
            device.arrays[v][:] = values.reshape(-1)[:1]
        else:
            v.get_value()[...] = values.astype(v.dtype, copy=False)
    return missing


def _structure_key(desc, arrays):
    """What must not change between runs of one device network: objects, sizes,
    connectivity and delays."""
    parts = [repr(desc["dt"]), repr([(g["kind"], g["name"], g["N"]) for g in desc["groups"]])]
    for s in desc["synapses"]:
        parts.append(s["name"])
        for k in ("pre", "post"):
            parts.append(arrays[s[k]].tobytes())
        for p in s["pathways"]:
            parts.append(np.asarray(arrays[p["delay"]]).tobytes())
    parts.append(repr([(m["kind"], m["name"]) for m in desc["monitors"]]))
    import hashlib

    h = hashlib.sha1()
    for p in parts:
        h.update(p if isinstance(p, bytes) else p.encode())
    return h.hexdigest()


def prepare_run(device, net, duration, namespace=None, level=0):
    """The set-up half of HipDevice.network_run, as Network.run does it
    (core/network.py:1100-1145): the clock's interval, the namespace,
    ``before_run`` (which creates the hip code objects), then ``describe``.
    Returns (desc, arrays, code_objects, clock, n_steps, t_end)."""
    from brian2.core.namespace import get_local_namespace

    if duration < 0:
        raise ValueError(f"Function 'run' expected a non-negative duration but got '{duration}'")
    all_objects = net.sorted_objects
    clocks = {obj.clock for obj in all_objects}
    if len(clocks) != 1:
        raise NotImplementedError("the hip target runs networks with one clock")
    clock = next(iter(clocks))
    net._clocks = clocks
    t_end = net.t + duration
    clock.set_interval(net.t, t_end)
    if namespace is None:
        namespace = get_local_namespace(level=level + 2)
    net.before_run(namespace)
    desc, arrays, code_objects = describe(net, device, seed=device.hip_seed)
    n_steps = int(clock._i_end - clock.variables["timestep"].get_value()[0])
    return desc, arrays, code_objects, clock, n_steps, t_end


def finish_run(device, net, state, clock, t_end):
    """The hand-back half: export_state's arrays into Brian's variables, the clock
    at the end of the run (clocks.py:315-325), ``after_run``."""
    missing = copy_back(net, device, state)
    if missing:
        raise RuntimeError(f"arrays with no Brian variable: {missing}")
    clock.variables["timestep"].get_value()[0] = clock._i_end
    clock.variables["t"].get_value()[0] = clock._i_end * float(clock.dt_)
    net.t_ = float(t_end)
    device._last_run_time = 0.0
    device._last_run_completed_fraction = 1.0
    net.after_run()


def _network_run(device, net, duration, namespace=None, level=0):
    """HipDevice.network_run: Network.run (core/network.py:1047-1290) with the step
    loop on the GPU."""
    import weakref

    if not net.sorted_objects:
        return
    desc, arrays, code_objects, clock, n_steps, t_end = prepare_run(device, net, duration, namespace,
                                                                    level=level + 1)
    key = _structure_key(desc, arrays)
    entry = device._networks.get(id(net))
    if entry is not None and (entry[0]() is not net or entry[2] != key):
        if entry[0]() is net:
            raise NotImplementedError("the hip target keeps a network's connectivity, delays and dt fixed "
                                      "between runs")
        entry = None
    if entry is None:
        hnet = frontend.from_specs(desc, arrays, code_objects, **device.network_options)
        device._networks[id(net)] = (weakref.ref(net), hnet, key)
    else:
        hnet = entry[1]
        frontend.update_state(hnet, desc, arrays)
    hnet.run(n_steps)
    finish_run(device, net, frontend.export_state(hnet), clock, t_end)
