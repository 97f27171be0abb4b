"""Device runtime of the `hip` target: groups, pathways, monitors, network.

The objects mirror the Brian objects of the clock-driven step (NeuronGroup,
Synapses with an on_pre pathway, SpikeGeneratorGroup, SpikeMonitor,
StateMonitor, Network.run) and are driven by Brian's optimised statements,
in the JSON form that Brian's code generator receives them
(``brian2_amd/specs``, captured from Brian by tests/golden/make_golden.py).

State lives in device memory as torch tensors (one per variable, structure of
arrays); kernels are generated per group/pathway, compiled with hiprtc and
replayed per step from a hipGraph-captured plan.
"""
import ctypes
import hashlib
import json
import os

import numpy as np

from . import _lib
from . import kernels as _kernels
from .codegen import ctype_of
from .kernels import (
    LDS_BYTES,
    CodeSpec,
    FusedStepKernel,
    GatherKernel,
    NeuronKernel,
    PostPathKernel,
    SummedKernel,
    fold_capacity,
    identifiers,
    rec_words,
    scalar_values,
    synapse_record,
)

L = _lib.lib
S_POISSON_GROUP = 17  # generator stream of PoissonGroup draws (hb_rng.h)
SPEC_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "specs")

_NP = {"double": np.float64, "float": np.float32, "int": np.int32, "long long": np.int64, "bool": np.uint8}


def _torch():
    import torch

    return torch


_TDT = None


def torch_dtype(ct):
    torch = _torch()
    return {"double": torch.float64, "float": torch.float32, "int": torch.int32, "long long": torch.int64,
            "bool": torch.uint8, "unsigned long long": torch.int64, "unsigned int": torch.int32}[ct]


# ---------------------------------------------------------------------------
class Device:
    """One HIP context per process (device 0 unless given)."""

    _instance = None

    def __init__(self, device=0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("the hip target needs a GPU (torch.cuda.is_available() is False)")
        torch.cuda.set_device(device)
        self.device = device
        self.ctx = ctypes.c_void_p()
        _lib.check(L.hb_ctx_create(device, ctypes.byref(self.ctx)), "hb_ctx_create")
        self.n_cu = int(torch.cuda.get_device_properties(device).multi_processor_count)
        self._modules = {}

    @classmethod
    def get(cls):
        if cls._instance is None:
            cls._instance = Device()
        return cls._instance

    @property
    def stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def function(self, source, name):
        key = hashlib.sha1(source.encode()).hexdigest()
        mod = self._modules.get(key)
        if mod is None:
            mod = ctypes.c_void_p()
            log = ctypes.create_string_buffer(1 << 16)
            rc = L.hb_module_load(self.ctx, source.encode(), b"", ctypes.byref(mod), log, len(log))
            if rc != 0:
                raise _lib.HipBrianError("kernel compilation failed:\n" + _lib.last_error() + "\n--- source ---\n"
                                         + source)
            self._modules[key] = mod
        fn = ctypes.c_void_p()
        _lib.check(L.hb_module_function(mod, name.encode(), ctypes.byref(fn)), f"function {name}")
        return fn

    def function_info(self, fn):
        """(VGPRs per thread, spill bytes per thread, static LDS bytes) of a generated kernel."""
        r, sc, lds = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.hb_function_info(fn, ctypes.byref(r), ctypes.byref(sc), ctypes.byref(lds)), "hb_function_info")
        return r.value, sc.value, lds.value

    def spills(self, fn):
        """Whether a generated kernel spills vector registers to memory: by the code object's
        VGPR spill count, or by its private segment when that is unknown (scalar registers
        spill into vector register lanes)."""
        v, s = ctypes.c_int(), ctypes.c_int()
        _lib.check(L.hb_function_spills(fn, ctypes.byref(v), ctypes.byref(s)), "hb_function_spills")
        if v.value < 0:
            return self.function_info(fn)[1] > 0
        return v.value > 0

    def sync(self):
        _torch().cuda.synchronize()


def dev(arr, ct=None):
    torch = _torch()
    a = np.ascontiguousarray(arr)
    if ct is not None:
        a = a.astype(_NP[ct], copy=False)
    return torch.from_numpy(a.copy()).to("cuda")


def zeros(n, ct):
    return _torch().zeros(max(int(n), 1), dtype=torch_dtype(ct), device="cuda")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


# ---------------------------------------------------------------------------
class Model:
    """Brian's statements for one model: group code objects and pathway code objects."""

    def __init__(self, d):
        self.name = d.get("model", "model")
        self.group = {k: CodeSpec(v) for k, v in d["group"].items()}
        self.pathways = {k: CodeSpec(v) for k, v in d.get("pathways", {}).items()}
        self.summed = {k: CodeSpec(v) for k, v in d.get("summed", {}).items()}

    @classmethod
    def load(cls, name):
        path = name if os.path.exists(name) else os.path.join(SPEC_DIR, f"{name}.json")
        with open(path) as f:
            return cls(json.load(f))


# ---------------------------------------------------------------------------
_names = {}


def _unique(base):
    k = _names.get(base, 0)
    _names[base] = k + 1
    return base if k == 0 else f"{base}_{k}"


class _GroupBase:
    def __getitem__(self, sl):
        if not isinstance(sl, slice) or sl.step not in (None, 1):
            raise TypeError("subgroups need a contiguous slice")
        a, b, _ = sl.indices(self.N)
        return Subgroup(self, a, b)


class Subgroup:
    """Contiguous slice of a group (brian2/groups/subgroup.py:10-60)."""

    def __init__(self, group, start, stop):
        self.group, self.start, self.stop = group, start, stop
        self.N = stop - start


class NeuronGroup(_GroupBase):
    """State, stateupdate, threshold and reset of one group.

    ``model`` holds the group's stateupdate/threshold/reset code objects.
    ``namespace`` overrides the values of Brian Constants (e.g. a parameter sweep).
    ``float_t`` is "double" (fp64 verification mode) or "float" (fp32 mode).
    """

    def __init__(self, N, model=None, float_t="double", name=None, namespace=None, variables=None):
        if isinstance(model, str):
            model = Model.load(model)
        if model is None:
            model = Model({"group": {}})
        self.N = int(N)
        self.name = name or _unique("neurongroup")
        self.float_t = float_t
        self.su = model.group.get("stateupdate")
        self.th = model.group.get("threshold")
        self.rs = model.group.get("reset")
        self.namespace = dict(namespace or {})
        self.vars = {}
        owner = None
        for spec in (self.su, self.th, self.rs):
            if spec is None:
                continue
            for n, v in spec.variables.items():
                if v["kind"] in ("ArrayVariable", "DynamicArrayVariable") and not v.get("scalar") \
                        and spec.indices.get(n, "_idx") == "_idx" and n not in ("i", "N") \
                        and v.get("owner") not in (None, "defaultclock"):
                    owner = owner or v["owner"]
                    if v["owner"] == owner:
                        self.vars[n] = ctype_of(v["dtype"], float_t, n)
        for n, dt_ in (variables or {}).items():  # plain state variables, e.g. "v : 1"
            self.vars[n] = ctype_of(dt_, float_t, n)
        self.vars = dict(sorted(self.vars.items()))
        self.uses_refractory = "lastspike" in self.vars and "not_refractory" in self.vars
        self._host = {}
        for n, ct in self.vars.items():
            if n == "lastspike":
                self._host[n] = np.full(self.N, -1e4)
            elif n == "not_refractory":
                self._host[n] = np.ones(self.N, np.uint8)
            else:
                self._host[n] = np.zeros(self.N, _NP[ct])
        self._dev = None
        self.net = None

    # state access (VariableView get/set, brian2/core/variables.py:946)
    def get(self, name):
        if self._dev is not None:
            return self._dev[name].cpu().numpy().copy()
        return self._host[name].copy()

    def set(self, name, values):
        if name not in self.vars:
            raise KeyError(f"{self.name} has no state variable {name!r}")
        a = np.broadcast_to(np.asarray(values), (self.N,)).astype(_NP[self.vars[name]])
        if self._dev is not None:
            self._dev[name].copy_(dev(a))
        else:
            self._host[name] = a.copy()

    def __getattr__(self, k):
        if k.startswith("_") or k not in self.__dict__.get("vars", {}):
            raise AttributeError(k)
        return self.get(k)

    def __setattr__(self, k, v):
        if k in self.__dict__.get("vars", {}):
            self.set(k, v)
        else:
            object.__setattr__(self, k, v)

    def state_tensor(self, name):
        """The device tensor of a state variable (zero-copy handoff)."""
        return self._dev[name]


class SpikeGeneratorGroup(_GroupBase):
    """Replays (index, time) spikes (brian2/input/spikegeneratorgroup.py:21)."""

    def __init__(self, N, indices, times=None, timebins=None, dt=1e-4, period_bins=0, name=None):
        self.N = int(N)
        self.name = name or _unique("spikegeneratorgroup")
        idx = np.asarray(indices, np.int32)
        self._idx_in = idx
        self.times = None if times is None else np.asarray(times, np.float64)
        if timebins is None:
            timebins = self.bins(dt)
        tb = np.asarray(timebins, np.int32)
        order = np.lexsort((idx, tb))
        self.indices, self.timebins = idx[order], tb[order]
        self.period_bins = int(period_bins)
        self.vars = {}
        self._dev = None
        self.net = None
        self.uses_refractory = False


def _spikegen_bins(self, dt):
    """spikegeneratorgroup.py:241-245: times shifted by 1e-3 dt, then truncated."""
    return np.asarray((self.times + 1e-3 * dt) / dt, dtype=np.int32)


def _spikegen_rebin(self, dt):
    """The bins at a new dt (before_run recomputes them when dt changed)."""
    if self.times is None:
        raise NotImplementedError(f"{self.name}: spikes given as time bins cannot follow a change of dt")
    tb = self.bins(dt)
    order = np.lexsort((self._idx_in, tb))
    self.indices, self.timebins = self._idx_in[order], tb[order]


SpikeGeneratorGroup.bins = _spikegen_bins
SpikeGeneratorGroup.rebin = _spikegen_rebin


class PoissonGroup(_GroupBase):
    """N independent Poisson spike sources (brian2/input/poissongroup.py:24): neuron i
    spikes in a step when a uniform draw is below rates[i] * dt (poissongroup.py:106).
    The draws come from a counter-based generator keyed by (seed, step, i), so the
    spike trains depend on nothing but the seed and the step (numpy's generator, which
    Brian uses, is not reproduced: agreement with Brian is statistical)."""

    def __init__(self, N, rates, seed=1234, name=None):
        self.N = int(N)
        self.name = name or _unique("poissongroup")
        self.rates = np.broadcast_to(np.asarray(rates, np.float64), (self.N,)).copy()
        if np.any(self.rates < 0) or not np.all(np.isfinite(self.rates)):
            raise ValueError("rates must be finite and non-negative")
        self.seed = int(seed)
        self.vars = {}
        self._dev = None
        self.net = None
        self.uses_refractory = False


class Synapses:
    """One synaptic object with an on_pre pathway and optionally an on_post one.

    ``model`` is the on_pre pathway's synapses code object (CodeSpec): one
    update of a postsynaptic variable, any number of updates of synaptic
    variables (plasticity, event-driven traces: brian2/synapses/synapses.py:
    359-383), temporaries.  ``on_post`` is the on_post pathway's code object
    (synaptic variables only), run for the synapses of every postsynaptic spike
    after the on_pre pathway of the same step (synapses.py:205-216: order -1
    before 1), with a scalar delay in seconds ``post_delay``.  Connectivity is
    given explicitly with connect(i, j) (synapses.py:1704-1710, synapse index =
    array order), per-synapse variables with set() (unset ones start at 0, as in
    Brian), and delays in seconds with set_delay (scalar or per synapse).
    ``storage`` may store a float64 synaptic variable as float32 on the device
    when every value is exactly representable (checked)."""

    def __init__(self, source, target, model=None, name=None, storage=None, order=0, count_events=False, on_post=None,
                 post_delay=0.0, summed=None):
        self.name = name or _unique("synapses")
        self.count_events = count_events
        self.pathway_name = f"{self.name}_pre"
        self.order = order
        self.post_spec = on_post
        self.post_delay = float(post_delay)
        self.summed = dict(summed or {})  # {target variable: summed_variable code object}
        if isinstance(source, Subgroup):
            self.source, self.src_range = source.group, (source.start, source.stop)
        else:
            self.source, self.src_range = source, (0, source.N)
        # a subgroup target is its group with an index range; connect()'s target indices are
        # relative to it (_target_offset, synapses.py:1393-1396,2005-2006)
        if isinstance(target, Subgroup):
            self.target, self.tgt_range = target.group, (target.start, target.stop)
        else:
            self.target, self.tgt_range = target, (0, target.N)
        self.spec = model
        self.storage = dict(storage or {})
        self._analyse()
        self.pre = None
        self.post = None
        self.delay = 0.0
        self.syn = {}
        self.csr = None
        self.net = None
        self._queue_size = 1      # CSpikeQueue's queue length (spikequeue.h:30-36, 131-149)
        self.carry = None         # events carried over a change of dt (see Network.run)
        self._events_base = 0     # events delivered by earlier builds of the network

    @staticmethod
    def _synaptic_vars(spec, skip, what):
        """Synaptic variables the statements use (reads of neuron variables are not
        supported: propagation runs beside the group updates)."""
        out = set()
        for n in identifiers(spec.vector) - set(skip):
            v = spec.variables.get(n)
            if v is None or v["kind"] in ("Constant", "Function", "AuxiliaryVariable") or n in ("t", "dt"):
                continue
            ix = spec.indices.get(n, "_idx")
            if ix != "_idx":
                raise NotImplementedError(f"{what} code reads {n!r} of a neuron group")
            if n == "_postsynaptic_idx":
                continue
            out.add(n)
        return out

    @staticmethod
    def _synaptic_writes(spec, skip_index, what):
        out = set()
        for k, st in enumerate(spec.vector):
            if k == skip_index or st["op"] == ":=":
                continue
            if spec.indices.get(st["var"], "_idx") != "_idx":
                raise NotImplementedError(f"{what} code may update one postsynaptic variable only")
            out.add(st["var"])
        return out

    def _analyse(self):
        self.plastic = set()
        self.summed_reads = {}  # {summed target: ({pre name: group var}, {post name: group var})}
        used = set()
        for tv, sp in self.summed.items():
            if tv not in self.target.vars:
                raise ValueError(f"summed variable target {tv!r} is not a variable of {self.target.name}")
            pre_r, post_r = {}, {}
            for n in identifiers(sp.vector):
                v = sp.variables.get(n)
                if v is None or v["kind"] in ("Constant", "Function", "AuxiliaryVariable") or n in ("t", "dt"):
                    continue
                ix = sp.indices.get(n, "_idx")
                if ix == "_presynaptic_idx" and n.endswith("_pre"):
                    pre_r[n] = n[:-4]
                elif ix == "_postsynaptic_idx" and n.endswith("_post"):
                    post_r[n] = n[:-5]
                elif ix == "_idx" and n not in ("_presynaptic_idx", "_postsynaptic_idx"):
                    used.add(n)
                elif n not in ("_presynaptic_idx", "_postsynaptic_idx"):
                    raise NotImplementedError(f"summed variable code reads {n!r}")
            self.summed_reads[tv] = (pre_r, post_r)
        if self.spec is None:
            if self.post_spec is not None:
                raise NotImplementedError("an on_post pathway without an on_pre pathway")
            self.target_var = self.op = None
            self.syn_vars = sorted(used)
            return
        vs = self.spec.vector
        if not vs:
            raise NotImplementedError("empty on_pre code")
        idx = self.spec.indices
        tgt = [k for k, st in enumerate(vs) if idx.get(st["var"]) == "_postsynaptic_idx"]
        if len(tgt) != 1 or vs[tgt[0]]["op"] not in ("+=", "-=", "*=", "/=", "="):
            raise NotImplementedError("on_pre code must update exactly one postsynaptic variable")
        self.target_index = tgt[0]
        self.target_var = vs[tgt[0]]["var"]
        self.op = vs[tgt[0]]["op"]
        # synaptic variables written by either pathway (plasticity) are read and written in place
        self.plastic = self._synaptic_writes(self.spec, self.target_index, "on_pre")
        used |= self._synaptic_vars(self.spec, {self.target_var}, "on_pre")
        if self.post_spec is not None:
            if any(self.post_spec.indices.get(st["var"], "_idx") != "_idx" for st in self.post_spec.vector):
                raise NotImplementedError("on_post code may update synaptic variables only")
            self.plastic |= self._synaptic_writes(self.post_spec, -1, "on_post")
            used |= self._synaptic_vars(self.post_spec, (), "on_post")
        self.syn_vars = sorted(used | self.plastic)

    def connect(self, i, j):
        """Synapses from source i to target j (indices relative to the source and target,
        which may be subgroups; synapse index = position in the arrays)."""
        i = np.asarray(i, np.int64)
        j = np.asarray(j, np.int64)
        n_src = self.src_range[1] - self.src_range[0]
        n_tgt = self.tgt_range[1] - self.tgt_range[0]
        if len(i) != len(j):
            raise ValueError("i and j must have the same length")
        if len(i) and (i.min() < 0 or i.max() >= n_src or j.min() < 0 or j.max() >= n_tgt):
            raise IndexError("synapse indices out of range")
        self.pre = (i + self.src_range[0]).astype(np.int32)
        self.post = (j + self.tgt_range[0]).astype(np.int32)

    def set(self, name, values):
        if self.pre is None:
            raise RuntimeError("connect() first")
        a = np.broadcast_to(np.asarray(values, np.float64), (len(self.pre),)).copy()
        self.syn[name] = a
        if self.csr is not None and name in self.csr["syn"]:
            # built: the device copy is in CSR storage order
            t = self.csr["syn"][name]
            order = self.csr.get("order")
            v = a[order] if order is not None else a
            t[: self.csr["n_syn"]].copy_(dev(v.astype(np.float32 if t.dtype == _torch().float32 else np.float64)))

    def set_delay(self, delay):
        self.delay = delay

    def connect_device_csr(self, seg_ptr, post, syn, dmin, n_seg, n_syn):
        """Connectivity built on the device (see synth.py): CSR by (source, delay);
        synapse index = storage position."""
        if self.post_spec is not None or self.summed:
            raise NotImplementedError("on_post pathways and summed variables need connect(i, j)")
        self.csr = dict(seg_ptr=seg_ptr, post=post, syn=syn, dmin=int(dmin), n_seg=int(n_seg), n_syn=int(n_syn))

    def get(self, name):
        """Values of a synaptic variable in synapse order (after a run: as updated by
        the synapses' code)."""
        if self.csr is None:
            if name in self.syn:
                return self.syn[name].copy()
            raise RuntimeError("not built yet")
        t = self.csr["syn"][name][: self.csr["n_syn"]].cpu().numpy().astype(np.float64)
        order = self.csr.get("order")
        if order is None:
            return t
        out = np.empty_like(t)
        out[order] = t
        return out

    # CSpikeQueue::prepare counterpart (spikequeue.h:48-105) on the host
    def _build_csr(self, dt):
        if self.csr is not None:
            return
        S = len(self.pre)
        d = self.delay
        if np.ndim(d) == 0:
            dsteps = np.full(S, int(float(d) / dt + 0.5), np.int64)  # spikequeue.h:90
        else:
            dd = np.asarray(d, np.float64)
            dsteps = (dd / dt + 0.5).astype(np.int64)
        dmin = int(dsteps.min()) if S else 0
        dmax = int(dsteps.max()) if S else 0
        n_seg = dmax - dmin + 1
        self._dsteps = dsteps
        self._queue_size = max(self._queue_size, dmax + 1)  # ensure_delay (spikequeue.h:143-149)
        nsrc = self.src_range[1] - self.src_range[0]
        seg = (self.pre.astype(np.int64) - self.src_range[0]) * n_seg + (dsteps - dmin)
        order = np.lexsort((np.arange(S), seg))  # per (source, delay): synapse order
        counts = np.bincount(seg, minlength=nsrc * n_seg)
        seg_ptr = np.zeros(nsrc * n_seg + 1, np.uint64)
        np.cumsum(counts, out=seg_ptr[1:])
        if S >= 2**32 - 1:
            raise NotImplementedError("more than 2^32 - 1 synapses in one pathway")
        syn = {}
        for n in self.syn_vars:
            v = self.syn[n][order] if n in self.syn else np.zeros(S)  # Brian's variables start at 0
            st = self.storage.get(n)
            if st is not None and np.dtype(st) == np.float32 and n in self.plastic:
                raise ValueError(f"{n!r} is written by the synapses' code: it cannot be stored as float32")
            if st is not None and np.dtype(st) == np.float32:
                v32 = v.astype(np.float32)
                if not np.array_equal(v32.astype(np.float64), v):
                    raise ValueError(f"{n!r} is not exactly representable in float32")
                syn[n] = dev(v32)
            else:
                syn[n] = dev(v)
        self.csr = dict(seg_ptr=dev(seg_ptr.astype(np.uint32).view(np.int32)), post=dev(self.post[order]),
                        syn=syn, dmin=dmin, n_seg=n_seg, n_syn=S, order=order)
        if self.post_spec is not None or self.summed:
            # CSR by target for the on_post pathway and summed variables: storage positions
            # of each target's synapses in synapse order (the post pathway's queue order,
            # summed_variable.pyx's loop order)
            pos_post = self.post[order].astype(np.int64)
            perm = np.lexsort((order, pos_post))
            ptr_ = np.zeros(self.target.N + 1, np.uint64)
            np.cumsum(np.bincount(pos_post, minlength=self.target.N), out=ptr_[1:])
            self.csr["post_ptr"] = dev(ptr_.astype(np.uint32).view(np.int32))
            self.csr["post_syn"] = dev(perm.astype(np.int32))
            self.csr["pre_of"] = dev(self.pre[order].astype(np.int32))

    @property
    def n_events(self):
        """Synapse ids delivered so far (Brian's 'synaptic events', synapses.pyx:42-50)."""
        n = int(self.events_dev.sum().item()) + self._events_base
        if getattr(self, "carry_events", None) is not None:
            n += int(self.carry_events.item())
        return n

    def event_counts(self):
        """Per-synapse delivered-event counts in synapse order (count_events=True)."""
        c = self.syn_count_dev.cpu().numpy()[: self.csr["n_syn"]]
        if "order" in self.csr:
            out = np.empty_like(c)
            out[self.csr["order"]] = c
            return out
        return c

    def syn_ctypes(self):
        out = {}
        for n in self.syn_vars:
            t = self.csr["syn"][n]
            out[n] = "float" if t.dtype == _torch().float32 else "double"
        return out


class SpikeMonitor:
    """Records (i, t) of spikes and a per-neuron count (brian2/monitors/spikemonitor.py:472)."""

    def __init__(self, source, name=None):
        if isinstance(source, Subgroup):
            self.source, self.range = source.group, (source.start, source.stop)
        else:
            self.source, self.range = source, (0, source.N)
        self.name = name or _unique("spikemonitor")
        self._i = []
        self._step = []
        self._t = []  # times of the records (step * dt of the run that recorded them)
        self.count_dev = None
        self.net = None

    def _collect(self, n_new):
        pass

    @property
    def i(self):
        i, s = self._sorted()
        return i

    @property
    def step(self):
        i, s = self._sorted()
        return s

    def t(self, dt):
        return self.step * dt

    @property
    def t_(self):
        """Spike times in seconds (step * dt of the run that recorded them)."""
        return self._sorted(with_t=True)[2]

    def _sorted(self, with_t=False):
        if self.net is not None:
            self.net._drain_monitors()
        i = np.concatenate(self._i) if self._i else np.zeros(0, np.int32)
        s = np.concatenate(self._step) if self._step else np.zeros(0, np.int64)
        t = np.concatenate(self._t) if self._t else np.zeros(0, np.float64)
        # by time (steps of runs with different dt do not compare), then index: spikes of one
        # step are ascending (threshold.pyx:18-32)
        o = np.lexsort((i, t))
        return (i[o], s[o], t[o]) if with_t else (i[o], s[o])

    @property
    def count(self):
        return self.count_dev.cpu().numpy()[: self.range[1] - self.range[0]].copy()

    @property
    def num_spikes(self):
        return len(self.i)


class StateMonitor:
    """Records variables at the start of every step (brian2/monitors/statemonitor.py:82)."""

    def __init__(self, source, variables, record, name=None, when="start"):
        if when not in ("start", "end"):
            raise NotImplementedError("StateMonitor supports when='start' and when='end'")
        self.when = when
        self.source = source
        self.variables = list(variables) if not isinstance(variables, str) else [variables]
        rec = np.asarray(record, np.int32)
        self.record = rec
        self.sorted_idx = np.unique(rec)
        self.name = name or _unique("statemonitor")
        self._rows = []
        self._times = []  # per run: the recorded steps' times
        self.net = None

    @property
    def t_(self):
        """Times of the recorded rows (seconds)."""
        return np.concatenate(self._times) if self._times else np.zeros(0)

    def values(self):
        """Array [n_vars, n_record, n_steps] in the order of `record`."""
        if not self._rows:
            return np.zeros((len(self.variables), len(self.record), 0))
        a = np.concatenate(self._rows, axis=0)  # [steps, vars, sorted_idx]
        pos = np.searchsorted(self.sorted_idx, self.record)
        return np.transpose(a[:, :, pos], (1, 2, 0))


class PopulationRateMonitor:
    """Population rate of a group or subgroup at every step, recorded after the
    resets (brian2/monitors/ratemonitor.py:172, template ratemonitor.pyx:7-48):
    rate = spikes / dt / neurons.  The device counts the spikes of each step; the
    rates are formed on the host with the template's expression."""

    def __init__(self, source, name=None):
        if isinstance(source, Subgroup):
            self.source, self.range = source.group, (source.start, source.stop)
        else:
            self.source, self.range = source, (0, source.N)
        self.name = name or _unique("ratemonitor")
        self._counts = []
        self._steps = []
        self._dts = []  # per run: dt of its steps
        self.net = None

    @property
    def count(self):
        """Spikes of the (sub)group per recorded step."""
        return np.concatenate(self._counts) if self._counts else np.zeros(0, np.int64)

    @property
    def step(self):
        return np.concatenate(self._steps) if self._steps else np.zeros(0, np.int64)

    def t(self, dt):
        return self.step * dt

    @property
    def t_(self):
        """Times of the recorded steps (seconds, with the dt of their run)."""
        if not self._dts:
            return np.zeros(0)
        return np.concatenate([st * d[0] if len(d) else st * 0.0 for st, d in zip(self._steps, self._dts)])

    def rate(self, dt=None):
        """Hz, as ratemonitor.pyx:47: _num_spikes / _clock_dt / _num_source_neurons (dt:
        that of each step's run when not given)."""
        n = self.range[1] - self.range[0]
        if dt is None:
            dt = np.concatenate(self._dts) if self._dts else np.zeros(0)
        return self.count.astype(np.float64) / dt / n

    def smooth_rate(self, dt, window="gaussian", width=None):
        """The rate convolved with a normalised window (ratemonitor.py:63-160):
        "gaussian" (standard deviation `width`, cut at two of them), "flat" (`width`
        rounded to an odd number of steps) or an array of odd length."""
        if width is None and isinstance(window, str):
            raise TypeError("Need a width when using a predefined window.")
        if width is not None and not isinstance(window, str):
            raise TypeError("Can only specify a width for a predefined window")
        if isinstance(window, str):
            if window == "gaussian":
                width_dt = int(np.round(2 * width / dt))
                window = np.exp(-(np.arange(-width_dt, width_dt + 1) ** 2) * 1.0 / (2 * (width / dt) ** 2))
            elif window == "flat":
                window = np.ones(int(np.round(width / (2 * dt))) * 2 + 1)
            else:
                raise NotImplementedError(f'Unknown pre-defined window "{window}"')
        else:
            window = np.asarray(window)
            if window.ndim != 1 or len(window) % 2 != 1:
                raise TypeError("The window has to be one-dimensional with an odd number of values.")
        window = window * 1.0 / sum(window)
        return np.convolve(self.rate(dt), window, mode="same")


# ---------------------------------------------------------------------------
class Network:
    """Runs the step loop on the device (Network.run, brian2/core/network.py:1047-1292)."""

    def __init__(self, *objs, dt=1e-4, bucket_cap=None, graph_steps=100, bucket_mem_frac=0.4, tile=None,
                 gather_pairs=None, npt=None, overlap=True, fuse=True):
        """bucket_cap: events of a tile sorted in LDS at a time (default: what fits two
        workgroups per CU); bucket_mem_frac: share of free device memory the event
        staging may use; tile: threads per workgroup of the step kernels (default:
        1024 for groups of at least 65536 neurons, else 256); npt: neurons per thread
        (default 1); overlap: run the gathers concurrently with the group updates when
        every delay is at least one step (default True); fuse: with overlap, run a
        group of 1024-thread tiles and the gathers into it as one launch per step
        (default True; a dict sets the fused launch's shape: events_per_thread and
        gather_blocks of its gathers, tiles_first, allow_spills); gather_pairs: (delay,
        spike) pairs per gather unit (default:
        about one register window of events).  These shape the launches only: results
        do not depend on them."""
        self.objs = []
        self.dt = float(dt)
        self.bucket_cap = None if bucket_cap is None else int(bucket_cap)
        self.bucket_mem_frac = float(bucket_mem_frac)
        self.graph_steps = int(graph_steps)
        self.tile = tile
        self.npt = npt
        self.gather_pairs = gather_pairs
        self.overlap = bool(overlap)
        self.fuse = isinstance(fuse, dict) or bool(fuse)
        # fused launches: gather events per thread and workgroups per pathway (None: by the
        # build), tiles before gathers in the grid, a two-per-CU build kept despite spills
        self.fuse_cfg = {"events_per_thread": None, "gather_blocks": None, "tiles_first": False,
                         "allow_spills": False}
        if isinstance(fuse, dict):
            unknown = set(fuse) - set(self.fuse_cfg)
            if unknown:
                raise TypeError(f"unknown fused-launch settings {sorted(unknown)}")
            self.fuse_cfg.update(fuse)
        gb = self.fuse_cfg["gather_blocks"]
        if gb is not None and (not isinstance(gb, (list, tuple)) or any(int(b) < 1 for b in gb)):
            raise ValueError("gather_blocks: a list of positive workgroup counts, one per pathway into the group")
        self.t_step = 0
        self._built = False
        for o in objs:
            self.add(o)

    def add(self, *objs):
        for o in objs:
            if self._built:
                raise RuntimeError("cannot add objects after the first run")
            o.net = self
            self.objs.append(o)

    @property
    def t(self):
        return self.t_step * self.dt

    # -- build --------------------------------------------------------------
    def _build(self):
        torch = _torch()
        D = self.device = Device.get()
        groups = [o for o in self.objs if isinstance(o, (NeuronGroup, SpikeGeneratorGroup, PoissonGroup))]
        all_syn = sorted([o for o in self.objs if isinstance(o, Synapses)], key=lambda s: (s.order, s.pathway_name))
        paths = [s for s in all_syn if s.spec is not None]  # synapses with an on_pre pathway
        self.summed_syn = [s for s in all_syn if s.summed]
        smons = [o for o in self.objs if isinstance(o, SpikeMonitor)]
        stmons = [o for o in self.objs if isinstance(o, StateMonitor)]
        self.rmons = [o for o in self.objs if isinstance(o, PopulationRateMonitor)]
        for m in smons + stmons + self.rmons:
            if m.source not in groups:
                raise RuntimeError(f"{m.source.name} recorded by {m.name} is not in the network")
        for s in all_syn:
            for g in (s.source, s.target):
                if g not in groups:
                    raise RuntimeError(f"{g.name} used by {s.name} is not in the network")
        self.groups, self.paths, self.smons, self.stmons = groups, paths, smons, stmons
        self.clock = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.clock.fill_(self.t_step)
        self.status = torch.zeros(1, dtype=torch.int32, device="cuda")
        for s in all_syn:
            s._build_csr(self.dt)
            s.post_delay_steps = int(s.post_delay / self.dt + 0.5)  # spikequeue.h:90
        # per group: device state, spike ring
        for g in groups:
            # threads per workgroup, neurons per thread; a workgroup owns a tile of block * npt neurons
            g.block = int(self.tile or (1024 if g.N >= 65536 else 256))
            g.npt = int(self.npt or 1)
            g.tile = g.block * g.npt
            dmax = max([s.csr["dmin"] + s.csr["n_seg"] - 1 for s in paths if s.source is g]
                       + [s.post_delay_steps for s in paths if s.target is g and s.post_spec is not None] + [0])
            g.ring_slots = dmax + 2
            # the ring of a neuron group is split at the bounds of its pathways' source ranges,
            # so that each gather enumerates only spikes of its own range
            cuts = {0}
            if isinstance(g, NeuronGroup):
                for s in paths:
                    if s.source is g:
                        cuts.update(b for b in s.src_range if 0 < b < g.N)
            g.ring_parts = tuple(sorted(cuts))
            g.ring_spk = zeros(g.ring_slots * g.N, "int")
            g.ring_cnt = zeros(g.ring_slots * len(g.ring_parts), "int")
            if isinstance(g, NeuronGroup):
                g._dev = {n: dev(g._host[n], ct) for n, ct in g.vars.items()}
        # incoming pathways -> buckets
        for g in groups:
            incoming = [s for s in paths if s.target is g]
            g.incoming = incoming
            if not incoming:
                continue
            if not isinstance(g, NeuronGroup):
                raise NotImplementedError("pathways into a spike source")
            vts = {self._value_type(s) for s in incoming}
            if len(vts & {"int", "long long"}) and len(vts) > 1:
                raise NotImplementedError("integer and floating-point pathway targets in one group")
            val_t = "double" if "double" in vts else ("long long" if "long long" in vts else
                                                        ("int" if "int" in vts else "float"))
            g.val_t = val_t
            self._alloc_staging(g, incoming)
        self._plan = ctypes.c_void_p()
        _lib.check(L.hb_plan_create(D.ctx, ptr(self.clock), ctypes.byref(self._plan)), "hb_plan_create")
        self._launches = []
        # with every delay at least one step, the gathers of step s read only spikes of earlier
        # steps: they run concurrently with the group updates of step s and with each other
        # (three side lanes, taken in turn), or inside the target group's launch
        self.concurrent_gathers = (bool(paths) and all(s.csr["dmin"] >= 1 for s in paths) and self.overlap
                                   and not any(s.plastic for s in paths))
        for g in groups:
            g.fused = None
            g.fuse_ok = (self.fuse and self.concurrent_gathers and isinstance(g, NeuronGroup)
                         and bool(g.incoming) and g.block == _kernels.GATHER_BLOCK)
        for s in paths:
            s.fused_into = s.target if s.target.fuse_ok else None
        # summed variables first: their updaters run in the 'groups' slot before the target's
        # state update (synapses.py:146-160, order target.order - 1)
        for s in self.summed_syn:
            self._build_summed(s)
        for g in groups:
            rm = [m for m in self.rmons if m.source is g]
            if len(rm) > 1:
                raise NotImplementedError("more than one PopulationRateMonitor per group")
            g.rmon = rm[0] if rm else None
            if g.rmon is not None:
                g.rmon.buf = None
            if isinstance(g, NeuronGroup):
                self._build_group(g)
            elif isinstance(g, PoissonGroup):
                self._build_poisson(g)
            else:
                self._build_spikegen(g)
        for s in paths:
            if s.fused_into is None:
                self._build_gather(s)
        # on_post pathways after every on_pre pathway of the step (synapses.py:205-216)
        for s in paths:
            if s.post_spec is not None:
                self._build_post(s)
        if self.concurrent_gathers:
            for i, s in enumerate([s for s in paths if s.fused_into is None]):
                _lib.check(L.hb_plan_set_lane(self._plan, s.launch_index, 1 + i % 3), "hb_plan_set_lane")
                if s.carry is not None:
                    _lib.check(L.hb_plan_set_lane(self._plan, s.carry_index, 1 + i % 3), "hb_plan_set_lane")
        self._built = True

    def _alloc_staging(self, g, incoming):
        """Event staging of one target group.

        stg_rec: three rotating buffers of event records (order key and value; one
        written by this step's gathers, one folded, one cleared), each split into
        one region per target tile (tile_off: region starts, n_tiles + 1 offsets).  Each synapse delivers
        at most one event per step, so a region sized by the tile's incoming
        synapses never overflows; when the three buffers would exceed the memory
        budget, every region is scaled down by the same factor and an overflow is
        reported (status bit 1) instead of writing out of bounds.  tile_cnt: events
        of each tile in each buffer.  stg_ord: per-tile index lists for targets with
        more events in one step than half the fold's LDS capacity, bounded by the
        tile's in-degree (status bit 4 when capped)."""
        torch = _torch()
        T = g.tile
        g.n_tiles = (g.N + T - 1) // T
        vbytes = 8 if g.val_t in ("double", "long long") else 4
        # tiles of 1024 neurons, one per thread: two workgroups per CU when the step kernel fits
        # in 64 VGPRs (checked when it is built), so each gets half the LDS
        g.two_per_cu = g.block == 1024 and g.npt == 1
        cap = fold_capacity(g.block, g.val_t, g.npt, wgs_per_cu=2 if g.two_per_cu else None)
        g.lds_cap = int(min(self.bucket_cap, cap) if self.bucket_cap else cap)
        window = _kernels.GATHER_BLOCK * _kernels.GATHER_EVENTS_PER_THREAD
        total = sum(int(s.csr["n_syn"]) for s in incoming)
        for s in incoming:
            self._gather_units(s, window)
        # in-degree of every target over the incoming pathways (N_incoming, synapses.py:1401-1410):
        # hb_csr_count with the targets as sources and one delay segment is a histogram
        indeg_n = torch.zeros(g.N, dtype=torch.int64, device="cuda")
        cnt = zeros(g.N, "int")
        for s in incoming:
            cnt.zero_()
            _lib.check(L.hb_csr_count(ptr(s.csr["post"]), None, int(s.csr["n_syn"]), 0, g.N, 0, 1, ptr(cnt),
                                      self.device.stream), "in-degree count")
            indeg_n += cnt.long()
        del cnt
        indeg_t = torch.nn.functional.pad(indeg_n, (0, g.n_tiles * T - g.N)).view(g.n_tiles, T).sum(1)
        indeg_t = indeg_t.cpu().numpy().astype(np.int64)
        need_tile = int(indeg_t.max()) if total else 0
        need_target = int(indeg_n.max().item()) if total else 0
        del indeg_n
        free, _ = torch.cuda.mem_get_info()
        budget = int(free * self.bucket_mem_frac)
        big = need_target > g.lds_cap // 2
        # two index lists per tile (a target's events, then the same ordered by key)
        g.tile_cap = int(0 if not big else min(need_tile, (budget // 4) // (8 * g.n_tiles)))
        stg_budget = budget - 8 * g.n_tiles * g.tile_cap
        # positions are 32-bit: a region's end plus a gather window's offset into it (below
        # 2^17 events) stays below 2^32 - 1, which marks an event that does not fit
        limit = int(min(total, 2**32 - 2**17, stg_budget // (3 * (8 + vbytes))))
        region = indeg_t if limit >= total else (indeg_t * limit) // max(total, 1)
        off = np.zeros(g.n_tiles + 1, np.int64)
        off[1:] = np.cumsum(region)
        g.stg_cap = int(max(1, off[-1]))
        g.tile_off = dev(off.astype(np.uint32).view(np.int32))
        g.tile_cnt = zeros(3 * g.n_tiles, "int")
        g.stg_rec = torch.empty(3 * g.stg_cap * rec_words(g.val_t), dtype=torch.int32, device="cuda")
        g.stg_ord = torch.empty(max(1, 2 * g.n_tiles * g.tile_cap), dtype=torch.int32, device="cuda")

    def _gather_units(self, s, window):
        """(delay, spike) pairs per gather unit: about one window of events per unit on
        average (s.pairs), and the most units one step can have (s.units_max).  A unit's
        events are a sum over its pairs' segments, so they vary little about the mean:
        units of 95 % of a window rarely need a second window (measured on the 1M-neuron
        network: 0.95 of a window 46-48 us/step against 49 for 0.85)."""
        nsrc = s.src_range[1] - s.src_range[0]
        pairs_max = nsrc * s.csr["n_seg"]
        avg = s.csr["n_syn"] / max(1, pairs_max)
        pc = int(min(_kernels.GATHER_BLOCK, max(32, 0.95 * window / max(avg, 1e-9))))
        s.pairs = int(self.gather_pairs or pc)
        s.units_max = -(-pairs_max // s.pairs)

    def _value_type(self, s):
        from .codegen import Renderer

        types = dict(s.syn_ctypes())
        types["t"] = "double"
        for n, v in s.spec.variables.items():
            if v["kind"] == "Constant":
                types[n] = "double"
        for st in s.spec.scalar:
            types[st["var"]] = ctype_of(st["dtype"], s.target.float_t, st["var"])
        for k, st in enumerate(s.spec.vector):
            if k != s.target_index:
                types.setdefault(st["var"], ctype_of(st["dtype"], s.target.float_t, st["var"]))
        _, t = Renderer(types).render(s.spec.vector[s.target_index]["expr"])
        tv = s.target.vars[s.target_var]
        if tv in ("int", "long long", "bool"):
            return "long long" if tv == "long long" else "int"
        if tv == "float":
            return "float"
        return "float" if t == "float" else "double"

    def _group_scalars(self, g):
        host = {"dt": self.dt, "N": g.N}
        host.update(g.namespace)
        sc = {}
        for tag, spec in (("su", g.su), ("th", g.th), ("rs", g.rs)):
            if spec is None:
                continue
            vals = scalar_values(spec, host, g.float_t)
            for n, vc in vals.items():
                if n in g.namespace:
                    vc = (g.namespace[n], vc[1])
                sc[(tag, n)] = vc
        return sc

    def _reset_mode(self, g):
        if g.rs is None:
            return "none"
        written = set(g.rs.vector_write) | {st["var"] for st in g.rs.vector}
        read = set(g.rs.vector_read) | identifiers(g.rs.vector)
        touched = {s.target_var for s in g.incoming} if getattr(g, "incoming", None) else set()
        return "deferred" if (written | read) & touched else "fused"

    def _build_group(self, g):
        D = self.device
        scalars = self._group_scalars(g)
        fold = []
        for rank, s in enumerate(g.incoming):
            fold.append((rank, s.target_var, s.op))
        mode = self._reset_mode(g)
        if mode == "deferred":
            g.spiked = zeros(g.N, "bool")
        smon = [m for m in self.smons if m.source is g]
        if len(smon) > 1:
            raise NotImplementedError("more than one SpikeMonitor per group")
        stmon = [m for m in self.stmons if m.source is g]
        if len(stmon) > 1:
            raise NotImplementedError("more than one StateMonitor per group")
        g.smon = smon[0] if smon else None
        g.stmon = stmon[0] if stmon else None
        for st in (g.stmon,) if g.stmon else ():
            for v in st.variables:
                if v not in g.vars:
                    raise KeyError(f"StateMonitor variable {v!r} is not a state variable of {g.name}")
        def step_kernel(waves):
            return NeuronKernel(g.name, g.N, g.vars, g.float_t, g.su, g.th, g.rs, mode, scalars, fold,
                                getattr(g, "lds_cap", 0), getattr(g, "val_t", "double"),
                                spikemon=g.smon.range if g.smon else None,
                                statemon=g.stmon.variables if g.stmon else None,
                                uses_refractory=g.uses_refractory,
                                statemon_when=g.stmon.when if g.stmon else "start", tile=g.block, npt=g.npt,
                                ratemon=g.rmon.range if g.rmon else None, waves_per_eu=waves,
                                ring_parts=g.ring_parts)

        # two workgroups of 1024 threads per CU need at most 64 VGPRs: keep that build only
        # if the compiler fits it without spilling
        g.waves = 8 if getattr(g, "two_per_cu", False) else 0
        if g.fuse_ok:
            K = self._build_fused(g, step_kernel)
        else:
            K = step_kernel(g.waves)
            g.fn = D.function(K.source, K.kernel_name)
            if g.waves and D.spills(g.fn):
                g.waves = 0
                K = step_kernel(0)
                g.fn = D.function(K.source, K.kernel_name)
        g.kernel = K
        g.struct_cls = K.struct.ctypes_struct()
        if g.smon:
            self._alloc_spike_monitor(g, g.smon)
        if g.stmon:
            m = g.stmon
            m.rec_idx = dev(m.sorted_idx, "int")
            m.rec_buf = None
        g.launch_index = len(self._launches)
        if g.fused is not None:
            F = g.fused
            g.fused_cls = F.ctypes_struct(g.struct_cls, [s.struct_cls for s in g.incoming])
            for s in g.incoming:
                s.launch_index = g.launch_index
            self._add_launch(g.fn, F.grid, F.block, self._launch_args(g), g.fused_cls)
            # carried ids after the fused launch of the step (its gathers and theirs write the
            # same staging buffer; the keys order them)
            for s in g.incoming:
                if s.carry is not None:
                    self._build_carry(s)
        else:
            self._add_launch(g.fn, (g.N + g.tile - 1) // g.tile, g.block, self._group_args(g), g.struct_cls)
        end_mon = g.stmon is not None and g.stmon.when == "end"
        if g.incoming or mode == "deferred" or end_mon:
            F = NeuronKernel(g.name, g.N, g.vars, g.float_t, g.su, g.th, g.rs, mode, scalars, fold,
                             getattr(g, "lds_cap", 0), getattr(g, "val_t", "double"), flush=True,
                             uses_refractory=g.uses_refractory,
                             statemon=g.stmon.variables if end_mon else None, statemon_when="end", tile=g.block,
                             npt=g.npt, waves_per_eu=g.waves)
            g.flush_fn = D.function(F.source, F.kernel_name)
            g.flush_struct_cls = F.struct.ctypes_struct()
        else:
            g.flush_fn = None

    def _build_fused(self, g, step_kernel):
        """The step kernel of g and the gathers of its incoming pathways as one kernel
        (kernels.FusedStepKernel).  Two workgroups per CU when the whole kernel fits
        in 64 VGPRs without spilling, with gather windows of 4, 3 or 2 events per
        thread (the largest that fits; their LDS then fits beside the fold's share),
        else one, with the gathers' usual windows."""
        D = self.device
        cfg = self.fuse_cfg
        ept = cfg["events_per_thread"]
        tries = [(g.waves, u) for u in ((ept,) if ept else (4, 3, 2))] if g.waves else []
        tries.append((0, ept or _kernels.GATHER_EVENTS_PER_THREAD))
        for k, (waves, ept) in enumerate(tries):
            K = step_kernel(waves)
            for s in g.incoming:
                self._gather_units(s, _kernels.GATHER_BLOCK * ept)
                self._gather_setup(s, ept)
            # 3/8 of the workgroup slots for the gathers, split over the pathways by their
            # synapses (their shares of the events when the sources fire alike).  Measured on
            # the 1M-neuron network (80 % of the synapses excitatory): more gather workgroups
            # hold slots the tiles need, fewer leave the excitatory gather the last to finish
            wgs = 2 if waves else 1
            total = max(len(g.incoming), self.device.n_cu * wgs * 3 // 8)
            nsyn = [max(1, int(s.csr["n_syn"])) for s in g.incoming]
            blocks = cfg["gather_blocks"] or [max(1, round(total * n / sum(nsyn))) for n in nsyn]
            if len(blocks) != len(g.incoming):
                raise ValueError(f"gather_blocks has {len(blocks)} entries for {len(g.incoming)} pathways into "
                                 f"{g.name}")
            blocks = [int(max(1, min(b, s.units_max))) for b, s in zip(blocks, g.incoming)]
            F = FusedStepKernel(K, [s.kernel for s in g.incoming], blocks, g.n_tiles, cfg["tiles_first"])
            fn = D.function(F.source, F.kernel_name)
            spills = D.spills(fn) and not cfg["allow_spills"]
            if waves and k + 1 < len(tries) and (spills or F.lds_bytes > LDS_BYTES // 2):
                continue
            g.waves, g.fn, g.fused = waves, fn, F
            for s, b in zip(g.incoming, blocks):
                s.fn, s.grid = None, b
                s.events_dev = _torch().zeros(b, dtype=_torch().int64, device="cuda")
                s.args.events = s.events_dev.data_ptr()
            return K
        raise AssertionError("unreachable")

    def _launch_args(self, g):
        """Argument struct of g's launch: its step kernel's, or the fused launch's."""
        if g.fused is None:
            return self._group_args(g)
        a = g.fused_cls()
        a.step_off = 0
        a.s = self._group_args(g)
        for i, s in enumerate(g.incoming):
            setattr(a, f"g{i}", s.args)
        return a

    def _alloc_spike_monitor(self, g, m):
        m.cap = max(1 << 20, 2 * g.N * self.graph_steps)
        m.mon_i = zeros(m.cap, "int")
        m.mon_step = zeros(m.cap, "int")
        m.mon_n = zeros(1, "long long")
        m.count_dev = zeros(m.range[1] - m.range[0], "int")
        m._drained = 0

    def _group_args(self, g, flush=False):
        if isinstance(g, PoissonGroup):
            return self._poisson_args(g)
        cls = g.flush_struct_cls if flush else g.struct_cls
        v = {"clock": self.clock.data_ptr(), "step_off": 0, "dt": self.dt, "N": g.N, "ring_slots": g.ring_slots,
             "ring_spk": g.ring_spk.data_ptr(), "ring_cnt": g.ring_cnt.data_ptr(), "ring_stride": g.N,
             "status": self.status.data_ptr()}
        for n in g.vars:
            v[f"a_{n}"] = g._dev[n].data_ptr()
        for (tg, n), (val, ct) in g.kernel.scalars.items():
            v[f"{tg}_{n}"] = val
        if g.incoming:
            v.update(stg_rec=g.stg_rec.data_ptr(), stg_ord=g.stg_ord.data_ptr(),
                     tile_off=g.tile_off.data_ptr(), tile_cnt=g.tile_cnt.data_ptr(), n_tiles=g.n_tiles,
                     tile_cap=g.tile_cap, stg_stride=g.stg_cap)
        if getattr(g, "spiked", None) is not None:
            v["spiked"] = g.spiked.data_ptr()
        if not flush and g.smon:
            m = g.smon
            v.update(mon_i=m.mon_i.data_ptr(), mon_step=m.mon_step.data_ptr(), mon_n=m.mon_n.data_ptr(),
                     mon_count=m.count_dev.data_ptr(), mon_cap=m.cap, mon_start=m.range[0], mon_stop=m.range[1])
        if not flush and g.rmon is not None:
            m = g.rmon
            v.update(rate_cnt=m.buf.data_ptr() if m.buf is not None else 0, rate_row0=getattr(m, "row0", 0),
                     rate_rows=getattr(m, "rows", 0), rate_start=m.range[0], rate_stop=m.range[1])
        if g.stmon and (not flush or g.stmon.when == "end"):
            m = g.stmon
            v.update(rec_idx=m.rec_idx.data_ptr(), rec_buf=m.rec_buf.data_ptr() if m.rec_buf is not None else 0,
                     rec_row0=getattr(m, "row0", 0), rec_rows=getattr(m, "rows", 0), n_rec=len(m.sorted_idx))
        s = cls()
        for name, _ in cls._fields_:
            if name in v:
                setattr(s, name, v[name])
        return s

    def _build_spikegen(self, g):
        fn = ctypes.c_void_p()
        _lib.check(L.hb_spikegen_function(ctypes.byref(fn)))
        g.fn = fn
        if g.times is not None:
            g.rebin(self.dt)
        g.gen_tb = dev(g.timebins, "int")
        g.gen_idx = dev(g.indices, "int")
        g.gen_state = zeros(1, "long long")
        first = int(np.searchsorted(g.timebins, self.t_step, side="left"))
        g.gen_state.fill_(first)

        class A(ctypes.Structure):
            _fields_ = [("timebins", ctypes.c_void_p), ("indices", ctypes.c_void_p), ("n", ctypes.c_int64),
                        ("state", ctypes.c_void_p), ("ring_spk", ctypes.c_void_p), ("ring_cnt", ctypes.c_void_p),
                        ("clock", ctypes.c_void_p), ("step_offset", ctypes.c_int64), ("ring_slots", ctypes.c_int32),
                        ("ring_stride", ctypes.c_int32), ("period_bins", ctypes.c_int32), ("pad", ctypes.c_int32)]

        a = A(g.gen_tb.data_ptr(), g.gen_idx.data_ptr(), len(g.timebins), g.gen_state.data_ptr(),
              g.ring_spk.data_ptr(), g.ring_cnt.data_ptr(), self.clock.data_ptr(), 0, g.ring_slots, g.N,
              g.period_bins, 0)
        g.struct_cls = A
        g.launch_index = len(self._launches)
        self._add_launch(fn, 1, 256, a, A, step_field="step_offset")
        g.flush_fn = None
        g.smon = None
        g.stmon = None

    class _PoissonArgs(ctypes.Structure):
        """struct hb_poisson_args (hipbrian.hip)."""
        _fields_ = [("rdt", ctypes.c_void_p), ("clock", ctypes.c_void_p), ("step_offset", ctypes.c_int64),
                    ("seed", ctypes.c_uint64), ("stream", ctypes.c_uint64), ("ring_spk", ctypes.c_void_p),
                    ("ring_cnt", ctypes.c_void_p), ("ring_slots", ctypes.c_int32), ("ring_stride", ctypes.c_int32),
                    ("N", ctypes.c_int32), ("pad", ctypes.c_int32), ("mon_i", ctypes.c_void_p),
                    ("mon_step", ctypes.c_void_p), ("mon_n", ctypes.c_void_p), ("mon_count", ctypes.c_void_p),
                    ("mon_cap", ctypes.c_int64), ("mon_start", ctypes.c_int32), ("mon_stop", ctypes.c_int32),
                    ("rate_cnt", ctypes.c_void_p), ("rate_row0", ctypes.c_int64), ("rate_rows", ctypes.c_int64),
                    ("rate_start", ctypes.c_int32), ("rate_stop", ctypes.c_int32), ("status", ctypes.c_void_p)]

    def _build_poisson(self, g):
        fn = ctypes.c_void_p()
        _lib.check(L.hb_poisson_function(ctypes.byref(fn)))
        g.fn = fn
        # poissongroup.py:106: the threshold compares rand() with rates * dt (an fp64 product)
        g.rdt = dev(g.rates * self.dt, "double")
        smon = [m for m in self.smons if m.source is g]
        if len(smon) > 1:
            raise NotImplementedError("more than one SpikeMonitor per group")
        g.smon = smon[0] if smon else None
        if g.smon:
            self._alloc_spike_monitor(g, g.smon)
        if any(m.source is g for m in self.stmons):
            raise NotImplementedError("a PoissonGroup has no state variables to record")
        g.stmon = None
        g.struct_cls = self._PoissonArgs
        g.launch_index = len(self._launches)
        self._add_launch(fn, (g.N + 255) // 256, 256, self._poisson_args(g), g.struct_cls, step_field="step_offset")
        g.flush_fn = None

    def _poisson_args(self, g):
        m, r = g.smon, g.rmon
        return self._PoissonArgs(
            g.rdt.data_ptr(), self.clock.data_ptr(), 0, g.seed, S_POISSON_GROUP, g.ring_spk.data_ptr(),
            g.ring_cnt.data_ptr(), g.ring_slots, g.N, g.N, 0,
            m.mon_i.data_ptr() if m else 0, m.mon_step.data_ptr() if m else 0, m.mon_n.data_ptr() if m else 0,
            m.count_dev.data_ptr() if m else 0, m.cap if m else 0, m.range[0] if m else 0, m.range[1] if m else 0,
            r.buf.data_ptr() if r is not None and r.buf is not None else 0, getattr(r, "row0", 0) if r else 0,
            getattr(r, "rows", 0) if r else 0, r.range[0] if r else 0, r.range[1] if r else 0,
            self.status.data_ptr())

    def _build_gather(self, s):
        self._gather_setup(s)
        if s.carry is not None:
            self._build_carry(s)  # before the gather: both write s's synaptic variables and counters
        s.fn = self.device.function(s.kernel.source, s.kernel.kernel_name)
        s.events_dev = _torch().zeros(s.grid, dtype=_torch().int64, device="cuda")
        s.args.events = s.events_dev.data_ptr()
        s.launch_index = len(self._launches)
        self._add_launch(s.fn, s.grid, s.kernel.block, s.args, s.struct_cls)

    @staticmethod
    def _ring_parts_read(s):
        """(parts of the source's ring, first part read, starts of the parts read) for the
        gather of s: the parts that cover its source range."""
        parts = s.source.ring_parts
        bounds = list(parts) + [s.source.N]
        read = [r for r in range(len(parts)) if bounds[r] >= s.src_range[0] and bounds[r + 1] <= s.src_range[1]]
        if not read:  # an empty source range: every part (the gather's range check skips them all)
            read = list(range(len(parts)))
        return (len(parts), read[0], tuple(parts[r] for r in read))

    def _gather_setup(self, s, events_per_thread=None):
        """The gather kernel of pathway s (not compiled), its synapse records and its
        argument struct (s.kernel, s.args, s.struct_cls, s.grid)."""
        g = s.target
        rank = g.incoming.index(s)
        host = {"dt": self.dt}
        host.update(getattr(s, "namespace", {}) or {})
        sc = scalar_values(s.spec, host, g.float_t)
        types_syn = s.syn_ctypes()
        K = GatherKernel(s.name, rank, s.csr["n_seg"], s.csr["dmin"], s.spec.vector, types_syn, sc, g.val_t,
                         g.float_t, s.target_var, g.n_tiles, g.tile, count_synapses=s.count_events,
                         pairs=s.pairs, target_index=s.target_index, plastic=s.plastic,
                         events_per_thread=events_per_thread, ring_parts=self._ring_parts_read(s))
        s.kernel = K
        s.csr["rec"] = self._synapse_records(s, K.rec_arrays)
        if s.csr["rec"].numel() != max(1, s.csr["n_syn"]) * K.rec_words and s.csr["n_syn"] > 0:
            raise RuntimeError("synapse records do not match the gather's layout")
        cls = K.struct.ctypes_struct()
        s.struct_cls = cls
        v = {"clock": self.clock.data_ptr(), "step_off": 0, "ring_spk": s.source.ring_spk.data_ptr(),
             "ring_cnt": s.source.ring_cnt.data_ptr(), "ring_stride": s.source.N, "ring_slots": s.source.ring_slots,
             "src0": s.src_range[0], "src1": s.src_range[1], "seg_ptr": s.csr["seg_ptr"].data_ptr(),
             "syn_rec": s.csr["rec"].data_ptr(), "stg_rec": g.stg_rec.data_ptr(),
             "tile_off": g.tile_off.data_ptr(), "tile_cnt": g.tile_cnt.data_ptr(),
             "status": self.status.data_ptr(), "stg_stride": g.stg_cap}
        for n, (val, ct) in sc.items():
            v[f"c_{n}"] = val
        if s.plastic:
            v["dt"] = self.dt
            for n in s.syn_vars:
                v[f"s_{n}"] = s.csr["syn"][n].data_ptr()
        # enough workgroups to fill every CU (units are taken in turn)
        grid = int(min(self.device.n_cu * K.wgs_per_cu, max(1, s.units_max)))
        if s.count_events and getattr(s, "syn_count_dev", None) is None:
            s.syn_count_dev = zeros(s.csr["n_syn"], "int")
        if s.count_events:
            v["syn_count"] = s.syn_count_dev.data_ptr()
        a = cls()
        for name, _ in cls._fields_:
            if name in v:
                setattr(a, name, v[name])
        s.args, s.grid = a, grid

    def _build_summed(self, s):
        g = s.target
        for tv, (pre_r, post_r) in s.summed_reads.items():
            for grp, names in ((s.source, pre_r), (g, post_r)):
                for gv in names.values():
                    if not isinstance(grp, NeuronGroup) or gv not in grp.vars:
                        raise ValueError(f"summed variable of {s.name} reads {gv!r}, not a variable of {grp.name}")
                    # the fold of a step's events and deferred resets happen at the start of the next
                    # group update: a summed variable computed before it would read stale values
                    if any(p.target is grp and p.target_var == gv for p in self.paths) or (
                            self._reset_mode(grp) == "deferred"):
                        raise NotImplementedError(f"summed variable reading {gv!r}, which synapses or a deferred "
                                                  "reset also update")
            if any(p.target is g and p.target_var == tv for p in self.paths):
                raise NotImplementedError(f"{tv!r} is both a summed variable and the target of a pathway")
        s.summed_kernels = []
        types_syn = s.syn_ctypes()
        for tv, sp in sorted(s.summed.items()):
            pre_r, post_r = s.summed_reads[tv]
            var_types = {n: s.source.vars[gv] for n, gv in pre_r.items()}
            var_types.update({n: g.vars[gv] for n, gv in post_r.items()})
            used = {n: ct for n, ct in types_syn.items() if n in identifiers(sp.vector)}
            host = {"dt": self.dt}
            host.update(getattr(s, "namespace", {}) or {})
            sc = scalar_values(sp, host, g.float_t)
            K = SummedKernel(f"{s.name}_{tv}", sp.vector, used, pre_r, post_r, var_types, g.vars[tv], sc, g.float_t)
            fn = self.device.function(K.source, K.kernel_name)
            cls = K.struct.ctypes_struct()
            t0, t1 = s.tgt_range
            v = {"clock": self.clock.data_ptr(), "step_off": 0, "dt": self.dt, "t0": t0, "t1": t1,
                 "post_ptr": s.csr["post_ptr"].data_ptr(), "post_syn": s.csr["post_syn"].data_ptr(),
                 "pre_of": s.csr["pre_of"].data_ptr(), "target": g._dev[tv].data_ptr()}
            for n in used:
                v[f"s_{n}"] = s.csr["syn"][n].data_ptr()
            for n, gv in pre_r.items():
                v[f"a_{n}"] = s.source._dev[gv].data_ptr()
            for n, gv in post_r.items():
                v[f"a_{n}"] = g._dev[gv].data_ptr()
            for n, (val, ct) in sc.items():
                v[f"c_{n}"] = val
            a = cls()
            for name, _ in cls._fields_:
                if name in v:
                    setattr(a, name, v[name])
            s.summed_kernels.append(K)
            self._add_launch(fn, max(1, (t1 - t0 + K.BLOCK - 1) // K.BLOCK), K.BLOCK, a, cls)

    def _build_post(self, s):
        host = {"dt": self.dt}
        host.update(getattr(s, "namespace", {}) or {})
        sc = scalar_values(s.post_spec, host, s.target.float_t)
        K = PostPathKernel(s.name, s.post_spec.vector, s.syn_ctypes(), s.plastic, sc, s.target.float_t,
                           s.post_delay_steps, ring_parts=s.target.ring_parts)
        s.post_kernel = K
        s.post_fn = self.device.function(K.source, K.kernel_name)
        cls = K.struct.ctypes_struct()
        g = s.target
        v = {"clock": self.clock.data_ptr(), "step_off": 0, "dt": self.dt, "ring_spk": g.ring_spk.data_ptr(),
             "ring_cnt": g.ring_cnt.data_ptr(), "ring_stride": g.N, "ring_slots": g.ring_slots,
             "post_ptr": s.csr["post_ptr"].data_ptr(), "post_syn": s.csr["post_syn"].data_ptr()}
        for n in s.syn_vars:
            v[f"s_{n}"] = s.csr["syn"][n].data_ptr()
        for n, (val, ct) in sc.items():
            v[f"c_{n}"] = val
        a = cls()
        for name, _ in cls._fields_:
            if name in v:
                setattr(a, name, v[name])
        s.post_launch_index = len(self._launches)
        self._add_launch(s.post_fn, int(min(256, max(1, g.N))), K.BLOCK, a, cls)

    @staticmethod
    def _synapse_records(s, types_syn):
        """The gather's synapse records in CSR order (layout: kernels.synapse_record):
        target index and the on_pre statement's synaptic variables side by side, so
        that one load per synapse reads them from one cache line."""
        torch = _torch()
        fields, words, _ = synapse_record(types_syn)
        S = int(s.csr["n_syn"])
        if S == 0:
            return torch.zeros(words, dtype=torch.int32, device="cuda")
        cols = []
        for name, ct in fields:
            if name == "post":
                cols.append(s.csr["post"][:S].view(torch.int32).view(S, 1))
            elif name == "pad_":
                cols.append(torch.zeros((S, 1), dtype=torch.int32, device="cuda"))
            else:
                cols.append(s.csr["syn"][name][:S].contiguous().view(torch.int32).view(S, -1))
        return torch.cat(cols, dim=1).contiguous().view(-1)

    def _add_launch(self, fn, grid, block, args, cls, step_field="step_off"):
        off = getattr(cls, step_field).offset
        buf = bytes(args)
        _lib.check(L.hb_plan_add(self._plan, fn, grid, block, 0, buf, len(buf), off), "hb_plan_add")
        self._launches.append((fn, grid, block, cls, off))

    def _set_launch_args(self, index, args):
        buf = bytes(args)
        _lib.check(L.hb_plan_set_args(self._plan, index, buf, len(buf)), "hb_plan_set_args")

    # -- store / restore (Network.store/restore, brian2/core/network.py:600-740) ------
    def _mutable(self):
        """Device tensors that a run changes, by (object, attribute) or (dict, key)."""
        out = [(self, "clock"), (self, "status")]
        for g in self.groups:
            out += [(g, "ring_spk"), (g, "ring_cnt")]
            for a in ("spiked", "gen_state", "tile_cnt"):
                if getattr(g, a, None) is not None:
                    out.append((g, a))
            m = getattr(g, "smon", None)
            if m is not None:
                out += [(m, "mon_n"), (m, "count_dev")]
        for s in self.paths:
            out.append((s, "events_dev"))
            for n in sorted(s.plastic):
                out.append((s.csr["syn"], n))
            if getattr(s, "syn_count_dev", None) is not None:
                out.append((s, "syn_count_dev"))
            if s.carry is not None:
                out.append((s, "carry_events"))
        return out

    def store(self, name="default"):
        """Keeps a device copy of the network's state (neuron state, spike rings,
        clock, counters, monitor positions) under `name`.  Between runs no event
        is staged (every run ends by folding the last step's events), so the
        staging area needs no copy."""
        if not self._built:
            self._build()
        snap = {"t_step": self.t_step, "tensors": {}, "state": {}, "monitors": {}}
        for o, a in self._mutable():
            snap["tensors"][(id(o), a)] = (o[a] if isinstance(o, dict) else getattr(o, a)).clone()
        for g in self.groups:
            if isinstance(g, NeuronGroup):
                snap["state"][id(g)] = {n: t.clone() for n, t in g._dev.items()}
            m = getattr(g, "smon", None)
            if m is not None:
                snap["monitors"][id(m)] = (list(m._i), list(m._step), list(m._t), m._drained)
            m = getattr(g, "stmon", None)
            if m is not None:
                snap["monitors"][id(m)] = (list(m._rows), list(m._times))
            m = getattr(g, "rmon", None)
            if m is not None:
                snap["monitors"][id(m)] = (list(m._counts), list(m._steps), list(m._dts))
        if not hasattr(self, "_stored"):
            self._stored = {}
        self._stored[name] = snap

    def restore(self, name="default"):
        """Returns the network to the state kept by store(name)."""
        snap = self._stored[name]
        for o, a in self._mutable():
            (o[a] if isinstance(o, dict) else getattr(o, a)).copy_(snap["tensors"][(id(o), a)])
        for g in self.groups:
            if isinstance(g, NeuronGroup):
                for n, t in g._dev.items():
                    t.copy_(snap["state"][id(g)][n])
            m = getattr(g, "smon", None)
            if m is not None:
                i, s, t, d = snap["monitors"][id(m)]
                m._i, m._step, m._t, m._drained = list(i), list(s), list(t), d
            m = getattr(g, "stmon", None)
            if m is not None:
                r, t = snap["monitors"][id(m)]
                m._rows, m._times = list(r), list(t)
            m = getattr(g, "rmon", None)
            if m is not None:
                c, st, d = snap["monitors"][id(m)]
                m._counts, m._steps, m._dts = list(c), list(st), list(d)
        self.t_step = snap["t_step"]

    # -- a change of dt between runs ---------------------------------------------
    EPSILON_DT = 1e-4  # Clock.epsilon_dt (core/clocks.py:290)

    @classmethod
    def calc_timestep(cls, t, dt):
        """Clock._calc_timestep (core/clocks.py:350-371): round(t / dt), or ceil(t / dt)
        when that step's time is not within epsilon_dt * dt of t."""
        new_i = int(np.int64(np.round(t / dt)))
        new_t = new_i * dt
        if new_t == t or abs(new_t - t) / dt <= cls.EPSILON_DT:
            return new_i
        return int(np.int64(np.ceil(t / dt)))

    def _pending_slots(self, s):
        """The synapse ids in s's spike queue after the last step that ran, by relative
        slot (slot r is delivered at step t_step + r), each in the reference queue's
        order: ids carried from an earlier change of dt first, then by push step,
        spiking source (ascending) and synapse index (CSpikeQueue::push,
        spikequeue.h:151-191).  Returns a list of s._queue_size int64 arrays."""
        now = self.t_step
        Q = s._queue_size
        parts = [[] for _ in range(Q)]
        c = s.carry
        if c is not None:
            for rc in range(len(c["ptr"]) - 1):
                r = c["t0"] + rc - now
                if 0 <= r < Q and c["ptr"][rc + 1] > c["ptr"][rc]:
                    parts[r].append(c["syn"][c["ptr"][rc]:c["ptr"][rc + 1]])
        g = s.source
        P = len(g.ring_parts)
        cnt = g.ring_cnt.cpu().numpy().reshape(g.ring_slots, P)
        spk = g.ring_spk.cpu().numpy().reshape(g.ring_slots, g.N)
        bounds = list(g.ring_parts) + [g.N]
        a, b = s.src_range
        d = s._dsteps
        dmax = int(d.max()) if len(d) else 0
        by_src = np.argsort(s.pre, kind="stable")  # each source's synapses in synapse order
        first = np.searchsorted(s.pre[by_src], np.arange(g.N + 1))
        ev_r, ev_syn = [], []
        for st in range(max(0, now - dmax), now):  # push steps, oldest first
            slot = st % g.ring_slots
            x = np.concatenate([spk[slot, bounds[p]:bounds[p] + cnt[slot, p]] for p in range(P)])
            x = np.sort(x[(x >= a) & (x < b)])
            for i in x:
                syn = by_src[first[i]:first[i + 1]]
                r = st + d[syn] - now
                keep = (r >= 0) & (r < Q)
                ev_r.append(r[keep])
                ev_syn.append(syn[keep])
        if ev_r:
            r = np.concatenate(ev_r)
            syn = np.concatenate(ev_syn)
            o = np.argsort(r, kind="stable")
            r, syn = r[o], syn[o]
            cuts = np.searchsorted(r, np.arange(Q + 1))
            for k in range(Q):
                if cuts[k + 1] > cuts[k]:
                    parts[k].append(syn[cuts[k]:cuts[k + 1]])
        return [np.concatenate(p_).astype(np.int64) if p_ else np.zeros(0, np.int64) for p_ in parts]

    def _change_dt(self, new_dt):
        """A run with another dt (Network.run after defaultclock.dt changed): the clock
        moves to the same time at the new dt (Clock._set_t_update_dt,
        core/clocks.py:373-400); every pathway re-bins its spikes in flight the way
        CSpikeQueue::prepare does (spikequeue.h:58-74: the ids of relative slot i move
        to slot (int)(i * old_dt / new_dt + 0.5) of a queue of (int)(size * old_dt /
        new_dt) + 1 slots, a later slot replacing an earlier one that maps to the same
        place), and the network is built again for the new dt (delay steps, integrator
        constants, spike-generator bins, Poisson probabilities) with its state, counters
        and monitors carried over.  The re-binned ids are delivered by the pathways'
        carry kernels before every later push."""
        new_dt = float(new_dt)
        old_dt = self.dt
        new_step = self.calc_timestep(self.t_step * old_dt, new_dt)
        if not self._built:
            self.dt, self.t_step = new_dt, new_step
            return
        torch = _torch()
        all_syn = [o for o in self.objs if isinstance(o, Synapses)]
        for S in all_syn:
            if S.post_spec is not None:
                raise NotImplementedError(f"{S.name}: a change of dt with an on_post pathway")
            if S.spec is not None and getattr(S, "_dsteps", None) is None:
                raise NotImplementedError(f"{S.name}: a change of dt needs connectivity given by connect(i, j)")
        for g in self.groups:
            if isinstance(g, SpikeGeneratorGroup) and g.times is None:
                raise NotImplementedError(f"{g.name}: spikes given as time bins cannot follow a change of dt")
        torch.cuda.synchronize()
        self._drain_monitors()
        factor = old_dt / new_dt
        for S in self.paths:
            slots = self._pending_slots(S)
            newsize = int(S._queue_size * factor) + 1
            new = [np.zeros(0, np.int64)] * newsize
            for i, ids in enumerate(slots):
                k = int(i * factor + 0.5)
                if k < newsize:  # a slot mapped past the end is lost, as in the reference
                    new[k] = ids
            S._queue_size = newsize
            while len(new) > 1 and len(new[-1]) == 0:
                new.pop()
            ptr = np.zeros(len(new) + 1, np.int64)
            np.cumsum([len(x) for x in new], out=ptr[1:])
            S.carry = dict(t0=new_step, ptr=ptr, syn=np.concatenate(new)) if ptr[-1] else None
            # state that outlives the build: counters, synaptic variables the code writes
            S._events_base = S.n_events
            if S.count_events:
                S._count_host = S.event_counts()
            for n in S.plastic:
                S.syn[n] = S.get(n)
        for g in self.groups:
            if isinstance(g, NeuronGroup):
                g._host = {n: g.get(n) for n in g.vars}
                g._dev = None
            elif isinstance(g, SpikeGeneratorGroup):
                g.rebin(new_dt)
            m = getattr(g, "smon", None)
            if m is not None:
                m._count_host = m.count
        _lib.check(L.hb_plan_destroy(self._plan), "hb_plan_destroy")
        self._plan = None
        for S in all_syn:
            S.csr = None
            S.syn_count_dev = None
            S.carry_events = None  # counted in _events_base; a new carry brings its own counter
        self.dt, self.t_step = new_dt, new_step
        self._built = False
        self._stored = {}  # snapshots of the previous build do not apply to the new one
        self._build()
        for S in self.paths:
            if getattr(S, "_count_host", None) is not None:
                S.syn_count_dev[: S.csr["n_syn"]].copy_(dev(S._count_host[S.csr["order"]].astype(np.int32)))
                S._count_host = None
        for g in self.groups:
            m = getattr(g, "smon", None)
            if m is not None and getattr(m, "_count_host", None) is not None:
                m.count_dev[: len(m._count_host)].copy_(dev(m._count_host.astype(np.int32)))
                m._count_host = None

    def _build_carry(self, s):
        """The launch delivering s's carried ids (after a change of dt): CSR positions of
        the listed synapses, slot by slot, for kernels.CarryKernel."""
        c = s.carry
        inv = np.empty(s.csr["n_syn"], np.int64)
        inv[s.csr["order"]] = np.arange(s.csr["n_syn"])
        s.carry_pos = dev(inv[c["syn"]].astype(np.uint32).view(np.int32))
        s.carry_ptr = dev(c["ptr"].astype(np.uint32).view(np.int32))
        s.carry_events = _torch().zeros(1, dtype=_torch().int64, device="cuda")
        K = _kernels.CarryKernel(s.kernel)
        fn = self.device.function(K.source, K.kernel_name)
        cls = K.struct.ctypes_struct()
        g = s.target
        v = {"clock": self.clock.data_ptr(), "step_off": 0, "t0": c["t0"], "n_slots": len(c["ptr"]) - 1,
             "slot_ptr": s.carry_ptr.data_ptr(), "pos": s.carry_pos.data_ptr(), "syn_rec": s.csr["rec"].data_ptr(),
             "stg_rec": g.stg_rec.data_ptr(), "tile_off": g.tile_off.data_ptr(), "tile_cnt": g.tile_cnt.data_ptr(),
             "stg_stride": g.stg_cap, "status": self.status.data_ptr(), "events": s.carry_events.data_ptr()}
        for n, (val, ct) in s.kernel.scalars.items():
            v[f"c_{n}"] = val
        if s.plastic:
            v["dt"] = self.dt
            for n in s.syn_vars:
                v[f"s_{n}"] = s.csr["syn"][n].data_ptr()
        if s.count_events:
            v["syn_count"] = s.syn_count_dev.data_ptr()
        a = cls()
        for name, _ in cls._fields_:
            if name in v:
                setattr(a, name, v[name])
        most = int(np.diff(c["ptr"]).max())
        grid = max(1, min(4 * self.device.n_cu, -(-most // K.BLOCK)))
        s.carry_index = len(self._launches)
        self._add_launch(fn, grid, K.BLOCK, a, cls)

    # -- run ------------------------------------------------------------------
    def run(self, n_steps=None, duration=None, graph_steps=None, timing=False, dt=None):
        """Advances the network by n_steps time steps (or duration seconds).

        dt: the time step of this run, when it differs from the previous run's (see
        _change_dt); timing: record HIP events on the launch stream around every enqueued
        chunk of steps; `last_run_gpu_ms` is then the device time of the step launches (the
        end-of-run fold excluded), without the host round trips between chunks."""
        torch = _torch()
        if dt is not None and float(dt) != self.dt:
            self._change_dt(dt)
        if n_steps is None:
            n_steps = int(round(duration / self.dt))
        n_steps = int(n_steps)
        if not self._built:
            self._build()
        gs = int(graph_steps or self.graph_steps)
        stream = self.device.stream
        # state monitors: one device row per step of this run
        for g in self.groups:
            m = getattr(g, "stmon", None)
            if m is not None:
                m.rows = n_steps
                m.row0 = self.t_step
                m.rec_buf = torch.zeros(max(n_steps, 1) * len(m.variables) * len(m.sorted_idx), dtype=torch.float64,
                                        device="cuda")
                self._set_launch_args(g.launch_index, self._launch_args(g))
            # population rate monitors: spike counts of this run's steps
            m = getattr(g, "rmon", None)
            if m is not None and not isinstance(g, SpikeGeneratorGroup):
                m.rows, m.row0 = n_steps, self.t_step
                m.buf = zeros(max(n_steps, 1), "int")
                self._set_launch_args(g.launch_index, self._launch_args(g))
        done = 0
        if timing:  # HIP events on the plan's stream around each chunk's launches
            _lib.check(L.hb_plan_timing(self._plan, 1, None, None), "hb_plan_timing")
        while done < n_steps:
            chunk = n_steps - done
            for g in self.groups:
                m = getattr(g, "smon", None)
                if m is not None:
                    chunk = min(chunk, self._ensure_monitor_room(g, m, gs))
            # whole graphs of gs steps, then the rest as one graph of its own length (a run
            # shorter than gs steps is still replayed from a captured graph)
            k = gs * (chunk // gs) if chunk >= gs else chunk
            _lib.check(L.hb_plan_run(self._plan, stream, k, gs if k >= gs else k), "hb_plan_run")
            done += k
            self.t_step += k
            st = int(self.status.cpu().item())  # synchronises
            if st:
                raise _lib.HipBrianError(self._status_message(st))
        if timing:
            ms, nr = ctypes.c_double(), ctypes.c_int64()
            _lib.check(L.hb_plan_timing(self._plan, 0, ctypes.byref(ms), ctypes.byref(nr)), "hb_plan_timing")
            self.last_run_gpu_ms = ms.value
        # fold the events delivered by the last step into the state (end of run)
        for g in self.groups:
            if getattr(g, "flush_fn", None) is not None:
                a = self._group_args(g, flush=True)
                off = type(a).step_off.offset
                a.step_off = 0
                buf = bytes(a)
                grid = (g.N + g.tile - 1) // g.tile
                _lib.check(L.hb_launch(g.flush_fn, grid, g.block, 0, stream, buf, len(buf)), "flush")
                if getattr(g, "incoming", None):
                    g.tile_cnt.zero_()  # the flushed events are not folded again by the next run
        for g in self.groups:
            m = getattr(g, "stmon", None)
            if m is not None and n_steps > 0:
                nv, nr = len(m.variables), len(m.sorted_idx)
                m._rows.append(m.rec_buf.view(n_steps, nv, nr).cpu().numpy().copy())
                # rows of steps t_step - n .. t_step - 1, stamped with the step's time (the
                # template records the clock's t, statemonitor.pyx:6-38)
                first = self.t_step - n_steps
                m._times.append(np.arange(first, first + n_steps, dtype=np.float64) * self.dt)
            m = getattr(g, "rmon", None)
            if m is not None and n_steps > 0:
                first = self.t_step - n_steps
                if isinstance(g, SpikeGeneratorGroup):
                    counts = self._spikegen_counts(g, m.range, first, n_steps)
                else:
                    counts = m.buf[:n_steps].cpu().numpy().astype(np.int64)
                m._counts.append(counts)
                m._steps.append(np.arange(first, first + n_steps, dtype=np.int64))
                m._dts.append(np.full(n_steps, self.dt))
        st = int(self.status.cpu().item())
        if st:
            raise _lib.HipBrianError(self._status_message(st))

    @property
    def n_events(self):
        """Synaptic events delivered since the network was built (0 before its first run)."""
        if not self._built:
            return 0
        return sum(s.n_events for s in self.paths)

    def profile(self, n_steps):
        """Runs n_steps as individual launches timed with HIP events (advances the
        simulation).  Returns [(kind, name, total_ms)] per launch of the step."""
        if not self._built:
            self._build()
        for g in self.groups:
            m = getattr(g, "smon", None)
            if m is not None:
                self._ensure_monitor_room(g, m, int(n_steps))
            if getattr(g, "stmon", None) is not None:
                raise NotImplementedError("profile() of a network with a StateMonitor")
            if getattr(g, "rmon", None) is not None:
                raise NotImplementedError("profile() of a network with a PopulationRateMonitor")
        ms = (ctypes.c_float * len(self._launches))()
        _lib.check(L.hb_plan_profile(self._plan, self.device.stream, int(n_steps), ms), "hb_plan_profile")
        self.t_step += int(n_steps)
        st = int(self.status.cpu().item())
        if st:
            raise _lib.HipBrianError(self._status_message(st))
        out = []
        for obj in list(self.groups) + list(self.paths):
            if isinstance(obj, Synapses):
                if obj.fused_into is None:
                    out.append(("gather", obj.name, float(ms[obj.launch_index])))
            else:
                kind = "fused" if getattr(obj, "fused", None) is not None else "step"
                out.append((kind, obj.name, float(ms[obj.launch_index])))
        return out

    @staticmethod
    def _spikegen_counts(g, rng, first, n):
        """Spikes of a SpikeGeneratorGroup in [rng) per step first .. first + n - 1, from
        its (bin, index) lists (the same bins the device replays)."""
        sel = (g.indices >= rng[0]) & (g.indices < rng[1])
        tb = g.timebins[sel].astype(np.int64)
        steps = np.arange(first, first + n, dtype=np.int64)
        if g.period_bins > 0:
            steps = steps % g.period_bins
        per_bin = np.bincount(tb, minlength=int(steps.max()) + 1 if n else 0) if len(tb) else None
        if per_bin is None:
            return np.zeros(n, np.int64)
        out = np.zeros(n, np.int64)
        ok = steps < len(per_bin)
        out[ok] = per_bin[steps[ok]]
        return out

    def _status_message(self, st):
        msg = []
        if st & 1:
            msg.append("event staging overflow (more synaptic events or gather chunks in one step than the "
                       "staging area holds); build the Network with a larger bucket_mem_frac")
        if st & 4:
            msg.append("event burst overflow (a tile received more synaptic events in one step than its index "
                       "list holds); build the Network with a larger bucket_mem_frac")
        if st & 2:
            msg.append("spike monitor buffer overflow")
        if st & 8:
            msg.append("synapse data names a target index outside the target group")
        return "; ".join(msg) or f"device status {st}"

    def _ensure_monitor_room(self, g, m, gs):
        n = int(m.mon_n.cpu().item())
        free = m.cap - n
        if free < g.N * gs:
            torch = _torch()
            newcap = max(2 * m.cap, n + 2 * g.N * gs)
            ni = zeros(newcap, "int")
            ns = zeros(newcap, "int")
            ni[:n] = m.mon_i[:n]
            ns[:n] = m.mon_step[:n]
            m.mon_i, m.mon_step, m.cap = ni, ns, newcap
            self._set_launch_args(g.launch_index, self._launch_args(g))
            free = m.cap - n
        return max(1, free // max(g.N, 1))

    def _drain_monitors(self):
        for g in getattr(self, "groups", []):
            m = getattr(g, "smon", None)
            if m is None:
                continue
            n = int(m.mon_n.cpu().item())
            if n > m._drained:
                m._i.append(m.mon_i[m._drained:n].cpu().numpy().copy())
                m._step.append(m.mon_step[m._drained:n].cpu().numpy().astype(np.int64))
                m._t.append(m._step[-1] * self.dt)
                m._drained = n
