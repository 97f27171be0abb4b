"""Batched parameter sweeps: many independent networks in one persistent kernel.

A sweep (e.g. 1024 CUBA networks of uneven size, each with its own
connectivity and parameters) is advanced by one kernel per launch in which
every workgroup owns one network and steps it through ``steps_per_launch``
time steps without any grid-wide synchronisation.  Per step and network the
workgroup

  1. updates its neurons (Brian's stateupdate / threshold / reset statements,
     fused; the reset only touches variables no pathway writes),
  2. compacts the spikes in ascending neuron order into LDS (wave ballot +
     prefix over waves),
  3. propagates the spikes in that order, one source row at a time with a
     barrier in between, so that every target receives its events in the order
     of Brian's spike queue (brian2/synapses/spikequeue.h:151-191 with delay 0)
     and the per-target sums are bit-identical to the cython runtime's.

Launches of the step kernel are replayed from a hipGraph.  The block size and
the number of steps per launch are chosen by a small autotuner that times the
candidates on the real networks (results do not depend on either).
"""
import ctypes
import time

import numpy as np

from . import _lib
from .codegen import HostEval, ParamStruct, cast_value, ctype_of, source_header, translate_block
from .kernels import identifiers
from .runtime import Device, Model, dev, ptr, zeros

L = _lib.lib


class SweepNetwork:
    """One network of the sweep: size, exc/inh split, seed, constant overrides."""

    def __init__(self, N, NE, seed, namespace):
        self.N, self.NE, self.seed, self.namespace = int(N), int(NE), int(seed), dict(namespace)


class BatchedCUBA:
    """Sweep of independent networks of one neuron model with two constant-weight
    pathways (exc: sources [0, NE) -> `exc_var`, inh: [NE, N) -> `inh_var`)."""

    def __init__(self, nets, model="cuba", p=0.02, dt=1e-4, float_t="double", record_spikes=True,
                 spike_capacity_hz=200.0, horizon_steps=10000):
        if isinstance(model, str):
            model = Model.load(model)
        self.model = model
        self.nets = list(nets)
        self.p = float(p)
        self.dt = float(dt)
        self.float_t = float_t
        self.record = record_spikes
        self.su = model.group.get("stateupdate")
        self.th = model.group.get("threshold")
        self.rs = model.group.get("reset")
        ex, ih = model.pathways["exc"], model.pathways["inh"]
        self.exc_var, self.inh_var = ex.vector[-1]["var"], ih.vector[-1]["var"]
        self.exc_expr, self.inh_expr = ex.vector[-1]["expr"], ih.vector[-1]["expr"]
        self.exc_spec, self.inh_spec = ex, ih
        if len(ex.vector) != 1 or len(ih.vector) != 1 or ex.vector[0]["op"] != "+=" or ih.vector[0]["op"] != "+=":
            raise NotImplementedError("batched sweeps support on_pre = '<var> += <constant expression>'")
        # state variables of the group (as in NeuronGroup)
        from .runtime import NeuronGroup

        proto = NeuronGroup(1, model, float_t=float_t)
        self.vars = proto.vars
        self.uses_refractory = proto.uses_refractory
        self.t_step = 0
        self.spike_capacity_hz = spike_capacity_hz
        self.horizon_steps = horizon_steps
        self._built = False
        self.block = 256
        self.steps_per_launch = 100

    # ------------------------------------------------------------------ build
    def _scalars(self):
        """Per-network values of every scalar the vector code needs (loop
        invariants of each code object evaluated on the host, and constants)."""
        names = []
        table = []
        for k, net in enumerate(self.nets):
            row = {}
            host = {"dt": self.dt, "N": net.N}
            host.update(net.namespace)
            for tag, spec in (("su", self.su), ("th", self.th), ("rs", self.rs), ("pe", self.exc_spec),
                              ("pi", self.inh_spec)):
                if spec is None:
                    continue
                vals = dict(host)
                for n, v in spec.variables.items():
                    if v["kind"] == "Constant" and n not in vals:
                        vals[n] = v["value"]
                ev = HostEval(vals)
                types = {}
                for st in spec.scalar:
                    ct = ctype_of(st["dtype"], self.float_t, st["var"])
                    ev.values[st["var"]] = cast_value(ev.eval(st["expr"]), ct)
                    types[st["var"]] = ct
                for n in sorted(identifiers(spec.vector)):
                    if n in types or (n in spec.variables and spec.variables[n]["kind"] == "Constant"):
                        ct = types.get(n)
                        v = ev.values[n]
                        if ct is None:
                            ct = "bool" if isinstance(v, bool) else ("long long" if isinstance(v, int) else "double")
                        row[(tag, n)] = (v, ct)
            if k == 0:
                names = sorted(row)
                self.scalar_types = {key: row[key][1] for key in names}
            table.append([float(row[key][0]) for key in names])
        self.scalar_names = names
        return np.array(table, np.float64)

    def _source(self):
        T = self.block
        NW = T // 64
        NS = len(self.scalar_names)
        nmax = max(n.N for n in self.nets)
        ps = ParamStruct("hbb_params")
        ps.ptr("clock", "long long", const=True)
        ps.val("step_off", "long long")
        ps.val("dt", "double")
        ps.val("n_steps", "int")
        ps.val("n_nets", "int")
        ps.ptr("order", "int", const=True)
        ps.ptr("n_off", "int", const=True)
        ps.ptr("ne", "int", const=True)
        ps.ptr("row_ptr", "unsigned int", const=True)
        ps.ptr("post", "int", const=True)
        ps.ptr("scal", "double", const=True)
        for v, ct in self.vars.items():
            ps.ptr(f"a_{v}", "unsigned char" if ct == "bool" else ct)
        ps.ptr("spk_i", "int")
        ps.ptr("spk_step", "int")
        ps.ptr("spk_off", "long long", const=True)
        ps.ptr("spk_cap", "long long", const=True)
        ps.ptr("spk_n", "long long")
        ps.ptr("status", "int")
        self.struct = ps

        def block(tag, spec, extra=None):
            types = dict(self.vars)
            types.update({"t": "double", "dt": "double", "N": "int", "i": "int", "_idx": "int"})
            types.update(extra or {})
            names = {}
            for (tg, n), ct in self.scalar_types.items():
                if tg == tag:
                    types[n] = ct
                    names[n] = f"{tg}_{n}"
            lines = translate_block(spec.vector, spec.conditional_write, types, set(), self.float_t)
            import re

            if names:
                pat = re.compile(r"(?<![\w.])(" + "|".join(re.escape(k) for k in sorted(names, key=len, reverse=True))
                                 + r")(?!\w)")
                lines = [pat.sub(lambda m: names[m.group(1)], ln) for ln in lines]
            return lines

        from .codegen import Renderer

        def value_expr(tag, spec):
            types = {}
            names = {}
            for (tg, n), ct in self.scalar_types.items():
                if tg == tag:
                    types[n] = ct
                    names[n] = f"{tg}_{n}"
            c, _ = Renderer(types).render(spec.vector[-1]["expr"])
            import re

            if names:
                pat = re.compile(r"(?<![\w.])(" + "|".join(re.escape(k) for k in sorted(names, key=len, reverse=True))
                                 + r")(?!\w)")
                c = pat.sub(lambda m: names[m.group(1)], c)
            return c

        k = [source_header(T), ps.c_decl()]
        k.append(f'extern "C" __global__ void __launch_bounds__({T}) batched_step(const hbb_params p) {{')
        k.append("    const int net = p.order[blockIdx.x];")
        k.append("    const int off = p.n_off[net];")
        k.append("    const int N = p.n_off[net + 1] - off;")
        k.append("    const int NE = p.ne[net];")
        k.append("    const double dt = p.dt;")
        for idx, (tg, n) in enumerate(self.scalar_names):
            ct = self.scalar_types[(tg, n)]
            k.append(f"    const {ct} {tg}_{n} = ({ct})p.scal[(size_t)net * {NS} + {idx}];")
        k.append(f"    __shared__ int s_spk[{nmax}];")
        k.append(f"    __shared__ int s_w[{NW + 1}];")
        k.append("    __shared__ int s_n;")
        k.append("    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;")
        k.append("    const unsigned int* rp = p.row_ptr + off + net;  // N + 1 entries per network")
        for (tag, spec, var) in (("pe", self.exc_spec, self.exc_var), ("pi", self.inh_spec, self.inh_var)):
            k.append(f"    const {self.vars[var]} {tag}_val = ({self.vars[var]})({value_expr(tag, spec)});")
        k.append("    for (int s = 0; s < p.n_steps; s++) {")
        k.append("        // the plan's clock counts launches of n_steps steps each")
        k.append("        const long long step = (p.clock[0] + p.step_off) * p.n_steps + s;")
        k.append("        const double t = (double)step * dt;")
        k.append("        if (threadIdx.x == 0) s_n = 0;")
        k.append("        __syncthreads();")
        k.append(f"        for (int j0 = 0; j0 < N; j0 += {T}) {{")
        k.append("            const int i = j0 + threadIdx.x;")
        k.append("            const int _idx = i;")
        k.append("            (void)_idx; (void)t;")
        k.append("            bool hb_cond = false;")
        k.append("            if (i < N) {")
        body = []
        for v, ct in self.vars.items():
            body.append(f"bool {v} = p.a_{v}[off + i] != 0;" if ct == "bool" else f"{ct} {v} = p.a_{v}[off + i];")
        if self.su is not None:
            body.append("{")
            body += ["    " + x for x in block("su", self.su)]
            body.append("}")
        if self.th is not None:
            body.append("{")
            body += ["    " + x for x in block("th", self.th, {"_cond": "bool"})]
            body.append("    hb_cond = _cond;")
            body.append("}")
            if self.uses_refractory:
                body.append("if (hb_cond) { not_refractory = false; lastspike = t; }")
        if self.rs is not None:
            written = {st["var"] for st in self.rs.vector}
            if written & {self.exc_var, self.inh_var}:
                raise NotImplementedError("reset writes a synaptic target variable")
            body.append("if (hb_cond) {")
            body += ["    " + x for x in block("rs", self.rs)]
            body.append("}")
        for v, ct in self.vars.items():
            body.append(f"p.a_{v}[off + i] = {v} ? 1 : 0;" if ct == "bool" else f"p.a_{v}[off + i] = {v};")
        k += ["                " + x for x in body]
        k.append("            }")
        k.append("            const unsigned long long bal = __ballot(hb_cond);")
        k.append("            if (lane == 0) s_w[wid] = __popcll(bal);")
        k.append("            __syncthreads();")
        k.append("            if (threadIdx.x == 0) {")
        k.append("                int acc = s_n;")
        k.append(f"                for (int w = 0; w < {NW}; w++) {{ const int c = s_w[w]; s_w[w] = acc; acc += c; }}")
        k.append(f"                s_w[{NW}] = acc;")
        k.append("            }")
        k.append("            __syncthreads();")
        k.append("            if (hb_cond) s_spk[s_w[wid] + __popcll(bal & ((1ull << lane) - 1ull))] = i;")
        k.append(f"            if (threadIdx.x == 0) s_n = s_w[{NW}];")
        k.append("            __syncthreads();")
        k.append("        }")
        k.append("        const int nspk = s_n;")
        # record
        k.append("        if (nspk) {")
        k.append("            const long long base = p.spk_n[net];")
        k.append("            const long long cap = p.spk_cap[net];")
        k.append(f"            for (int q = threadIdx.x; q < nspk; q += {T}) {{")
        k.append("                if (base + q < cap) {")
        k.append("                    p.spk_i[p.spk_off[net] + base + q] = s_spk[q];")
        k.append("                    p.spk_step[p.spk_off[net] + base + q] = (int)step;")
        k.append("                }")
        k.append("            }")
        k.append("            __syncthreads();")
        k.append("            if (threadIdx.x == 0) { p.spk_n[net] = base + nspk; if (base + nspk > cap) atomicOr(p.status, 2); }")
        k.append("        }")
        # propagate in ascending source order, one row at a time
        ev, iv = self.exc_var, self.inh_var
        k.append("        for (int q = 0; q < nspk; q++) {")
        k.append("            const int src = s_spk[q];")
        k.append("            const unsigned int a = rp[src], b = rp[src + 1];")
        k.append("            if (src < NE) {")
        k.append(f"                for (unsigned int e = a + threadIdx.x; e < b; e += {T}) p.a_{ev}[off + p.post[e]] += pe_val;")
        k.append("            } else {")
        k.append(f"                for (unsigned int e = a + threadIdx.x; e < b; e += {T}) p.a_{iv}[off + p.post[e]] += pi_val;")
        k.append("            }")
        k.append("            __syncthreads();")
        k.append("        }")
        k.append("    }")
        k.append("}")
        return "\n".join(k) + "\n"

    def build(self):
        import torch

        from .synth import S_CONN, cuba_v0

        self.device = D = Device.get()
        sizes = np.array([n.N for n in self.nets], np.int64)
        self.n_off = np.zeros(len(sizes) + 1, np.int64)
        np.cumsum(sizes, out=self.n_off[1:])
        Ntot = int(self.n_off[-1])
        self.state = {}
        host = {}
        for v, ct in self.vars.items():
            if v == "lastspike":
                host[v] = np.full(Ntot, -1e4)
            elif v == "not_refractory":
                host[v] = np.ones(Ntot, np.uint8)
            else:
                host[v] = np.zeros(Ntot)
        for k, net in enumerate(self.nets):
            host["v"][self.n_off[k]:self.n_off[k + 1]] = cuba_v0(net.seed, net.N)
        for v, ct in self.vars.items():
            self.state[v] = dev(host[v], ct)
        # connectivity: per network Bernoulli(p) over all pairs, CSR by source
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        row_ptr = zeros(Ntot + len(sizes), "int")
        counts = []
        for k, net in enumerate(self.nets):
            base = int(self.n_off[k]) + k
            view = row_ptr[base: base + net.N + 1]
            _lib.check(L.hb_gen_bernoulli_rows(net.seed, S_CONN, 0, net.N, net.N, self.p, ptr(view), None, s))
            _lib.check(L.hb_scan_u32(ptr(view), net.N, s))
            counts.append(view[net.N: net.N + 1])
        totals = torch.cat(counts).cpu().numpy().astype(np.int64)
        syn_off = np.zeros(len(sizes) + 1, np.int64)
        np.cumsum(totals, out=syn_off[1:])
        if syn_off[-1] >= 2**32:
            raise NotImplementedError("more than 2^32 synapses in one sweep")
        self.n_syn = int(syn_off[-1])
        self.post = zeros(self.n_syn, "int")
        for k, net in enumerate(self.nets):
            base = int(self.n_off[k]) + k
            view = row_ptr[base: base + net.N + 1]
            pv = self.post[int(syn_off[k]):]
            _lib.check(L.hb_gen_bernoulli_rows(net.seed, S_CONN, 0, net.N, net.N, self.p, ptr(view), ptr(pv), s))
            view += int(syn_off[k])  # absolute positions into the concatenated post array
        self.row_ptr = row_ptr
        self.scal = dev(self._scalars(), "double")
        self.order = dev(np.argsort(-sizes, kind="stable").astype(np.int32), "int")
        self.n_off_d = dev(self.n_off.astype(np.int32), "int")
        self.ne_d = dev(np.array([n.NE for n in self.nets], np.int32), "int")
        cap = np.maximum(1024, (sizes * self.spike_capacity_hz * self.horizon_steps * self.dt).astype(np.int64))
        if not self.record:
            cap = np.zeros_like(cap)
        off = np.zeros(len(sizes), np.int64)
        off[1:] = np.cumsum(cap)[:-1]
        self.spk_cap = dev(cap, "long long")
        self.spk_off = dev(off, "long long")
        self._spk_off_h = off
        self.spk_n = zeros(len(sizes), "long long")
        tot = max(int(cap.sum()), 1)
        self.spk_i = zeros(tot, "int")
        self.spk_step = zeros(tot, "int")
        self.status = zeros(1, "int")
        self._compile()
        self._built = True

    def _compile(self):
        """Generates and compiles the step kernel for the current block size and
        steps per launch, and records it in a one-launch plan."""
        import torch

        D = self.device
        if self.t_step % self.steps_per_launch:
            raise ValueError("the current step is not a multiple of steps_per_launch")
        src = self._source()
        self.source = src
        self.fn = D.function(src, "batched_step")
        cls = self.struct.ctypes_struct()
        self.launch_clock = torch.full((1,), self.t_step // self.steps_per_launch, dtype=torch.int64, device="cuda")
        v = dict(clock=self.launch_clock.data_ptr(), step_off=0, dt=self.dt, n_steps=self.steps_per_launch,
                 n_nets=len(self.nets), order=self.order.data_ptr(), n_off=self.n_off_d.data_ptr(),
                 ne=self.ne_d.data_ptr(), row_ptr=self.row_ptr.data_ptr(), post=self.post.data_ptr(),
                 scal=self.scal.data_ptr(), spk_i=self.spk_i.data_ptr(), spk_step=self.spk_step.data_ptr(),
                 spk_off=self.spk_off.data_ptr(), spk_cap=self.spk_cap.data_ptr(), spk_n=self.spk_n.data_ptr(),
                 status=self.status.data_ptr())
        for n in self.vars:
            v[f"a_{n}"] = self.state[n].data_ptr()
        a = cls()
        for name, _ in cls._fields_:
            if name in v:
                setattr(a, name, v[name])
        self.cls, self.args = cls, a
        self.plan = ctypes.c_void_p()
        _lib.check(L.hb_plan_create(D.ctx, ptr(self.launch_clock), ctypes.byref(self.plan)))
        buf = bytes(a)
        _lib.check(L.hb_plan_add(self.plan, self.fn, len(self.nets), self.block, 0, buf, len(buf),
                                 cls.step_off.offset))

    # ------------------------------------------------------------------ run
    def run(self, n_steps, graph_launches=10):
        """Advances every network by n_steps (a multiple of steps_per_launch)."""
        if not self._built:
            self.build()
        S = self.steps_per_launch
        if n_steps % S:
            raise ValueError(f"n_steps must be a multiple of steps_per_launch={S}")
        self._run_launches(n_steps // S, graph_launches)
        self.t_step += n_steps
        st = int(self.status.item())
        if st:
            raise _lib.HipBrianError("spike record buffer overflow in the batched sweep" if st & 2 else f"status {st}")

    def _run_launches(self, launches, graph_launches):
        g = max(1, min(graph_launches, launches))
        _lib.check(L.hb_plan_run(self.plan, self.device.stream, launches, g if launches >= g else 1))

    def spikes(self, k):
        n = int(self.spk_n[k].item())
        o = int(self._spk_off_h[k])
        i = self.spk_i[o:o + n].cpu().numpy()
        s = self.spk_step[o:o + n].cpu().numpy().astype(np.int64)
        return i, s

    def get(self, var, k):
        a, b = int(self.n_off[k]), int(self.n_off[k + 1])
        return self.state[var][a:b].cpu().numpy()

    # ------------------------------------------------------------------ autotuner
    def snapshot(self):
        return ({v: t.clone() for v, t in self.state.items()}, self.spk_n.clone(), self.t_step)

    def restore(self, snap):
        st, n, t = snap
        for v, tt in st.items():
            self.state[v].copy_(tt)
        self.spk_n.copy_(n)
        self.t_step = t
        self.launch_clock.fill_(t // self.steps_per_launch)

    def autotune(self, blocks=(128, 256, 512, 1024), steps=(10, 50, 100), trial_steps=200):
        """Times each (block size, steps per launch) on the real networks from the
        current state, restores the state, and keeps the fastest."""
        import torch

        if not self._built:
            self.build()
        snap = self.snapshot()
        results = {}
        for T in blocks:
            for S in steps:
                if trial_steps % S:
                    continue
                self.block, self.steps_per_launch = T, S
                self.t_step = snap[2]
                self._compile()
                self.restore(snap)
                self._run_launches(trial_steps // S, max(1, (trial_steps // S)))  # warm (compile + capture)
                torch.cuda.synchronize()
                self.restore(snap)
                t0 = time.perf_counter()
                self._run_launches(trial_steps // S, max(1, (trial_steps // S)))
                torch.cuda.synchronize()
                results[(T, S)] = time.perf_counter() - t0
        best = min(results, key=results.get)
        self.block, self.steps_per_launch = best
        self.t_step = snap[2]
        self._compile()
        self.restore(snap)
        self.autotune_results = results
        return best, results
