"""Kernel sources generated per group and per synaptic pathway.

Step semantics follow Brian's schedule (brian2/core/network.py:42-50,
1194-1250): stateupdate -> threshold -> spike monitor -> [push, propagate]
per pathway -> reset -> clock.advance.  On the device one step is

  neuron kernel (per group)   fold of the events delivered by propagate(step-1)
                              [deferred reset of step-1] [StateMonitor, 'start']
                              stateupdate; threshold; [fused reset];
                              spike compaction into the ring slot of `step`;
                              [SpikeMonitor]
  gather kernel (per pathway) events of `step` (spikes of step-d whose synapses
                              have delay d, d in [dmin, dmax]) appended to the
                              target group's per-tile buckets

The fold of a target reproduces the cython runtime's sequential left fold in
queue order (brian2/codegen/runtime/cython_rt/templates/synapses.pyx:42-50):
each event carries the key (pathway rank, dmax - delay, CSR position), which is
ordered exactly like (pathway in schedule order, push step, source index,
synapse index), and every target sorts its own events by that key before it
applies them one by one.  No floating-point atomics are used.
"""
import ctypes

from .codegen import (
    ParamStruct,
    Renderer,
    c_literal,
    ctype_of,
    source_header,
    translate_block,
)

T_TILE = 256        # targets (neurons) per workgroup of the neuron kernel
KEY_TGT_SHIFT = 48  # event key: [63:48] target in tile, [47:44] rank, [43:32] delay, [31:0] CSR position
MAX_RANK = 16
MAX_SEG = 4096


class CodeSpec:
    """One captured Brian code object (scalar + vector statements, variables)."""

    def __init__(self, d):
        self.name = d["name"]
        self.template = d["template"]
        self.owner = d.get("owner")
        blocks = d["blocks"]
        b = blocks.get("None") or blocks.get("run") or next(iter(blocks.values()))
        self.scalar = b["scalar"]
        self.vector = b["vector"]
        self.conditional_write = b.get("conditional_write", {})
        self.vector_read = set(b.get("vector_read", []))
        self.vector_write = set(b.get("vector_write", []))
        self.variables = d["variables"]
        self.indices = d.get("indices", {})

    def to_dict(self):
        return {"name": self.name, "template": self.template, "owner": self.owner,
                "blocks": {"None": {"scalar": self.scalar, "vector": self.vector,
                                    "conditional_write": self.conditional_write,
                                    "vector_read": sorted(self.vector_read),
                                    "vector_write": sorted(self.vector_write)}},
                "variables": self.variables, "indices": self.indices}


def identifiers(statements):
    import ast

    out = set()
    for st in statements:
        out.add(st["var"])
        exprs = [st["expr"]] + [e for _, e in st.get("bool_exprs", [])]
        for e in exprs:
            for n in ast.walk(ast.parse(e.strip(), mode="eval")):
                if isinstance(n, ast.Name):
                    out.add(n.id)
    return out


def scalar_values(spec, host_values, float_t):
    """Evaluates a code object's scalar block on the host.

    Returns {name: (value, ctype)} for every scalar name the vector code needs
    (loop invariants and constants)."""
    from .codegen import HostEval, cast_value

    vals = dict(host_values)
    for n, v in spec.variables.items():
        if v["kind"] == "Constant" and n not in vals:
            vals[n] = v["value"]
    ev = HostEval(vals)
    types = {}
    for st in spec.scalar:
        ct = ctype_of(st["dtype"], float_t, st["var"])
        v = cast_value(ev.eval(st["expr"]), ct)
        ev.values[st["var"]] = v
        types[st["var"]] = ct
    needed = identifiers(spec.vector)
    out = {}
    for n in needed:
        if n in types:
            out[n] = (ev.values[n], types[n])
        elif n in spec.variables and spec.variables[n]["kind"] == "Constant":
            v = ev.values[n]
            ct = "bool" if isinstance(v, bool) else ("long long" if isinstance(v, int) else "double")
            out[n] = (v, ct)
    return out


# ---------------------------------------------------------------------------
class NeuronKernel:
    """Fused per-group step kernel (or the end-of-run flush kernel)."""

    def __init__(self, name, N, arrays, float_t, stateupdate=None, threshold=None, reset=None,
                 reset_mode="none", scalars=None, fold=None, bucket_cap=0, val_t="double",
                 spikemon=None, statemon=None, flush=False, uses_refractory=False, statemon_when="start"):
        self.name = name
        self.N = N
        self.arrays = arrays          # {var: ctype} of the group's state arrays
        self.float_t = float_t
        self.su, self.th, self.rs = stateupdate, threshold, reset
        self.reset_mode = reset_mode  # 'fused' | 'deferred' | 'none'
        self.scalars = scalars or {}  # {(tag, name): (value, ctype)}
        self.fold = fold or []        # [(rank, var, op)]
        self.cap = bucket_cap
        self.val_t = val_t
        self.spikemon = spikemon      # (start, stop) or None
        self.statemon = statemon      # [vars] or None
        self.statemon_when = statemon_when
        self.flush = flush
        self.uses_refractory = uses_refractory
        self.struct = None
        self.source = None
        self._build()

    # -- names used by the vector code of one code object ------------------
    def _block(self, tag, spec, extra_types):
        types = {}
        names = {}
        for v, ct in self.arrays.items():
            types[v] = ct
        types["t"] = "double"
        types["dt"] = "double"
        types["N"] = "int"
        types["i"] = "int"
        types["_idx"] = "int"
        types.update(extra_types)
        for (tg, n), (_, ct) in self.scalars.items():
            if tg == tag:
                types[n] = ct
                names[n] = f"p.{tag}_{n}"
        declared = set()
        lines = translate_block(spec.vector, spec.conditional_write, types, declared, self.float_t)
        # rename scalar names to struct fields (whole-identifier substitution)
        import re

        if names:
            pat = re.compile(r"(?<![\w.])(" + "|".join(re.escape(k) for k in sorted(names, key=len, reverse=True)) + r")(?!\w)")
            lines = [pat.sub(lambda m: names[m.group(1)], ln) for ln in lines]
        return lines

    def _build(self):
        ps = ParamStruct(f"hbp_{self.name}")
        ps.ptr("clock", "long long", const=True)
        ps.val("step_off", "long long")
        ps.val("dt", "double")
        ps.val("N", "int")
        ps.val("ring_slots", "int")
        for v, ct in self.arrays.items():
            ps.ptr(f"a_{v}", "unsigned char" if ct == "bool" else ct)
        for (tg, n), (_, ct) in sorted(self.scalars.items()):
            ps.val(f"{tg}_{n}", ct)
        ps.ptr("ring_spk", "int")
        ps.ptr("ring_cnt", "int")
        ps.val("ring_stride", "long long")
        if self.fold:
            ps.ptr("bkt_cnt", "int")
            ps.ptr("bkt_key", "unsigned long long")
            ps.ptr("bkt_val", self.val_t)
            ps.ptr("bkt_ord", "int")
            ps.val("bkt_cap", "int")
            ps.val("pad_bkt", "int")
        if self.reset_mode == "deferred":
            ps.ptr("spiked", "unsigned char")
        ps.ptr("status", "int")
        if self.spikemon:
            ps.ptr("mon_i", "int")
            ps.ptr("mon_step", "int")
            ps.ptr("mon_n", "unsigned long long")
            ps.ptr("mon_count", "int")
            ps.val("mon_cap", "long long")
            ps.val("mon_start", "int")
            ps.val("mon_stop", "int")
        if self.statemon:
            ps.ptr("rec_idx", "int", const=True)
            ps.ptr("rec_buf", "double")
            ps.val("rec_row0", "long long")
            ps.val("rec_rows", "long long")
            ps.val("n_rec", "int")
            ps.val("pad_rec", "int")
        self.struct = ps
        self.source = self._source(ps)

    def _source(self, ps):
        T = T_TILE
        nw = T // 64
        k = [source_header(T), ps.c_decl()]
        kname = f"{'flush' if self.flush else 'step'}_{self.name}"
        self.kernel_name = kname
        k.append(f'extern "C" __global__ void __launch_bounds__({T}) {kname}(const {ps.name} p) {{')
        k.append("    const long long step = p.clock[0] + p.step_off;")
        k.append("    const double t = (double)step * p.dt;")
        k.append("    const double dt = p.dt;")
        k.append("    const int N = p.N;")
        k.append(f"    const int i = blockIdx.x * {T} + threadIdx.x;")
        k.append("    const int _idx = i;")
        k.append("    const bool hb_active = i < N;")
        k.append("    bool hb_cond = false;")
        k.append("    (void)t; (void)dt; (void)_idx;")
        # ---- fold: this tile's events, counting-sorted by target, each run sorted by key.
        # Runs live in LDS when the tile received at most LDS_CAP events, otherwise as
        # an index permutation in global memory (bursts).
        if self.fold:
            cap, vt = self.cap, self.val_t
            k += [
                "    // ---- events delivered by the previous step ----",
                f"    __shared__ unsigned long long s_key[{cap}];",
                f"    __shared__ {vt} s_val[{cap}];",
                f"    __shared__ int s_pos[{T}];",
                f"    __shared__ int s_sh[{nw + 1}];",
                "    __shared__ int s_n;",
                "    if (threadIdx.x == 0) { const int c = p.bkt_cnt[blockIdx.x]; s_n = c < p.bkt_cap ? c : p.bkt_cap; }",
                "    s_pos[threadIdx.x] = 0;",
                "    __syncthreads();",
                "    const int hb_n = s_n;",
                f"    const bool hb_big = hb_n > {cap};",
                "    const size_t hb_off = (size_t)blockIdx.x * (size_t)p.bkt_cap;",
                "    const unsigned long long* hb_gk = p.bkt_key + hb_off;",
                f"    const {vt}* hb_gv = p.bkt_val + hb_off;",
                "    int* hb_go = p.bkt_ord + hb_off;",
                f"    for (int e = threadIdx.x; e < hb_n; e += {T}) atomicAdd(&s_pos[(int)(hb_gk[e] >> {KEY_TGT_SHIFT})], 1);",
                "    __syncthreads();",
                "    const int hb_cnt = s_pos[threadIdx.x];",
                "    int hb_tot;",
                "    const int hb_beg = hb_block_exscan(hb_cnt, s_sh, &hb_tot);",
                "    s_pos[threadIdx.x] = hb_beg;",
                "    __syncthreads();",
                "    if (!hb_big) {",
                f"        for (int e = threadIdx.x; e < hb_n; e += {T}) {{",
                "            const unsigned long long kk = hb_gk[e];",
                f"            const int q = atomicAdd(&s_pos[(int)(kk >> {KEY_TGT_SHIFT})], 1);",
                "            s_key[q] = kk;",
                "            s_val[q] = hb_gv[e];",
                "        }",
                "    } else {",
                f"        for (int e = threadIdx.x; e < hb_n; e += {T}) {{",
                f"            const int q = atomicAdd(&s_pos[(int)(hb_gk[e] >> {KEY_TGT_SHIFT})], 1);",
                "            hb_go[q] = e;",
                "        }",
                "    }",
                "    __syncthreads();",
                "    if (threadIdx.x == 0) p.bkt_cnt[blockIdx.x] = 0;",
                "    if (!hb_big) hb_sort_kv(s_key + hb_beg, s_val + hb_beg, hb_cnt);",
                "    else hb_sort_idx(hb_go + hb_beg, hb_gk, hb_cnt);",
            ]
        # ---- per-neuron body ------------------------------------------------------
        k.append("    if (hb_active) {")
        body = []
        for v, ct in self.arrays.items():
            body.append(f"bool {v} = p.a_{v}[i] != 0;" if ct == "bool" else f"{ct} {v} = p.a_{v}[i];")
        if self.fold:
            body.append("for (int e = hb_beg; e < hb_beg + hb_cnt; e++) {")
            body.append("    unsigned long long hb_k;")
            body.append(f"    {self.val_t} hb_v;")
            body.append("    if (!hb_big) { hb_k = s_key[e]; hb_v = s_val[e]; }")
            body.append("    else { const int r = hb_go[e]; hb_k = hb_gk[r]; hb_v = hb_gv[r]; }")
            body.append("    switch ((int)((hb_k >> 44) & 15)) {")
            for rank, var, op in self.fold:
                body.append(f"    case {rank}: {var} {op} hb_v; break;")
            body.append("    default: break;")
            body.append("    }")
            body.append("}")
        if self.reset_mode == "deferred":
            body.append("if (p.spiked[i]) { // reset of the previous step (after its propagation)")
            body += ["    " + x for x in self._block("rs", self.rs, {})]
            body.append("}")
        if self.statemon and (not self.flush or self.statemon_when == "end"):
            # when='start' records step s before its update; when='end' records the state
            # after step s-1's propagation and reset, i.e. the same point in the loop
            nv = len(self.statemon)
            row = "step - p.rec_row0" if self.statemon_when == "start" else "step - 1 - p.rec_row0"
            body += [
                f"{{ // StateMonitor (when='{self.statemon_when}')",
                "    int lo = 0, hi = p.n_rec;",
                "    while (lo < hi) { const int m = (lo + hi) >> 1; if (p.rec_idx[m] < i) lo = m + 1; else hi = m; }",
                f"    const long long row = {row};",
                "    if (lo < p.n_rec && p.rec_idx[lo] == i && row >= 0 && row < p.rec_rows) {",
            ]
            for a, v in enumerate(self.statemon):
                body.append(f"        p.rec_buf[((size_t)row * {nv} + {a}) * p.n_rec + lo] = (double){v};")
            body += ["    }", "}"]
        if not self.flush:
            if self.su is not None:
                body.append("{ // stateupdate")
                body += ["    " + x for x in self._block("su", self.su, {})]
                body.append("}")
            if self.th is not None:
                body.append("{ // threshold")
                body += ["    " + x for x in self._block("th", self.th, {"_cond": "bool"})]
                body.append("    hb_cond = _cond;")
                body.append("}")
                if self.uses_refractory:
                    body.append("if (hb_cond) { not_refractory = false; lastspike = t; }")
            if self.reset_mode == "fused":
                body.append("if (hb_cond) { // reset")
                body += ["    " + x for x in self._block("rs", self.rs, {})]
                body.append("}")
            if self.reset_mode == "deferred":
                body.append("p.spiked[i] = hb_cond ? 1 : 0;")
            if self.spikemon:
                body.append("if (hb_cond && i >= p.mon_start && i < p.mon_stop) p.mon_count[i - p.mon_start] += 1;")
        elif self.reset_mode == "deferred":
            body.append("p.spiked[i] = 0;")
        for v, ct in self.arrays.items():
            body.append(f"p.a_{v}[i] = {v} ? 1 : 0;" if ct == "bool" else f"p.a_{v}[i] = {v};")
        k += ["        " + x for x in body]
        k.append("    }")
        if self.flush:
            k.append("}")
            return "\n".join(k) + "\n"
        # ---- spikes of this step -> ring slot (ballot + workgroup prefix + one atomic)
        k += [
            "    {",
            "        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;",
            "        const int slot = (int)(step % p.ring_slots);",
            "        const unsigned long long bal = __ballot(hb_cond);",
            "        const int rnk = __popcll(bal & ((1ull << lane) - 1ull));",
            f"        __shared__ int s_w[{nw + 1}];",
            "        __shared__ int s_base;",
            "        if (lane == 0) s_w[wid] = __popcll(bal);",
        ]
        if self.spikemon:
            k += [
                "        const bool hb_m = hb_cond && i >= p.mon_start && i < p.mon_stop;",
                "        const unsigned long long balm = __ballot(hb_m);",
                "        const int rnkm = __popcll(balm & ((1ull << lane) - 1ull));",
                f"        __shared__ int s_wm[{nw + 1}];",
                "        __shared__ unsigned long long s_mbase;",
                "        if (lane == 0) s_wm[wid] = __popcll(balm);",
            ]
        k += [
            "        __syncthreads();",
            "        if (threadIdx.x == 0) {",
            "            int acc = 0;",
            f"            for (int w = 0; w < {nw}; w++) {{ const int c = s_w[w]; s_w[w] = acc; acc += c; }}",
            "            s_base = acc ? atomicAdd(&p.ring_cnt[slot], acc) : 0;",
        ]
        if self.spikemon:
            k += [
                "            int accm = 0;",
                f"            for (int w = 0; w < {nw}; w++) {{ const int c = s_wm[w]; s_wm[w] = accm; accm += c; }}",
                "            s_mbase = accm ? atomicAdd(p.mon_n, (unsigned long long)accm) : 0ull;",
                "            if (accm && s_mbase + (unsigned long long)accm > (unsigned long long)p.mon_cap) atomicOr(p.status, 2);",
            ]
        k += [
            "        }",
            "        __syncthreads();",
            "        if (hb_cond) p.ring_spk[(size_t)slot * p.ring_stride + s_base + s_w[wid] + rnk] = i;",
        ]
        if self.spikemon:
            k += [
                "        if (hb_m) {",
                "            const unsigned long long q = s_mbase + (unsigned long long)(s_wm[wid] + rnkm);",
                "            if (q < (unsigned long long)p.mon_cap) { p.mon_i[q] = i; p.mon_step[q] = (int)step; }",
                "        }",
            ]
        k += [
            "        if (blockIdx.x == 0 && threadIdx.x == 0) p.ring_cnt[(int)((step + 1) % p.ring_slots)] = 0;",
            "    }",
            "}",
        ]
        return "\n".join(k) + "\n"


# ---------------------------------------------------------------------------
class GatherKernel:
    """Per-pathway event gather: spikes of step-d -> CSR segments of delay d -> buckets."""

    def __init__(self, name, rank, n_seg, dmin, on_pre, syn_arrays, scalars, val_t, bucket_cap, float_t,
                 target_var, count_synapses=False):
        if rank >= MAX_RANK:
            raise NotImplementedError(f"more than {MAX_RANK} pathways into one group")
        if n_seg > MAX_SEG:
            raise NotImplementedError(f"delay range of more than {MAX_SEG} steps")
        self.name = name
        self.rank = rank
        self.n_seg = n_seg
        self.dmin = dmin
        self.on_pre = on_pre          # list of statement dicts, last one writes the target
        self.syn_arrays = syn_arrays  # {name: ctype}
        self.scalars = scalars        # {name: (value, ctype)}
        self.val_t = val_t
        self.cap = bucket_cap
        self.float_t = float_t
        self.target_var = target_var
        self.count_synapses = count_synapses
        self.kernel_name = f"gather_{name}"
        self._build()

    def _build(self):
        ps = ParamStruct(f"hbg_{self.name}")
        ps.ptr("clock", "long long", const=True)
        ps.val("step_off", "long long")
        ps.ptr("ring_spk", "int", const=True)
        ps.ptr("ring_cnt", "int", const=True)
        ps.val("ring_stride", "long long")
        ps.val("ring_slots", "int")
        ps.val("src0", "int")
        ps.val("src1", "int")
        ps.val("pad0", "int")
        ps.ptr("seg_ptr", "unsigned int", const=True)
        ps.ptr("post", "int", const=True)
        for v, ct in self.syn_arrays.items():
            ps.ptr(f"s_{v}", ct, const=True)
        for n, (_, ct) in sorted(self.scalars.items()):
            ps.val(f"c_{n}", ct)
        ps.ptr("bkt_cnt", "int")
        ps.ptr("bkt_key", "unsigned long long")
        ps.ptr("bkt_val", self.val_t)
        ps.val("bkt_cap", "int")
        ps.val("pad_bkt", "int")
        ps.ptr("status", "int")
        ps.ptr("events", "unsigned long long")  # per-workgroup delivered-event totals
        if self.count_synapses:
            ps.ptr("syn_count", "int")           # per-synapse event counts (CSR order)
        self.struct = ps
        T = T_TILE
        NS = self.n_seg
        cap = self.cap
        # value expression: statements before the target write are temporaries
        types = dict(self.syn_arrays)
        types["_postsynaptic_idx"] = "int"
        types["_idx"] = "int"
        for n, (_, ct) in self.scalars.items():
            types[n] = ct
        names = {n: f"p.c_{n}" for n in self.scalars}
        lines = []
        declared = set()
        from .codegen import translate_block as tb

        body = self.on_pre[:-1]
        lines += tb(body, {}, types, declared, self.float_t)
        R = Renderer(types)
        c, _ = R.render(self.on_pre[-1]["expr"])
        lines.append(f"const {self.val_t} hb_val = ({self.val_t})({c});")
        import re

        if names:
            pat = re.compile(r"(?<![\w.])(" + "|".join(re.escape(k) for k in sorted(names, key=len, reverse=True)) + r")(?!\w)")
            lines = [pat.sub(lambda m: names[m.group(1)], ln) for ln in lines]
        loads = [f"const {ct} {v} = p.s_{v}[e];" for v, ct in self.syn_arrays.items()]
        k = [source_header(T), ps.c_decl()]
        k.append(f'extern "C" __global__ void __launch_bounds__({T}) {self.kernel_name}(const {ps.name} p) {{')
        k.append("    const long long step = p.clock[0] + p.step_off;")
        k.append(f"    __shared__ int s_cum[{NS + 1}];")
        k.append(f"    for (int s = threadIdx.x; s < {NS}; s += {T}) {{")
        k.append(f"        const long long ps_ = step - {self.dmin} - s;")
        k.append("        s_cum[s + 1] = ps_ >= 0 ? p.ring_cnt[(int)(ps_ % p.ring_slots)] : 0;")
        k.append("    }")
        k.append("    if (threadIdx.x == 0) s_cum[0] = 0;")
        k.append("    __syncthreads();")
        k.append(f"    if (threadIdx.x == 0) for (int s = 1; s <= {NS}; s++) s_cum[s] += s_cum[s - 1];")
        k.append("    __syncthreads();")
        k.append(f"    const int total = s_cum[{NS}];")
        k.append("    __shared__ unsigned long long s_ev;")
        k.append("    if (threadIdx.x == 0) s_ev = 0ull;")
        k.append("    __syncthreads();")
        k.append("    unsigned long long hb_ev = 0ull;")
        k.append("    const int lane = threadIdx.x & 63;")
        k.append("    const int nwaves = (gridDim.x * blockDim.x) >> 6;")
        k.append("    for (int q = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < total; q += nwaves) {")
        k.append(f"        int lo = 0, hi = {NS};")
        k.append("        while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (s_cum[m] <= q) lo = m; else hi = m; }")
        k.append("        const int seg = lo;")
        k.append(f"        const long long sstep = step - {self.dmin} - seg;")
        k.append("        const int slot = (int)(sstep % p.ring_slots);")
        k.append("        const int src = p.ring_spk[(size_t)slot * p.ring_stride + (q - s_cum[seg])];")
        k.append("        if (src < p.src0 || src >= p.src1) continue;")
        k.append(f"        const size_t row = (size_t)(src - p.src0) * {NS} + seg;")
        k.append("        const unsigned int a = p.seg_ptr[row], b = p.seg_ptr[row + 1];")
        k.append("        hb_ev += (unsigned long long)(b - a);")
        k.append(f"        const unsigned long long khi = ((unsigned long long){self.rank} << 44) | "
                 f"((unsigned long long)({NS - 1} - seg) << 32);")
        k.append("        for (unsigned int e = a + lane; e < b; e += 64) {")
        k.append("            const int _postsynaptic_idx = p.post[e];")
        k.append("            const int _idx = (int)e;")
        k.extend("            " + x for x in loads)
        k.extend("            " + x for x in lines)
        k.append(f"            const int tile = _postsynaptic_idx / {T};")
        k.append(f"            const unsigned long long key = ((unsigned long long)(_postsynaptic_idx - tile * {T}) << {KEY_TGT_SHIFT}) | khi | (unsigned long long)e;")
        if self.count_synapses:
            k.append("            p.syn_count[e] += 1;")
        k.append("            const int pos = atomicAdd(&p.bkt_cnt[tile], 1);")
        k.append("            if (pos < p.bkt_cap) {")
        k.append("                p.bkt_key[(size_t)tile * (size_t)p.bkt_cap + pos] = key;")
        k.append("                p.bkt_val[(size_t)tile * (size_t)p.bkt_cap + pos] = hb_val;")
        k.append("            } else {")
        k.append("                atomicOr(p.status, 1);")
        k.append("            }")
        k.append("        }")
        k.append("    }")
        k.append("    if (lane == 0 && hb_ev) atomicAdd(&s_ev, hb_ev);")
        k.append("    __syncthreads();")
        k.append("    if (threadIdx.x == 0) p.events[blockIdx.x] += s_ev;")
        k.append("}")
        self.source = "\n".join(k) + "\n"


def struct_offset(struct_cls, field):
    return getattr(struct_cls, field).offset


def fill_struct(struct_cls, values):
    s = struct_cls()
    for name, _ in struct_cls._fields_:
        if name in values:
            v = values[name]
            setattr(s, name, v)
    return s


def struct_bytes(s):
    return ctypes.string_at(ctypes.addressof(s), ctypes.sizeof(s))


__all__ = ["CodeSpec", "NeuronKernel", "GatherKernel", "scalar_values", "T_TILE", "c_literal"]
