"""Kernel sources generated per group and per synaptic pathway.

Step semantics follow Brian's schedule (brian2/core/network.py:42-50,
1194-1250): stateupdate -> threshold -> spike monitor -> [push, propagate]
per pathway -> reset -> clock.advance.  On the device one step is

  neuron kernel (per group)   fold of the events delivered by propagate(step-1)
                              [deferred reset of step-1] [StateMonitor]
                              stateupdate; threshold; [fused reset];
                              spike compaction into the ring slot of `step`;
                              [SpikeMonitor]
  gather kernel (per pathway) events of `step`: spikes of step-d whose synapses
                              have delay d, d in [dmin, dmax] (the spike queue as
                              a ring of spike lists indexed by integer delay)

Event staging.  The gather works on units of (delay, spike) pairs.  A workgroup
flattens a unit's synapse segments into events and handles them in windows held
in registers: counted per target tile in LDS, space taken in each touched tile's
region of the staging buffer with one global atomic per tile, and the window's
events written there, sorted by tile.  No event is appended with its own atomic.

Fold.  The neuron kernel of a tile reads its region (contiguous), counting-sorts
the events by target in LDS, ranks every event among its target's by key and
applies them one by one in that order.  The key (pathway rank, dmax - delay, CSR
position) is ordered exactly like (pathway in schedule order, push step, source
index, synapse index), i.e. like the cython runtime's sequential left fold over
the spike queue (brian2/codegen/runtime/cython_rt/templates/synapses.pyx:42-50).
No floating-point atomics are used, so the sums are bit-identical to Brian's.
"""
import re

from .codegen import (
    ParamStruct,
    Renderer,
    c_literal,
    ctype_of,
    source_header,
    translate_block,
    translate_statement,
)

# event key: [63:48] target in tile, [47:44] pathway rank, [43:32] push-order field, [31:0] CSR
# position.  The push-order field is n_seg - (d - dmin) for an event of delay d (earlier pushes
# first); 0 is kept for events carried over a change of dt (they precede every later push)
KEY_TGT_SHIFT = 48
MAX_RANK = 16
MAX_SEG = 4095
BIG_LISTED = 64  # targets per tile and step ordered cooperatively when their events exceed cap/2
FOLD_LOADS_COUNT = 8  # fold: loads in flight per thread when counting the events beyond the registers
FOLD_LOADS_PASS = 3   # fold: loads in flight per thread when a pass reads them again
FOLD_RANK_UNROLL = 3  # fold: keys of a segment read per LDS round trip when ranking an event
FOLD_APPLY_BATCH = 3  # fold: events of a target read together before they are applied in order
GATHER_BLOCK = 1024
DECL_MARK = "// ---- generated declarations ----"  # the per-kernel part of a source starts here


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


def _rename(lines, names):
    """Whole-identifier substitution of scalar names by their struct fields."""
    if not names:
        return lines
    pat = re.compile(r"(?<![\w.])(" + "|".join(re.escape(k) for k in sorted(names, key=len, reverse=True))
                     + r")(?!\w)")
    return [pat.sub(lambda m: names[m.group(1)], ln) for ln in lines]


FOLD_EVENTS_PER_THREAD = 1    # fold: events a thread keeps in registers
GATHER_EVENTS_PER_THREAD = 8  # gather: events a thread keeps in registers per window
GATHER_STAGE_RUNS = True      # gather: sort a window's events by tile in LDS before writing them
LDS_BYTES = 163840            # LDS of one CU


def synapse_record(syn_arrays):
    """Layout of a pathway's synapse records: the target index, the 4-byte variables,
    a pad word if needed, the 8-byte variables.  Returns ([(field, ctype)], 32-bit
    words, alignment); records of 8 or 16 bytes are aligned to their size, so that
    one load reads a record."""
    small = [(n, ct) for n, ct in syn_arrays.items() if ct in ("float", "int")]
    big = [(n, ct) for n, ct in syn_arrays.items() if ct not in ("float", "int")]
    fields = [("post", "int")] + small
    words = len(fields)
    if big:
        if words % 2:
            fields.append(("pad_", "int"))
            words += 1
        fields += big
        words += 2 * len(big)
    align = 4 * words if words in (2, 4) else (8 if big else 4)
    return fields, words, align


def rec_type(val_t):
    """Staged event record of a value type: 12 bytes for 4-byte values, 16 for 8-byte."""
    return "hb_rec8" if val_t in ("double", "long long") else "hb_rec4"


def rec_words(val_t):
    return 4 if val_t in ("double", "long long") else 3


def fold_lds_bytes(T, cap, val_t, npt=1):
    """LDS of a step kernel's fold: key, value and order slot per event, first
    event and count per target, and a little for scans and flags."""
    vbytes = 8 if val_t in ("double", "long long") else 4
    return cap * (8 + vbytes + 2) + 8 * T * npt + 1024


def fold_capacity(T, val_t, npt=1, wgs_per_cu=None):
    """Events of one tile sorted in LDS at a time: the workgroups that share a CU
    (by default max(1, 1024 // T), i.e. 16 waves of up to 128 VGPRs) get equal
    shares of its LDS."""
    vbytes = 8 if val_t in ("double", "long long") else 4
    per_wg = LDS_BYTES // (wgs_per_cu or max(1, 1024 // T))
    cap = (per_wg - 8 * T * npt - 1024) // (8 + vbytes + 2)
    return int(max(128, cap - cap % 64))


def _fold_source(T, cap, vt, flush, npt=1):
    """Fold of the events staged by the previous step's gathers, for a tile of
    T*npt targets (thread x owns targets q*T + x, q < npt).

    The tile's events are contiguous in its region of the staging buffer (in the
    order the gathers' atomics gave them; the keys fix the delivery order).
    Phase A reads them, the first T*U into registers (U per thread, coalesced),
    and counts events per target in LDS.  Phase B computes each target's first
    event (block scan) and assigns the targets to passes: all of them at once
    when the tile's events fit in LDS, else by the position of their first event
    in windows of cap/2 (a pass then covers at most cap events).  In each pass
    the events of its targets are scattered into LDS and ordered by key, and the
    owner of each target applies them to the target's registers.  A target with
    more than cap/2 events in one step (an in-degree above cap/2) is sorted
    through an index list in global memory.  Events beyond the registers are
    read again by each pass."""
    U = FOLD_EVENTS_PER_THREAD
    XW = T * U
    W = cap // 2
    TN = T * npt
    k = [
        "    // ---- events delivered by propagate(step - 1): this tile's region of the staging ----",
        "    // staging buffers rotate over three steps: step s's gathers write buffer s % 3 (possibly",
        "    // while this kernel runs), this kernel reads buffer (s - 1) % 3 and clears the tile's",
        "    // counter of buffer (s + 1) % 3",
        "    const int hb_nxt = (int)((step + 1) % 3), hb_prv = (int)((step + 2) % 3);",
        "    (void)hb_nxt;",
        "    const unsigned int hb_rbase = p.tile_off[hb_tile];",
        f"    const {rec_type(vt)}* hb_st = (const {rec_type(vt)}*)p.stg_rec + (size_t)hb_prv * p.stg_stride + hb_rbase;",
        f"    __shared__ unsigned long long s_key[{cap}];",
        f"    __shared__ {vt} s_val[{cap}];",
        f"    __shared__ unsigned short s_x[{cap}];",
        f"    __shared__ int s_beg[{TN}];",
        f"    __shared__ int s_fill[{TN}];",
        "    __shared__ int s_cmax, s_alloc;",
        "    // events of the tile: the gathers' count, clamped to the region (a count beyond it",
        "    // was reported by the gather as an overflow)",
        "    int hb_tot;",
        "    {",
        "        const unsigned int c_ = p.tile_cnt[(size_t)hb_prv * p.n_tiles + hb_tile];",
        "        const unsigned int r_ = p.tile_off[hb_tile + 1] - hb_rbase;",
        "        const unsigned int n_ = c_ < r_ ? c_ : r_;",
        "        hb_tot = (int)(n_ < 0x7fffffffu ? n_ : 0x7fffffffu);",
        "    }",
    ]
    if not flush:
        k += ["    // the tile's counter of the next step's gathers",
              "    if (threadIdx.x == 0) p.tile_cnt[(size_t)hb_nxt * p.n_tiles + hb_tile] = 0u;"]
    pro = k
    k = [
        f"    for (int t_ = threadIdx.x; t_ < {TN}; t_ += {T}) s_fill[t_] = 0;",
        "    if (threadIdx.x == 0) { s_cmax = 0; s_alloc = 0; }",
        f"    unsigned long long hb_fk[{U}];",
        f"    {vt} hb_fv[{U}];",
        "    // phase A: events per target of this tile; event g in slot g / T of thread g % T.  The",
        "    // loads are issued before the barrier, so that they and the state loads complete together",
        "    #pragma unroll",
        f"    for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
        f"        const int g = hb_u * {T} + (int)threadIdx.x;",
        "        // loads under a condition are each waited for inside their branch: the index is",
        "        // clamped instead (hb_tot > 0 here), and only the use is conditional",
        f"        const auto r_ = hb_st[g < hb_tot ? g : hb_tot - 1]; hb_fk[hb_u] = hb_rkey(r_); hb_fv[hb_u] = hb_rval<{vt}>(r_);",
        "    }",
        "    __syncthreads();",
        "    #pragma unroll",
        f"    for (int hb_u = 0; hb_u < {U}; hb_u++)",
        f"        if (hb_u * {T} + (int)threadIdx.x < hb_tot) atomicAdd(&s_fill[(int)(hb_fk[hb_u] >> {KEY_TGT_SHIFT})], 1);",
        "    // events beyond the registers: counted here, read again by the passes; FOLD_LOADS_COUNT",
        "    // loads in flight per thread (a load waited for before the next is issued leaves the",
        "    // loop bound by the memory latency)",
        f"    for (int g0 = {XW} + (int)threadIdx.x; g0 < hb_tot; g0 += {FOLD_LOADS_COUNT * T}) {{",
        f"        unsigned int hb_kh[{FOLD_LOADS_COUNT}];",
        "        #pragma unroll",
        f"        for (int j = 0; j < {FOLD_LOADS_COUNT}; j++) {{",
        f"            const int g = g0 + j * {T};",
        "            hb_kh[j] = (unsigned int)(hb_rkey(hb_st[g < hb_tot ? g : hb_tot - 1]) >> 32);  // clamped: see phase A",
        "        }",
        "        #pragma unroll",
        f"        for (int j = 0; j < {FOLD_LOADS_COUNT}; j++)",
        f"            if (g0 + j * {T} < hb_tot) atomicAdd(&s_fill[(int)(hb_kh[j] >> {KEY_TGT_SHIFT - 32})], 1);",
        "    }",
        "    __syncthreads();",
        "    // phase B: first event of every target (segments in (thread, q) order); passes",
        f"    int hb_c[{npt}], hb_b[{npt}], hb_myp[{npt}];",
        f"    bool hb_bigt[{npt}];",
        "    int hb_ct = 0;",
        "    #pragma unroll",
        f"    for (int q = 0; q < {npt}; q++) {{ hb_c[q] = s_fill[q * {T} + threadIdx.x]; hb_b[q] = hb_ct; hb_ct += hb_c[q]; }}",
        "    // segments in LDS: an allocation, waves in any order (the passes need no particular order)",
        "    const int hb_tb = hb_block_alloc(hb_ct, &s_alloc);",
        f"    const bool hb_one = hb_tot <= {cap};",
        "    bool hb_anyb = false;",
        "    int hb_cm = 0;",
        "    #pragma unroll",
        f"    for (int q = 0; q < {npt}; q++) {{",
        "        hb_b[q] += hb_tb;",
        f"        hb_bigt[q] = !hb_one && hb_c[q] > {W};",
        f"        s_beg[q * {T} + threadIdx.x] = hb_bigt[q] ? ~hb_b[q] : hb_b[q];",
        f"        s_fill[q * {T} + threadIdx.x] = 0;",
        "        hb_anyb = hb_anyb || hb_bigt[q];",
        "        if (!hb_bigt[q] && hb_c[q] > hb_cm) hb_cm = hb_c[q];",
        "    }",
        "    if (!hb_one && hb_cm > 0) atomicMax(&s_cmax, hb_cm);",
        "    const bool hb_anybig = __syncthreads_or(hb_anyb) != 0;",
        "    // passes: targets by the window of hb_pw positions their first event falls in; a pass",
        "    // then spans at most hb_pw + (largest count of a target sorted in LDS) = cap events",
        f"    const int hb_pw = hb_one ? {cap} : {cap} - s_cmax;",
        "    const int hb_np = hb_one ? (hb_tot > 0 ? 1 : 0) : (hb_tot + hb_pw - 1) / hb_pw;",
        "    #pragma unroll",
        f"    for (int q = 0; q < {npt}; q++) hb_myp[q] = hb_one ? 0 : hb_b[q] / hb_pw;",
        "    // bursts beyond the tile's index list (only when memory capped it) are reported, not folded",
        "    const bool hb_ovf = hb_anybig && (long long)hb_tot > p.tile_cap;",
        "    if (hb_ovf && threadIdx.x == 0) atomicOr(p.status, 4);",
        "    unsigned int* hb_go = p.stg_ord + (size_t)hb_tile * (size_t)p.tile_cap;",
        "    // event (kk, vv) at position pos of the region: into LDS if its target is in this pass,",
        "    // into the target's index list if the target is sorted in global memory",
        "#define HB_SCATTER(kk, vv, pos) { \\",
        f"        const int t_ = (int)((kk) >> {KEY_TGT_SHIFT}); \\",
        "        const int bt_ = s_beg[t_]; \\",
        "        if (bt_ >= 0) { \\",
        "            if (hb_one || (unsigned int)(bt_ - hk_pb) < (unsigned int)hb_pw) { \\",
        "                const int q_ = bt_ + atomicAdd(&s_fill[t_], 1) - hk_pb; \\",
        "                s_key[q_] = (kk); s_val[q_] = (vv); } \\",
        "        } else if (hk_pass == 0 && !hb_ovf) hb_go[~bt_ + atomicAdd(&s_fill[t_], 1)] = (pos); }",
        "    // the same with the target's segment start already read (bt_)",
        "#define HB_SCATTER_B(kk, vv, pos, bt_) { \\",
        f"        const int t_ = (int)((kk) >> {KEY_TGT_SHIFT}); \\",
        "        if ((bt_) >= 0) { \\",
        "            if (hb_one || (unsigned int)((bt_) - hk_pb) < (unsigned int)hb_pw) { \\",
        "                const int q_ = (bt_) + atomicAdd(&s_fill[t_], 1) - hk_pb; \\",
        "                s_key[q_] = (kk); s_val[q_] = (vv); } \\",
        "        } else if (hk_pass == 0 && !hb_ovf) hb_go[~(bt_) + atomicAdd(&s_fill[t_], 1)] = (pos); }",
    ]
    return pro, k, U, W


def _fold_passes(T, U, W, cap, vt, apply_lines, npt=1):
    """Phase C of the fold (see _fold_source).  `apply_lines` apply (hb_k, hb_v) to
    target q's registers.

    Per pass: the events of the pass's targets are scattered into LDS by target
    (from registers, then the events beyond them from global memory); each event
    finds its rank among its target's (event-parallel rank counting, keys are
    distinct) and the order goes to s_x; the owner of each target applies its
    events in key order.  With several passes, LDS positions that no event of
    the pass wrote hold stale keys: an entry counts only if it lies inside the
    segment of a target of the pass, which the pass wrote in full.  Targets
    sorted in global memory come last."""
    XW = T * U
    TN = T * npt
    RU, AB = FOLD_RANK_UNROLL, FOLD_APPLY_BATCH
    k = [
        f"    const int hb_lim = hb_one ? hb_tot : {cap};  // LDS positions a pass may have written",
        "    for (int hk_pass = 0; hk_pass < hb_np; hk_pass++) {",
        "        const int hk_pb = hk_pass * hb_pw;",
        "        #pragma unroll",
        f"        for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
        f"            const int g = hb_u * {T} + (int)threadIdx.x;",
        "            if (g < hb_tot) HB_SCATTER(hb_fk[hb_u], hb_fv[hb_u], (unsigned int)g)",
        "        }",
        f"        for (int g0 = {XW} + (int)threadIdx.x; g0 < hb_tot; g0 += {FOLD_LOADS_PASS * T}) {{",
        f"            {rec_type(vt)} hb_rr[{FOLD_LOADS_PASS}];",
        "            #pragma unroll",
        f"            for (int j = 0; j < {FOLD_LOADS_PASS}; j++)",
        f"                hb_rr[j] = hb_st[g0 + j * {T} < hb_tot ? g0 + j * {T} : hb_tot - 1];  // clamped: see phase A",
        "            // the records' segment starts first (one LDS round trip for the group)",
        f"            int hb_bt[{FOLD_LOADS_PASS}];",
        "            #pragma unroll",
        f"            for (int j = 0; j < {FOLD_LOADS_PASS}; j++) hb_bt[j] = s_beg[(int)(hb_rkey(hb_rr[j]) >> {KEY_TGT_SHIFT})];",
        "            #pragma unroll",
        f"            for (int j = 0; j < {FOLD_LOADS_PASS}; j++) {{",
        f"                const int g = g0 + j * {T};",
        f"                if (g < hb_tot) HB_SCATTER_B(hb_rkey(hb_rr[j]), hb_rval<{vt}>(hb_rr[j]), (unsigned int)g, hb_bt[j])",
        "            }",
        "        }",
        "        __syncthreads();",
        f"        for (int x = threadIdx.x; x < hb_lim; x += {T}) {{",
        "            const unsigned long long kx = s_key[x];",
        f"            int t_ = (int)(kx >> {KEY_TGT_SHIFT});",
        f"            t_ = t_ < {TN} ? t_ : {TN - 1};",
        "            const int bt_ = s_beg[t_], b_ = bt_ - hk_pb, c_ = s_fill[t_], e_ = b_ + c_ - 1;",
        "            if (hb_one || (bt_ >= 0 && (unsigned int)b_ < (unsigned int)hb_pw && x >= b_ && x <= e_)) {",
        f"                // smaller keys of the segment ({RU} per LDS round trip)",
        "                int r = 0;",
        "                if (c_ > 1) {",
        "                    int j = b_;",
    ]
    if RU > 1:
        k += [
            f"                    for (; j + {RU - 1} <= e_; j += {RU}) {{",
            "                        " + " ".join(f"const unsigned long long k{i}_ = s_key[j + {i}];" for i in range(RU)),
            "                        r += " + " + ".join(f"(k{i}_ < kx ? 1 : 0)" for i in range(RU)) + ";",
            "                    }",
        ]
    k += [
        "                    for (; j <= e_; j++) r += s_key[j] < kx ? 1 : 0;",
        "                }",
        "                s_x[b_ + r] = (unsigned short)(x - b_);",
        "            }",
        "        }",
        "        __syncthreads();",
        "        // apply: the owner of each target applies its events in key order, FOLD_APPLY_BATCH",
        "        // at a time (their order indices, then their keys and values read together)",
        "        #pragma unroll",
        f"        for (int q = 0; q < {npt}; q++) {{",
        "            if (!hb_bigt[q] && hb_myp[q] == hk_pass) {",
        "                const int o = hb_b[q] - hk_pb, c = hb_c[q];",
        "                int r = 0;",
        "                // order slots one batch ahead (read at clamped positions: past the segment",
        "                // they are not used)",
        "                " + " ".join(f"int n{i}_ = (int)s_x[o + {i} < {cap - 1} ? o + {i} : {cap - 1}];" for i in range(AB)),
        f"                for (; r + {AB} <= c; r += {AB}) {{",
        "                    " + " ".join(f"const int a{i}_ = o + n{i}_;" for i in range(AB)),
        "                    " + " ".join(f"n{i}_ = (int)s_x[o + r + {AB + i} < {cap - 1} ? o + r + {AB + i} : {cap - 1}];" for i in range(AB)),
        "                    " + " ".join(f"const unsigned long long k{i}_ = s_key[a{i}_];" for i in range(AB)),
        "                    " + " ".join(f"const {vt} v{i}_ = s_val[a{i}_];" for i in range(AB)),
    ]
    for i in range(AB):
        k.append(f"                    {{ const unsigned long long hb_k = k{i}_; const {vt} hb_v = v{i}_;")
        k += ["                        " + x for x in apply_lines]
        k.append("                    }")
    k += [
        "                }",
        "                for (; r < c; r++) {",
        "                    const int a_ = o + (int)s_x[o + r];",
        "                    const unsigned long long hb_k = s_key[a_];",
        f"                    const {vt} hb_v = s_val[a_];",
    ]
    k += ["                    " + x for x in apply_lines]
    k += [
        "                }",
        "            }",
        "        }",
        "        // the next pass overwrites what this one's owners read (after the last pass the",
        "        // LDS is next written behind a barrier)",
        "        if (hk_pass + 1 < hb_np) __syncthreads();",
        "    }",
        "    // targets sorted in global memory (more than cap/2 events): the workgroup orders each",
        "    // one's index list by counting, for every event, the smaller keys of the target (keys",
        "    // staged in LDS chunks of cap), writing the ordered list to the tile's second list; the",
        "    // owner then applies it.  Up to BIG_LISTED targets per tile take this path, further",
        "    // ones are sorted by their owner alone.",
        "    if (hb_anybig && !hb_ovf) {",
        f"        int* s_bigb = s_beg; int* s_bigc = s_fill;  // free after the passes",
        "        if (threadIdx.x == 0) s_alloc = 0;",
        "        __syncthreads();",
        f"        bool hb_lst[{npt}];",
        "        #pragma unroll",
        f"        for (int q = 0; q < {npt}; q++) {{",
        "            hb_lst[q] = false;",
        "            if (hb_bigt[q]) {",
        "                const int j = atomicAdd(&s_alloc, 1);",
        f"                if (j < {min(BIG_LISTED, TN)}) {{ s_bigb[j] = hb_b[q]; s_bigc[j] = hb_c[q]; hb_lst[q] = true; }}",
        "            }",
        "        }",
        "        __syncthreads();",
        f"        const int hb_nbig = s_alloc < {min(BIG_LISTED, TN)} ? s_alloc : {min(BIG_LISTED, TN)};",
        "        unsigned int* hb_go2 = p.stg_ord + ((size_t)p.n_tiles + (size_t)hb_tile) * (size_t)p.tile_cap;",
        "        for (int j = 0; j < hb_nbig; j++) {",
        "            const int b0 = s_bigb[j], c0 = s_bigc[j];",
        f"            for (int x0 = 0; x0 < c0; x0 += {T}) {{",
        "                const int x = x0 + (int)threadIdx.x;",
        "                unsigned int ixx = 0u;",
        "                unsigned long long kx = 0ull;",
        "                if (x < c0) { ixx = hb_go[b0 + x]; kx = hb_rkey(hb_st[ixx]); }",
        "                int r = 0;",
        f"                for (int y0 = 0; y0 < c0; y0 += {cap}) {{",
        f"                    const int ny = c0 - y0 < {cap} ? c0 - y0 : {cap};",
        "                    __syncthreads();",
        f"                    for (int y = threadIdx.x; y < ny; y += {T}) s_key[y] = hb_rkey(hb_st[hb_go[b0 + y0 + y]]);",
        "                    __syncthreads();",
        "                    if (x < c0)",
        "                        for (int y = 0; y < ny; y++) r += s_key[y] < kx ? 1 : 0;",
        "                }",
        "                if (x < c0) hb_go2[b0 + r] = ixx;",
        "            }",
        "        }",
        "        __syncthreads();",
        "        #pragma unroll",
        f"        for (int q = 0; q < {npt}; q++) {{",
        "            if (!hb_bigt[q]) continue;",
        "            const unsigned int* hb_ol = hb_go2;",
        "            if (!hb_lst[q]) { hb_sort_idx(hb_go + hb_b[q], hb_st, hb_c[q]); hb_ol = hb_go; }",
        "            for (int e = hb_b[q]; e < hb_b[q] + hb_c[q]; e++) {",
        "                const unsigned int r = hb_ol[e];",
        "                const unsigned long long hb_k = hb_rkey(hb_st[r]);",
        f"                const {vt} hb_v = hb_rval<{vt}>(hb_st[r]);",
    ]
    k += ["                " + x for x in apply_lines]
    k += [
        "            }",
        "        }",
        "    }",
        "#undef HB_SCATTER",
        "#undef HB_SCATTER_B",
    ]
    return k


# ---------------------------------------------------------------------------
class NeuronKernel:
    """Fused per-group step kernel (or the end-of-run flush kernel)."""

    def __init__(self, name, N, arrays, float_t, stateupdate=None, threshold=None, reset=None,
                 reset_mode="none", scalars=None, fold=None, lds_cap=0, val_t="double",
                 spikemon=None, statemon=None, flush=False, uses_refractory=False, statemon_when="start",
                 tile=256, npt=1, ratemon=None, waves_per_eu=0, ring_parts=None):
        self.name = name
        self.N = N
        self.arrays = arrays          # {var: ctype} of the group's state arrays
        self.float_t = float_t
        self.su, self.th, self.rs = stateupdate, threshold, reset
        self.reset_mode = reset_mode  # 'fused' | 'deferred' | 'none'
        self.scalars = scalars or {}  # {(tag, name): (value, ctype)}
        self.fold = fold or []        # [(rank, var, op)]
        self.val_t = val_t
        self.cap = fold_capacity(tile, val_t, npt) if not lds_cap else int(lds_cap)
        if self.cap < 128 or self.cap % 2 or fold_lds_bytes(tile, self.cap, val_t, npt) > LDS_BYTES:
            raise ValueError("fold capacity must be an even number of events, at least 128, whose LDS fits a CU")
        self.spikemon = spikemon      # (start, stop) or None
        self.ratemon = None if flush else ratemon  # (start, stop) of a PopulationRateMonitor or None
        self.waves_per_eu = int(waves_per_eu)  # > 0: ask for that many waves per SIMD (at most 512/w VGPRs)
        self.statemon = statemon      # [vars] or None
        self.statemon_when = statemon_when
        self.flush = flush
        self.uses_refractory = uses_refractory
        # starts of the parts of the spike ring (neuron ranges, ascending, the first 0): the
        # spikes of part r go to slot * stride + start_r, counted in ring_cnt[slot * R + r]
        self.ring_parts = tuple(ring_parts or (0,))
        if self.ring_parts[0] != 0 or list(self.ring_parts) != sorted(set(self.ring_parts)) or self.ring_parts[-1] >= max(N, 1):
            raise ValueError("ring parts must be ascending neuron indices starting at 0")
        self.T = tile                 # threads per workgroup
        self.npt = npt                # neurons per thread: a workgroup owns tile * npt neurons
        if npt < 1 or tile * npt > 65536:
            raise ValueError("a tile holds at most 65536 neurons")
        self._build()

    def _block(self, tag, spec, extra_types):
        types = dict(self.arrays)
        types.update({"t": "double", "dt": "double", "N": "int", "i": "int", "_idx": "int"})
        types.update(extra_types)
        names = {}
        for (tg, n), (_, ct) in self.scalars.items():
            if tg == tag:
                types[n] = ct
                names[n] = f"p.{tag}_{n}"
        lines = translate_block(spec.vector, spec.conditional_write, types, set(), self.float_t)
        return _rename(lines, names)

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
            ps.ptr("stg_rec", "unsigned int", const=True)
            ps.ptr("stg_ord", "unsigned int")
            ps.ptr("tile_off", "unsigned int", const=True)
            ps.ptr("tile_cnt", "unsigned int")
            ps.val("n_tiles", "long long")
            ps.val("tile_cap", "long long")
            ps.val("stg_stride", "unsigned long long")
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
        if self.ratemon:
            ps.ptr("rate_cnt", "int")
            ps.val("rate_row0", "long long")
            ps.val("rate_rows", "long long")
            ps.val("rate_start", "int")
            ps.val("rate_stop", "int")
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
        T, NPT = self.T, self.npt
        TN = T * NPT
        k = [source_header(T), DECL_MARK, ps.c_decl()]
        kname = f"{'flush' if self.flush else 'step'}_{self.name}"
        self.kernel_name = kname
        occ = f" __attribute__((amdgpu_waves_per_eu({self.waves_per_eu})))" if self.waves_per_eu else ""
        k.append(f'extern "C" __global__ void __launch_bounds__({T}){occ} {kname}(const {ps.name} p) {{')
        k.append("    const long long step = p.clock[0] + p.step_off;")
        k.append("    const double t = (double)step * p.dt;")
        k.append("    const double dt = p.dt;")
        k.append("    const int N = p.N;")
        k.append("    const int hb_tile = blockIdx.x;")
        k += [
            f"    const int hb_i0 = hb_tile * {TN} + threadIdx.x;  // neuron q*{T} + hb_i0 for q < {NPT}",
        ]
        k.append("    (void)t; (void)dt;")
        if not self.flush:
            # spike counts of the workgroup (cleared here, before the first barrier)
            k += ["    __shared__ int s_nsp, s_nspm, s_nrt;", "    if (threadIdx.x == 0) { s_nsp = 0; s_nspm = 0; s_nrt = 0; }"]
            if len(self.ring_parts) > 1:
                R = len(self.ring_parts)
                k += [f"    __shared__ int s_npart[{R}], s_pbase[{R}];",
                      f"    if (threadIdx.x < {R}) s_npart[threadIdx.x] = 0;"]
        fold = _fold_source(T, self.cap, self.val_t, self.flush, NPT) if self.fold else None
        if fold:
            k += fold[0]  # the tile's event count first: phase A's loads depend on it, not on the state
        # state of this thread's neurons in registers (the fold updates them)
        for v, ct in self.arrays.items():
            k.append(f"    {'bool' if ct == 'bool' else ct} hb_a_{v}[{NPT}];")
        k.append(f"    bool hb_condq[{NPT}];")
        k.append("    #pragma unroll")
        k.append(f"    for (int q = 0; q < {NPT}; q++) {{")
        k.append(f"        const int i = hb_i0 + q * {T};")
        k.append("        const bool act = i < N;")
        for v, ct in self.arrays.items():
            if ct == "bool":
                k.append(f"        hb_a_{v}[q] = act ? p.a_{v}[i] != 0 : false;")
            else:
                k.append(f"        hb_a_{v}[q] = act ? p.a_{v}[i] : ({ct})0;")
        k.append("        hb_condq[q] = false;")
        k.append("    }")
        if self.fold:
            fk, U, W = fold[1], fold[2], fold[3]
            # hb_tot is the same for every thread of the workgroup: a tile without events skips
            # the fold (its barriers included)
            k.append("    if (hb_tot > 0) {")
            k += fk
            apply = ["switch ((int)((hb_k >> 44) & 15)) {"]
            for rank, var, op in self.fold:
                apply.append(f"case {rank}: hb_a_{var}[q] {op} hb_v; break;")
            apply += ["default: break;", "}"]
            k += _fold_passes(T, U, W, self.cap, self.val_t, apply, NPT)
            k += ["    } else {", "    __syncthreads();  // the workgroup's counters, cleared at the start", "    }"]
        if self.statemon and (not self.flush or self.statemon_when == "end"):
            # the recorded neurons of this tile: rec_idx[hb_rlo, hb_rhi), found by two searches
            # whose indices are the same for every thread (scalar loads), so that neurons of
            # tiles without recorded neurons search nothing
            k += [
                "    int hb_rlo, hb_rhi;",
                "    {",
                f"        const int a_ = hb_tile * {TN}, b_ = a_ + {TN};",
                "        int lo = 0, hi = p.n_rec;",
                "        while (lo < hi) { const int m = (lo + hi) >> 1; if (p.rec_idx[m] < a_) lo = m + 1; else hi = m; }",
                "        hb_rlo = __builtin_amdgcn_readfirstlane(lo);",
                "        hi = p.n_rec;",
                "        while (lo < hi) { const int m = (lo + hi) >> 1; if (p.rec_idx[m] < b_) lo = m + 1; else hi = m; }",
                "        hb_rhi = __builtin_amdgcn_readfirstlane(lo);",
                "    }",
            ]
        # ---- per-neuron body ------------------------------------------------------
        k.append("    #pragma unroll")
        k.append(f"    for (int q = 0; q < {NPT}; q++) {{")
        k.append(f"        const int i = hb_i0 + q * {T};")
        k.append("        const int _idx = i;")
        k.append("        (void)_idx;")
        k.append("        if (i < N) {")
        body = []
        for v, ct in self.arrays.items():
            body.append(f"{'bool' if ct == 'bool' else ct} {v} = hb_a_{v}[q];")
        body.append("bool hb_cond = false;")
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
                f"{{ // StateMonitor (when='{self.statemon_when}'): search only the tile's recorded neurons",
                "    int lo = hb_rlo, hi = hb_rhi;",
                "    while (lo < hi) { const int m = (lo + hi) >> 1; if (p.rec_idx[m] < i) lo = m + 1; else hi = m; }",
                f"    const long long row = {row};",
                "    if (lo < hb_rhi && p.rec_idx[lo] == i && row >= 0 && row < p.rec_rows) {",
            ]
            for a_, v in enumerate(self.statemon):
                body.append(f"        p.rec_buf[((size_t)row * {nv} + {a_}) * p.n_rec + lo] = (double){v};")
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
        body.append("hb_condq[q] = hb_cond;")
        k += ["            " + x for x in body]
        k.append("        }")
        k.append("    }")
        if self.flush:
            k.append("}")
            return "\n".join(k) + "\n"
        # ---- spikes of this step -> ring slot (workgroup allocation + one atomic per workgroup)
        k += [
            "    {",
            "        const int slot = (int)(step % p.ring_slots);",
            "        __shared__ int s_base;",
            "        int hb_n = 0;",
            "        #pragma unroll",
            f"        for (int q = 0; q < {NPT}; q++) hb_n += hb_condq[q] ? 1 : 0;",
        ]
        if not self.fold:
            k.append("        __syncthreads();  // s_nsp cleared")
        if self.spikemon:
            k += [
                "        __shared__ unsigned long long s_mbase;",
                "        int hb_nm = 0;",
                "        #pragma unroll",
                f"        for (int q = 0; q < {NPT}; q++) {{",
                f"            const int i = hb_i0 + q * {T};",
                "            hb_nm += (hb_condq[q] && i >= p.mon_start && i < p.mon_stop) ? 1 : 0;",
                "        }",
            ]
        RPARTS = self.ring_parts
        R = len(RPARTS)
        bounds = list(RPARTS) + [self.N]
        if R > 1:
            k += ["        // spikes per part of the ring"]
            for r in range(R):
                k += [
                    f"        int hb_np{r} = 0;",
                    "        #pragma unroll",
                    f"        for (int q = 0; q < {NPT}; q++) {{",
                    f"            const int i = hb_i0 + q * {T};",
                    f"            hb_np{r} += (hb_condq[q] && i >= {bounds[r]} && i < {bounds[r + 1]}) ? 1 : 0;",
                    "        }",
                    f"        int hb_op{r} = hb_block_alloc(hb_np{r}, &s_npart[{r}]);",
                ]
        packed = self.spikemon and TN < 32768 and R == 1
        if R > 1:
            if self.spikemon:
                k.append("        const int hb_offm = hb_block_alloc(hb_nm, &s_nspm);")
        elif packed:
            k += [
                "        // both counts in one allocation (each total is at most the tile size, below 2^15)",
                "        const int hb_po = hb_block_alloc(hb_n | (hb_nm << 16), &s_nsp);",
                "        const int hb_off = hb_po & 0xffff, hb_offm = hb_po >> 16;",
            ]
        else:
            k.append("        const int hb_off = hb_block_alloc(hb_n, &s_nsp);")
            if self.spikemon:
                k.append("        const int hb_offm = hb_block_alloc(hb_nm, &s_nspm);")
        if self.ratemon:
            k += [
                "        // spikes of the rate monitor's range: only their number is needed",
                "        int hb_nr = 0;",
                "        #pragma unroll",
                f"        for (int q = 0; q < {NPT}; q++) {{",
                f"            const int i = hb_i0 + q * {T};",
                "            hb_nr += (hb_condq[q] && i >= p.rate_start && i < p.rate_stop) ? 1 : 0;",
                "        }",
                "        (void)hb_block_alloc(hb_nr, &s_nrt);",
            ]
        k += ["        __syncthreads();", "        if (threadIdx.x == 0) {"]
        if R > 1:
            k += [f"            for (int r = 0; r < {R}; r++) s_pbase[r] = s_npart[r] ? atomicAdd(&p.ring_cnt[slot * {R} + r], s_npart[r]) : 0;"]
        else:
            k += [
                "            const int hb_ntot = s_nsp & 0xffff;" if packed else "            const int hb_ntot = s_nsp;",
                "            s_base = hb_ntot ? atomicAdd(&p.ring_cnt[slot], hb_ntot) : 0;",
            ]
        if self.spikemon:
            k += [
                "            const int hb_nmtot = s_nsp >> 16;" if packed else "            const int hb_nmtot = s_nspm;",
                "            s_mbase = hb_nmtot ? atomicAdd(p.mon_n, (unsigned long long)hb_nmtot) : 0ull;",
                "            if (hb_nmtot && s_mbase + (unsigned long long)hb_nmtot > (unsigned long long)p.mon_cap) "
                "atomicOr(p.status, 2);",
            ]
        if self.ratemon:
            k += [
                "            if (s_nrt && p.rate_cnt && step - p.rate_row0 >= 0 && step - p.rate_row0 < p.rate_rows)",
                "                atomicAdd(&p.rate_cnt[step - p.rate_row0], s_nrt);",
            ]
        k += ["        }", "        __syncthreads();"]
        if R > 1:
            k += [f"        hb_op{r} += s_pbase[{r}];" for r in range(R)]
        else:
            k.append("        int hb_o = s_base + hb_off;")
        if self.spikemon:
            k.append("        unsigned long long hb_om = s_mbase + (unsigned long long)hb_offm;")
        k += [
            "        #pragma unroll",
            f"        for (int q = 0; q < {NPT}; q++) {{",
            "            if (hb_condq[q]) {",
            f"                const int i = hb_i0 + q * {T};",
        ]
        if R > 1:
            for r in range(R):
                k.append(f"                {'else ' if r else ''}if (i < {bounds[r + 1]}) "
                         f"p.ring_spk[(size_t)slot * p.ring_stride + {bounds[r]} + hb_op{r}++] = i;")
        else:
            k += [
                "                p.ring_spk[(size_t)slot * p.ring_stride + hb_o] = i;",
                "                hb_o++;",
            ]
        if self.spikemon:
            k += [
                "                if (i >= p.mon_start && i < p.mon_stop) {",
                "                    if (hb_om < (unsigned long long)p.mon_cap) { p.mon_i[hb_om] = i; p.mon_step[hb_om] = (int)step; }",
                "                    hb_om++;",
                "                }",
            ]
        k += [
            "            }",
            "        }",
            "        if (blockIdx.x == 0 && threadIdx.x == 0) p.ring_cnt[(int)((step + 1) % p.ring_slots)] = 0;" if R == 1 else
            f"        if (blockIdx.x == 0 && threadIdx.x < {R}) p.ring_cnt[(int)((step + 1) % p.ring_slots) * {R} + threadIdx.x] = 0;",
            "    }",
            "}",
        ]
        return "\n".join(k) + "\n"


# ---------------------------------------------------------------------------
def onpre_lines(on_pre, target_index, syn_arrays, scalars, float_t, assign, val_t):
    """The on_pre statements in order (synapses.py:359-383): the postsynaptic update's
    value is assigned with `assign` (the fold applies it later), every other statement
    runs as it is; scalars are read from the parameter struct (p.c_<name>)."""
    types = dict(syn_arrays)
    types["_postsynaptic_idx"] = "int"
    types["_idx"] = "int"
    types["t"] = "double"
    for n, (_, ct) in scalars.items():
        types[n] = ct
    names = {n: f"p.c_{n}" for n in scalars}
    lines = []
    declared = set()
    for i_, st in enumerate(on_pre):
        if i_ == target_index:
            c, _ = Renderer(types).render(st["expr"])
            lines.append(f"{assign}({c});")
        else:
            lines += translate_statement(st, types, declared, float_t)
    return _rename(lines, names)


class GatherKernel:
    """Per-pathway event gather: spikes of step-d -> CSR segments of delay d ->
    the regions of the target tiles in the target group's staging buffer.

    Work is split in units of `pairs` (delay, spike) pairs (one per thread of the
    first `pairs` threads).  Their segments are flattened into events, handled in
    windows of GATHER_BLOCK * GATHER_EVENTS_PER_THREAD events held in registers:
    counted per target tile in LDS, space taken in each touched tile's region
    with one global atomic per tile, and the window's events written there
    (sorted by tile in LDS first when it fits, so that consecutive threads write
    consecutive positions)."""

    def __init__(self, name, rank, n_seg, dmin, on_pre, syn_arrays, scalars, val_t, float_t, target_var,
                 n_tiles, tile, count_synapses=False, pairs=256, target_index=None, plastic=(),
                 events_per_thread=None, ring_parts=None):
        if rank >= MAX_RANK:
            raise NotImplementedError(f"more than {MAX_RANK} pathways into one group")
        if n_seg > MAX_SEG:
            raise NotImplementedError(f"delay range of more than {MAX_SEG} steps")
        if n_tiles > 16384:
            raise NotImplementedError("more than 16384 target tiles")
        self.block = GATHER_BLOCK
        self.events_per_thread = int(events_per_thread or GATHER_EVENTS_PER_THREAD)
        self.window = self.block * self.events_per_thread
        if not 1 <= pairs <= self.block:
            raise ValueError(f"pairs per unit must be in [1, {self.block}]")
        self.name = name
        self.rank = rank
        self.n_seg = n_seg
        self.dmin = dmin
        self.on_pre = on_pre          # list of statement dicts, one of them writes the target
        self.target_index = len(on_pre) - 1 if target_index is None else int(target_index)
        self.plastic = set(plastic)   # synaptic variables the statements write (read and written in place)
        self.syn_arrays = syn_arrays  # {name: ctype}
        self.scalars = scalars        # {name: (value, ctype)}
        self.val_t = val_t
        self.float_t = float_t
        self.target_var = target_var
        self.n_tiles = n_tiles
        self.tile = tile
        self.pairs = pairs
        self.count_synapses = count_synapses
        # the source's spike ring split in parts (n_parts, first part read, starts of the parts
        # read): the spikes of part r of a slot are at slot * stride + start_r, counted in
        # ring_cnt[slot * n_parts + r]; the gather reads the parts covering its source range
        self.ring_parts = ring_parts or (1, 0, (0,))
        self.kernel_name = f"gather_{name}"
        self._build()

    def _build(self):
        self.rec_name = f"hb_syn_{self.name}"
        # synaptic variables that the statements write are not copied into the records
        self.rec_arrays = {} if self.plastic else dict(self.syn_arrays)  # the record layout's variables
        fields, words, align = synapse_record(self.rec_arrays)
        self.rec_words = words
        body = " ".join(f"{ct} {n};" for n, ct in fields)
        self.rec_decl = (f"// one synapse: target and the on_pre statement's synaptic variables ({4 * words} bytes)\n"
                         f"struct __attribute__((aligned({align}))) {self.rec_name} {{ {body} }};")
        ps = ParamStruct(f"hbg_{self.name}")
        ps.ptr("clock", "long long", const=True)
        ps.val("step_off", "long long")
        ps.ptr("ring_spk", "int", const=True)
        ps.ptr("ring_cnt", "int", const=True)
        ps.val("ring_stride", "long long")
        ps.val("ring_slots", "int")
        ps.val("src0", "int")
        ps.val("src1", "int")
        ps.ptr("seg_ptr", "unsigned int", const=True)
        ps.ptr("syn_rec", "unsigned int", const=True)
        if self.plastic:
            ps.val("dt", "double")
            for v, ct in self.syn_arrays.items():
                ps.ptr(f"s_{v}", ct)
        for n, (_, ct) in sorted(self.scalars.items()):
            ps.val(f"c_{n}", ct)
        ps.ptr("stg_rec", "unsigned int")
        ps.ptr("tile_off", "unsigned int", const=True)
        ps.ptr("tile_cnt", "unsigned int")
        ps.val("stg_stride", "unsigned long long")
        ps.ptr("status", "int")
        ps.ptr("events", "unsigned long long")  # per-workgroup delivered-event totals
        if self.count_synapses:
            ps.ptr("syn_count", "int")           # per-synapse event counts (CSR order)
        self.struct = ps
        GT = self.block
        U = self.events_per_thread
        WN = GT * U
        PC = self.pairs
        NS = self.n_seg
        RP, P0, PB = self.ring_parts
        NR = len(PB)  # parts read: segment s of the pair enumeration is (delay s // NR, part s % NR)
        NSR = NS * NR
        NT = self.n_tiles
        TT = self.tile
        VT = self.val_t
        # the on_pre statements in order; the target update's value goes to hb_gv (the fold
        # applies it), the other statements run as they are
        lines = onpre_lines(self.on_pre, self.target_index, self.syn_arrays, self.scalars, self.float_t,
                            f"hb_gv[hb_u] = ({VT})", VT)
        k = [source_header(GT), DECL_MARK, ps.c_decl(), self.rec_decl]
        k.append(f'extern "C" __global__ void __launch_bounds__({GT}) {self.kernel_name}(const {ps.name} p) {{')
        k += [
            "    const long long step = p.clock[0] + p.step_off;",
            "    const int par = (int)(step % 3);  // staging buffer of this step",
            f"    {rec_type(VT)}* hb_st = ({rec_type(VT)}*)p.stg_rec + (size_t)par * p.stg_stride;",
            f"    __shared__ int s_cum[{NSR + 1}];",
            f"    __shared__ unsigned int s_a[{PC}];",
            f"    __shared__ int s_o[{PC}];",
            f"    __shared__ unsigned short s_sg[{PC}];",
            f"    __shared__ int s_hist[{NT}];",
            f"    __shared__ unsigned int s_base[{NT}];",
            f"    __shared__ unsigned short s_pid[{WN}];",
        ]
        vb = 8 if VT in ("double", "long long") else 4
        lds = WN * 2 + NT * 8 + (NSR + 1) * 4 + PC * 10 + 1024
        if lds > LDS_BYTES:
            raise NotImplementedError("too many target tiles for the gather's LDS")
        self.staged = GATHER_STAGE_RUNS and lds + WN * (12 + vb) <= LDS_BYTES
        self.lds_bytes = lds + (WN * (12 + vb) if self.staged else 0)
        # workgroups that share a CU: LDS, and 16 waves of at most 128 VGPRs
        self.wgs_per_cu = max(1, min(LDS_BYTES // self.lds_bytes, 1024 // GT))
        if self.staged:
            k += [f"    __shared__ unsigned long long s_ok[{WN}];", f"    __shared__ {VT} s_ov[{WN}];",
                  f"    __shared__ unsigned int s_dst[{WN}];"]
        k += [
            f"    __shared__ __align__(16) int s_sh[{GT // 64 + 1}];",
            "    __shared__ int s_ovw;  // a run of this window overflows its region",
            "    // s_cum[s]: (delay, spike) pairs of the delay segments before s (spikes of step - dmin - s)",
            "    if (threadIdx.x == 0) { s_cum[0] = 0; s_ovw = 0; }",
            f"    for (int t = threadIdx.x; t < {NT}; t += {GT}) s_hist[t] = 0;",
            "    {",
            "        int carry = 0;",
            f"        for (int s0 = 0; s0 < {NSR}; s0 += {GT}) {{",
            "            const int s = s0 + (int)threadIdx.x;",
        ]
        if NR == 1 and RP == 1:
            k += [
                f"            const long long ps_ = step - {self.dmin} - s;",
                f"            const int n_ = s < {NS} && ps_ >= 0 ? p.ring_cnt[(int)(ps_ % p.ring_slots)] : 0;",
            ]
        else:
            k += [
                f"            const long long ps_ = step - {self.dmin} - s / {NR};",
                f"            const int n_ = s < {NSR} && ps_ >= 0 ? p.ring_cnt[(int)(ps_ % p.ring_slots) * {RP} + {P0} + s % {NR}] : 0;",
            ]
        k += [
            "            int tot;",
            "            const int x = hb_block_exscan(n_, s_sh, &tot) + carry;",
            f"            if (s < {NSR}) s_cum[s + 1] = x + n_;",
            "            carry += tot;",
            "        }",
            "    }",
            "    __syncthreads();",
            f"    const int P = s_cum[{NSR}];",
            f"    const int n_units = (P + {PC - 1}) / {PC};",
            "    unsigned long long hb_ev = 0ull;",
            "    // one (delay, spike) pair per thread of a unit: its delay segment and source (read",
            "    // from the ring), then its synapse segment; a unit's pairs are read while the",
            "    // previous unit is processed (the source at its start, the segment before its last",
            "    // window is written out)",
            "    // (the loads are unconditional, at a valid index for a thread without a pair, so that",
            "    // they stay in flight: a load under a condition is waited for inside its branch)",
            "#define HB_PAIR_SRC(c_, seg_, src_, ok_) { \\",
            f"        const int q_ = (c_) * {PC} + (int)threadIdx.x; \\",
            f"        const bool v_ = (int)threadIdx.x < {PC} && q_ < P; \\",
            f"        int lo = 0, hi = {NSR}; \\",
            "        if (v_) while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (s_cum[m] <= q_) lo = m; else hi = m; } \\",
        ]
        if NR == 1 and RP == 1:
            k += [
                "        seg_ = lo; \\",
                f"        const int slot = (int)((step - {self.dmin} - lo) % p.ring_slots); \\",
                "        src_ = p.ring_spk[v_ ? (size_t)slot * p.ring_stride + (q_ - s_cum[lo]) : (size_t)0]; ok_ = v_; }",
            ]
        else:
            start = " : ".join(f"hb_j == {j} ? {b}" for j, b in enumerate(PB[:-1])) + (" : " if NR > 1 else "") + str(PB[-1])
            k += [
                f"        seg_ = lo / {NR}; \\",
                f"        const int hb_j = lo % {NR}; \\",
                f"        const int slot = (int)((step - {self.dmin} - seg_) % p.ring_slots); \\",
                f"        src_ = p.ring_spk[v_ ? (size_t)slot * p.ring_stride + ({start}) + (q_ - s_cum[lo]) : (size_t)0]; \\",
                "        ok_ = v_; (void)hb_j; }",
            ]
        k += [
            "#define HB_PAIR_SEG(seg_, src_, ok_, a_, len_) { \\",
            "        asm volatile(\"\" : \"+v\"(src_)); /* the source is read here, not earlier */ \\",
            "        const bool v_ = ok_ && src_ >= p.src0 && src_ < p.src1; \\",
            f"        const size_t row = v_ ? (size_t)(src_ - p.src0) * {NS} + (seg_) : (size_t)0; \\",
            "        const unsigned int x0_ = p.seg_ptr[row], x1_ = p.seg_ptr[row + 1]; \\",
            "        a_ = v_ ? x0_ : 0u; len_ = v_ ? x1_ - x0_ : 0u; }",
            "    int n_seg = 0, n_src = -1;",
            "    bool n_ok = false;",
            "    unsigned int n_a = 0u, n_len = 0u;",
            "    if ((int)blockIdx.x < n_units) {",
            "        HB_PAIR_SRC((int)blockIdx.x, n_seg, n_src, n_ok)",
            "        HB_PAIR_SEG(n_seg, n_src, n_ok, n_a, n_len)",
            "    }",
            "    for (int c = blockIdx.x; c < n_units; c += gridDim.x) {",
            f"        const int q0 = c * {PC};",
            f"        const int nq = P - q0 < {PC} ? P - q0 : {PC};",
            "        const unsigned int a = n_a, len = n_len;",
            "        const int seg = n_seg;",
            "        const bool hb_more = c + (int)gridDim.x < n_units;",
            "        HB_PAIR_SRC(c + (int)gridDim.x, n_seg, n_src, n_ok)  // no pair for any thread without a next unit",
            "        int E;",
            "        const int ex = hb_block_exscan((int)len, s_sh, &E);",
            f"        if ((int)threadIdx.x < {PC}) {{ s_a[threadIdx.x] = a; s_o[threadIdx.x] = ex; s_sg[threadIdx.x] = (unsigned short)seg; }}",
            "        // each pair writes its index over its events of the window (the first window's here,",
            "        // before the barrier the unit needs anyway)",
            "        if ((int)threadIdx.x < nq) {",
            "            const int e_ = ex + (int)len;",
            f"            const int b_ = e_ < {WN} ? e_ : {WN};",
            "            for (int x_ = ex; x_ < b_; x_++) s_pid[x_] = (unsigned short)threadIdx.x;",
            "        }",
            "        __syncthreads();",
            f"        for (int w0 = 0; w0 < E; w0 += {WN}) {{",
            "            if (w0 > 0) {",
            "                if ((int)threadIdx.x < nq) {",
            "                    const int a_ = ex > w0 ? ex : w0;",
            "                    const int e_ = ex + (int)len;",
            f"                    const int b_ = e_ < w0 + {WN} ? e_ : w0 + {WN};",
            "                    for (int x_ = a_; x_ < b_; x_++) s_pid[x_ - w0] = (unsigned short)threadIdx.x;",
            "                }",
            "                __syncthreads();",
            "            }",
            f"            int hb_gt[{U}];",
            f"            unsigned long long hb_gk[{U}];",
            f"            {VT} hb_gv[{U}];",
            "            // the window's events in registers, counted per target tile: positions first,",
            "            // then all loads in flight at once, then keys and values",
            f"            unsigned int hb_ge[{U}];",
            f"            unsigned short hb_gs[{U}];",
            f"            {self.rec_name} hb_gr[{U}];",
        ]
        k += [
            "            #pragma unroll",
            f"            for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
            f"                const int ev = w0 + hb_u * {GT} + (int)threadIdx.x;",
            "                const int evc = ev < E ? ev : E - 1;",
            "                const int lo = s_pid[evc - w0];",
            "                hb_ge[hb_u] = s_a[lo] + (unsigned int)(evc - s_o[lo]);",
            "                hb_gs[hb_u] = s_sg[lo];",
            "                hb_gt[hb_u] = ev < E ? 0 : -1;",
            "            }",
            "            #pragma unroll",
            f"            for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
            f"                hb_gr[hb_u] = ((const {self.rec_name}*)p.syn_rec)[hb_ge[hb_u]];",
        ]
        k += [
            "            }",
            "            #pragma unroll",
            f"            for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
            "                if (hb_gt[hb_u] >= 0) {",
            "                    const unsigned int e = hb_ge[hb_u];",
            "                    const int _postsynaptic_idx = hb_gr[hb_u].post;",
            "                    const int _idx = (int)e;",
            "                    (void)_idx;",
        ]
        if self.plastic:
            k.append("                    const double t = (double)step * p.dt;")
            k += [f"                    {ct} {v} = p.s_{v}[e];" for v, ct in self.syn_arrays.items()]
        else:
            k += [f"                    const {ct} {v} = hb_gr[hb_u].{v};" for v, ct in self.syn_arrays.items()]
        k += ["                    " + x for x in lines]
        k += [f"                    p.s_{v}[e] = {v};" for v in self.syn_arrays if v in self.plastic]
        k += [
            f"                    const int tile = _postsynaptic_idx / {TT};",
            "                    hb_gt[hb_u] = tile;",
            "                    // a target outside the tiles (corrupt synapse data) is reported, not staged",
            f"                    if ((unsigned int)_postsynaptic_idx >= {NT * TT}u) {{ hb_gt[hb_u] = -1; atomicOr(p.status, 8); }}",
            f"                    hb_gk[hb_u] = ((unsigned long long)(_postsynaptic_idx - tile * {TT}) << {KEY_TGT_SHIFT})",
            f"                        | ((unsigned long long){self.rank} << 44) | ((unsigned long long)({NS} - hb_gs[hb_u]) << 32)",
            "                        | (unsigned long long)e;",
            "                }",
            "            }",
            "            #pragma unroll",
            f"            for (int hb_u = 0; hb_u < {U}; hb_u++) if (hb_gt[hb_u] >= 0) atomicAdd(&s_hist[hb_gt[hb_u]], 1);",
            "            __syncthreads();",
        ]
        fits = [
            "            const bool ovw = s_ovw != 0;",
            "            // an event at buffer position r of tile t fits unless r passes the region's end",
            "#define HB_FITS(t, r) (!ovw || (r) < (unsigned long long)p.tile_off[(t) + 1])",
        ]
        if not self.staged:
            k += [
                "            // space in the regions of the tiles the window touches: one global atomic per tile;",
                "            // s_base = buffer position of the window's run.  A run that does not fit is written",
                "            // up to the region's end (so the region has no unwritten hole) and reported",
                f"            for (int t = threadIdx.x; t < {NT}; t += {GT}) {{",
                "                const unsigned int c = (unsigned int)s_hist[t];",
                "                if (c > 0u) {",
                "                    const unsigned int lo = p.tile_off[t], hi = p.tile_off[t + 1];",
                f"                    const unsigned int b = atomicAdd(p.tile_cnt + (size_t)par * {NT} + t, c);",
                "                    if ((unsigned long long)b + c > hi - lo) { s_ovw = 1; atomicOr(p.status, 1); }",
                "                    s_base[t] = b < hi - lo ? lo + b : hi;  // position in the buffer",
                "                }",
                "            }",
                "            __syncthreads();",
            ] + fits
        else:
            k += [
                "            // space in the regions of the tiles the window touches (one global atomic per tile,",
                "            // in flight during the block scan) and the window sorted by tile in LDS with each",
                "            // event's buffer position (s_dst), then written out run by run, consecutive",
                "            // threads on consecutive positions.  s_hist[t] becomes the start of tile t's run in",
                "            // the sorted window and s_base[t] the buffer position of the run minus that start.",
                "            // A run that does not fit is written up to the region's end (so the region has no",
                "            // unwritten hole) and reported",
                "            int carry = 0;",
                f"            for (int t0 = 0; t0 < {NT}; t0 += {GT}) {{",
                "                const int t = t0 + (int)threadIdx.x;",
                f"                const unsigned int c = t < {NT} ? (unsigned int)s_hist[t] : 0u;",
                "                unsigned int lo = 0u, hi = 0u, b = 0u;",
                "                if (c > 0u) {",
                "                    lo = p.tile_off[t]; hi = p.tile_off[t + 1];",
                f"                    b = atomicAdd(p.tile_cnt + (size_t)par * {NT} + t, c);",
                "                }",
                "                int tot;",
                "                const int x = hb_block_exscan((int)c, s_sh, &tot) + carry;",
                f"                if (t < {NT}) s_hist[t] = x;",
                "                if (c > 0u) {",
                "                    if ((unsigned long long)b + c > hi - lo) { s_ovw = 1; atomicOr(p.status, 1); }",
                "                    s_base[t] = (b < hi - lo ? lo + b : hi) - (unsigned int)x;",
                "                }",
                "                carry += tot;",
                "            }",
                "            __syncthreads();",
            ] + fits + [
                "            #pragma unroll",
                f"            for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
                "                if (hb_gt[hb_u] >= 0) {",
                "                    const int t = hb_gt[hb_u];",
                "                    const int q = atomicAdd(&s_hist[t], 1);",
                "                    s_ok[q] = hb_gk[hb_u];",
                "                    s_ov[q] = hb_gv[hb_u];",
                "                    const unsigned int r = s_base[t] + (unsigned int)q;",
                "                    s_dst[q] = HB_FITS(t, (unsigned long long)r) ? r : 0xffffffffu;  // no position is 2^32 - 1",
            ]
            if self.count_synapses:
                k.append("                    p.syn_count[(unsigned int)hb_gk[hb_u]] += 1;")
            k += [
                "                }",
                "            }",
                "            __syncthreads();",
                f"            if (hb_more && w0 + {WN} >= E) HB_PAIR_SEG(n_seg, n_src, n_ok, n_a, n_len)",
                "            // carry = the events placed (the window's events with a valid target)",
                "            // (a window holds at most U events per thread: their positions, keys and values are",
                "            // read from LDS together, then written)",
                f"            unsigned int hb_wr[{U}];",
                f"            unsigned long long hb_wk[{U}];",
                f"            {VT} hb_wv[{U}];",
                "            #pragma unroll",
                f"            for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
                f"                const int x = hb_u * {GT} + (int)threadIdx.x;",
                "                const int xc = x < carry ? x : 0;",
                "                hb_wr[hb_u] = x < carry ? s_dst[xc] : 0xffffffffu; hb_wk[hb_u] = s_ok[xc]; hb_wv[hb_u] = s_ov[xc];",
                "            }",
                "            #pragma unroll",
                f"            for (int hb_u = 0; hb_u < {U}; hb_u++)",
                "                if (hb_wr[hb_u] != 0xffffffffu) hb_rset(hb_st[hb_wr[hb_u]], hb_wk[hb_u], hb_wv[hb_u]);",
                "            // the counts of the next window (the write-out reads neither)",
                f"            for (int t = threadIdx.x; t < {NT}; t += {GT}) s_hist[t] = 0;",
                "            if (threadIdx.x == 0) s_ovw = 0;",
                "            __syncthreads();",
            ]
        if not self.staged:
            k += [
                f"            for (int t = threadIdx.x; t < {NT}; t += {GT}) s_hist[t] = 0;",
                f"            if (hb_more && w0 + {WN} >= E) HB_PAIR_SEG(n_seg, n_src, n_ok, n_a, n_len)",
                "            __syncthreads();",
                "            #pragma unroll",
                f"            for (int hb_u = 0; hb_u < {U}; hb_u++) {{",
                "                if (hb_gt[hb_u] >= 0) {",
                "                    const int t = hb_gt[hb_u];",
                "                    const unsigned long long r = (unsigned long long)s_base[t] + (unsigned int)atomicAdd(&s_hist[t], 1);",
                "                    if (HB_FITS(t, r)) {",
                "                        hb_rset(hb_st[r], hb_gk[hb_u], hb_gv[hb_u]);",
                "                    }",
            ]
            if self.count_synapses:
                k.append("                    p.syn_count[(unsigned int)hb_gk[hb_u]] += 1;")
            k += [
                "                }",
                "            }",
                "            __syncthreads();",
                f"            for (int t = threadIdx.x; t < {NT}; t += {GT}) s_hist[t] = 0;",
                "            if (threadIdx.x == 0) s_ovw = 0;",
                "            __syncthreads();",
            ]
        k += [
            "#undef HB_FITS",
            "        }",
            "        if (E == 0 && hb_more) HB_PAIR_SEG(n_seg, n_src, n_ok, n_a, n_len)",
            "        if (threadIdx.x == 0) hb_ev += (unsigned long long)E;",
            "    }",
            "#undef HB_PAIR_SRC",
            "#undef HB_PAIR_SEG",
            "    if (threadIdx.x == 0) p.events[blockIdx.x] += hb_ev;",
            "}",
        ]
        self.source = "\n".join(k) + "\n"


class CarryKernel:
    """Events carried over a change of dt: CSpikeQueue::prepare re-bins the spikes
    in flight (spikequeue.h:58-74), so after the change a pathway delivers, at the
    start of each of the next few steps, a fixed list of synapses (the re-binned
    slot) before the events of later pushes.  One thread per listed synapse: the
    on_pre statements run as in the pathway's gather (same synapse records), and
    the event is staged with push-order field 0 and its position in the slot's
    list as the last key field, so that the fold applies the slot in its order and
    before every event of the gathers."""

    BLOCK = 256

    def __init__(self, gather):
        g = gather
        self.name = g.name
        self.kernel_name = f"carry_{g.name}"
        ps = ParamStruct(f"hbc_{g.name}")
        ps.ptr("clock", "long long", const=True)
        ps.val("step_off", "long long")
        ps.val("t0", "long long")          # the step of relative slot 0
        ps.val("n_slots", "long long")
        ps.ptr("slot_ptr", "unsigned int", const=True)  # n_slots + 1 offsets into pos
        ps.ptr("pos", "unsigned int", const=True)       # CSR positions, slot by slot, in slot order
        ps.ptr("syn_rec", "unsigned int", const=True)
        if g.plastic:
            ps.val("dt", "double")
            for v, ct in g.syn_arrays.items():
                ps.ptr(f"s_{v}", ct)
        for n, (_, ct) in sorted(g.scalars.items()):
            ps.val(f"c_{n}", ct)
        ps.ptr("stg_rec", "unsigned int")
        ps.ptr("tile_off", "unsigned int", const=True)
        ps.ptr("tile_cnt", "unsigned int")
        ps.val("stg_stride", "unsigned long long")
        ps.ptr("status", "int")
        ps.ptr("events", "unsigned long long")
        if g.count_synapses:
            ps.ptr("syn_count", "int")
        self.struct = ps
        VT, NT, TT = g.val_t, g.n_tiles, g.tile
        lines = onpre_lines(g.on_pre, g.target_index, g.syn_arrays, g.scalars, g.float_t, f"hb_val = ({VT})", VT)
        k = [source_header(self.BLOCK), DECL_MARK, ps.c_decl(), g.rec_decl,
             f'extern "C" __global__ void __launch_bounds__({self.BLOCK}) {self.kernel_name}(const {ps.name} p) {{',
             "    const long long step = p.clock[0] + p.step_off;",
             "    const long long r = step - p.t0;",
             "    if (r < 0 || r >= p.n_slots) return;",
             "    const unsigned int a = p.slot_ptr[r], b = p.slot_ptr[r + 1];",
             "    const int par = (int)(step % 3);",
             f"    {rec_type(VT)}* hb_st = ({rec_type(VT)}*)p.stg_rec + (size_t)par * p.stg_stride;",
             "    unsigned long long hb_ev = 0ull;",
             "    for (unsigned int j = a + blockIdx.x * blockDim.x + threadIdx.x; j < b; j += gridDim.x * blockDim.x) {",
             "        const unsigned int e = p.pos[j];",
             f"        const {g.rec_name} hb_r = ((const {g.rec_name}*)p.syn_rec)[e];",
             "        const int _postsynaptic_idx = hb_r.post;",
             "        const int _idx = (int)e;",
             "        (void)_idx;",
             f"        {VT} hb_val;"]
        if g.plastic:
            k.append("        const double t = (double)step * p.dt;")
            k += [f"        {ct} {v} = p.s_{v}[e];" for v, ct in g.syn_arrays.items()]
        else:
            k += [f"        const {ct} {v} = hb_r.{v};" for v, ct in g.syn_arrays.items()]
        k += ["        " + x for x in lines]
        k += [f"        p.s_{v}[e] = {v};" for v in g.syn_arrays if v in g.plastic]
        k += [
            f"        if ((unsigned int)_postsynaptic_idx >= {NT * TT}u) {{ atomicOr(p.status, 8); continue; }}",
            f"        const int tile = _postsynaptic_idx / {TT};",
            f"        const unsigned long long key = ((unsigned long long)(_postsynaptic_idx - tile * {TT}) << {KEY_TGT_SHIFT})",
            f"            | ((unsigned long long){g.rank} << 44) | (unsigned long long)(j - a);",
            f"        const unsigned int c = atomicAdd(p.tile_cnt + (size_t)par * {NT} + tile, 1u);",
            "        const unsigned int lo = p.tile_off[tile], hi = p.tile_off[tile + 1];",
            "        if (c >= hi - lo) { atomicOr(p.status, 1); continue; }",
            "        hb_rset(hb_st[lo + c], key, hb_val);",
        ]
        if g.count_synapses:
            k.append("        p.syn_count[e] += 1;  // a slot lists a synapse at most once")
        k += [
            "        hb_ev++;",
            "    }",
            "    if (hb_ev) atomicAdd(p.events, hb_ev);",
            "}",
        ]
        self.source = "\n".join(k) + "\n"


class PostPathKernel:
    """A pathway's on_post code (brian2/synapses/synapses.py:205-216; template
    synapses.pyx:16-53 run on the post pathway's queue): for every spike of the
    target group at step s - delay, the statements run for each of the neuron's
    incoming synapses (CSR by target, synapse order).  The statements may read
    and write synaptic variables only, so synapses are independent: one thread
    per synapse, one workgroup per spike at a time."""

    BLOCK = 256

    def __init__(self, name, statements, syn_arrays, plastic, scalars, float_t, delay_steps, ring_parts=None):
        self.name = name
        self.ring_parts = tuple(ring_parts or (0,))  # part starts of the target's spike ring
        self.statements = statements
        self.syn_arrays = syn_arrays
        self.plastic = set(plastic)
        self.scalars = scalars
        self.float_t = float_t
        self.delay = int(delay_steps)
        self.kernel_name = f"post_{name}"
        ps = ParamStruct(f"hbo_{name}")
        ps.ptr("clock", "long long", const=True)
        ps.val("step_off", "long long")
        ps.val("dt", "double")
        ps.ptr("ring_spk", "int", const=True)
        ps.ptr("ring_cnt", "int", const=True)
        ps.val("ring_stride", "long long")
        ps.val("ring_slots", "int")
        ps.val("pad", "int")
        ps.ptr("post_ptr", "unsigned int", const=True)
        ps.ptr("post_syn", "unsigned int", const=True)
        for v, ct in syn_arrays.items():
            ps.ptr(f"s_{v}", ct)
        for n, (_, ct) in sorted(scalars.items()):
            ps.val(f"c_{n}", ct)
        self.struct = ps
        types = dict(syn_arrays)
        types["_idx"] = "int"
        types["t"] = "double"
        for n, (_, ct) in scalars.items():
            types[n] = ct
        lines = _rename(translate_block(statements, {}, types, set(), float_t), {n: f"p.c_{n}" for n in scalars})
        B = self.BLOCK
        k = [source_header(B), ps.c_decl()]
        k.append(f'extern "C" __global__ void __launch_bounds__({B}) {self.kernel_name}(const {ps.name} p) {{')
        k += [
            "    const long long step = p.clock[0] + p.step_off;",
            "    const double t = (double)step * p.dt;",
            f"    const long long ps_ = step - {self.delay};",
            "    if (ps_ < 0) return;",
            "    const int slot = (int)(ps_ % p.ring_slots);",
        ]
        R = len(self.ring_parts)
        if R == 1:
            k += [
                "    const int P = p.ring_cnt[slot];",
                "    for (int c = blockIdx.x; c < P; c += gridDim.x) {",
                "        const int j = p.ring_spk[(size_t)slot * p.ring_stride + c];",
            ]
        else:
            # the parts' spikes one after the other: part r holds spikes [c_r, c_{r+1}) of the slot
            k += ["    const int hb_c0 = 0;"]
            k += [f"    const int hb_c{r + 1} = hb_c{r} + p.ring_cnt[slot * {R} + {r}];" for r in range(R)]
            pos = "c - hb_c0 + " + str(self.ring_parts[0])
            for r in range(1, R):
                pos = f"c >= hb_c{r} ? c - hb_c{r} + {self.ring_parts[r]} : {pos}"
            k += [
                f"    const int P = hb_c{R};",
                "    for (int c = blockIdx.x; c < P; c += gridDim.x) {",
                f"        const int j = p.ring_spk[(size_t)slot * p.ring_stride + ({pos})];",
            ]
        k += [
            "        const unsigned int lo = p.post_ptr[j], hi = p.post_ptr[j + 1];",
            f"        for (unsigned int k_ = lo + threadIdx.x; k_ < hi; k_ += {B}) {{",
            "            const unsigned int e = p.post_syn[k_];",
            "            const int _idx = (int)e;",
            "            (void)_idx; (void)t;",
        ]
        k += [f"            {ct} {v} = p.s_{v}[e];" for v, ct in syn_arrays.items()]
        k += ["            " + x for x in lines]
        k += [f"            p.s_{v}[e] = {v};" for v in syn_arrays if v in self.plastic]
        k += ["        }", "    }", "}"]
        self.source = "\n".join(k) + "\n"


class SummedKernel:
    """A summed variable (synapses.py:1146-1200, SummedVariableUpdater; template
    summed_variable.pyx): target[j] = sum over the synapses of j, in synapse order,
    of the expression, starting from 0, for the targets j of the synapses' target
    (sub)group.  One thread per target, the synapses of the target (CSR by target,
    synapse order) folded sequentially, as the template's loop over all synapses
    does for each target's subsequence."""

    BLOCK = 256

    def __init__(self, name, statements, syn_arrays, pre_reads, post_reads, var_types, target_t, scalars,
                 float_t):
        self.kernel_name = f"summed_{name}"
        ps = ParamStruct(f"hbs_{name}")
        ps.ptr("clock", "long long", const=True)
        ps.val("step_off", "long long")
        ps.val("dt", "double")
        ps.val("t0", "int")  # the targets [t0, t1) of the group (a subgroup target: its range)
        ps.val("t1", "int")
        ps.ptr("post_ptr", "unsigned int", const=True)
        ps.ptr("post_syn", "unsigned int", const=True)
        ps.ptr("pre_of", "int", const=True)
        ps.ptr("target", target_t)
        for v, ct in syn_arrays.items():
            ps.ptr(f"s_{v}", ct, const=True)
        for n, gv in sorted(pre_reads.items()):
            ps.ptr(f"a_{n}", var_types[n], const=True)
        for n, gv in sorted(post_reads.items()):
            ps.ptr(f"a_{n}", var_types[n], const=True)
        for n, (_, ct) in sorted(scalars.items()):
            ps.val(f"c_{n}", ct)
        self.struct = ps
        types = dict(syn_arrays)
        types.update({n: var_types[n] for n in list(pre_reads) + list(post_reads)})
        types.update({"_idx": "int", "_presynaptic_idx": "int", "_postsynaptic_idx": "int", "t": "double"})
        for n, (_, ct) in scalars.items():
            types[n] = ct
        lines = _rename(translate_block(statements, {}, types, set(), float_t), {n: f"p.c_{n}" for n in scalars})
        B = self.BLOCK
        k = [source_header(B), ps.c_decl()]
        k.append(f'extern "C" __global__ void __launch_bounds__({B}) {self.kernel_name}(const {ps.name} p) {{')
        k += [
            "    const long long step = p.clock[0] + p.step_off;",
            "    const double t = (double)step * p.dt;",
            "    (void)t;",
            f"    const int j = p.t0 + (int)(blockIdx.x * {B} + threadIdx.x);",
            "    if (j >= p.t1) return;",
            "    const int _postsynaptic_idx = j;",
            f"    {target_t} hb_acc = ({target_t})0;",
            "    for (unsigned int k_ = p.post_ptr[j]; k_ < p.post_ptr[j + 1]; k_++) {",
            "        const unsigned int e = p.post_syn[k_];",
            "        const int _idx = (int)e;",
            "        const int _presynaptic_idx = p.pre_of[e];",
            "        (void)_idx; (void)_presynaptic_idx; (void)_postsynaptic_idx;",
        ]
        k += [f"        const {ct} {v} = p.s_{v}[e];" for v, ct in syn_arrays.items()]
        k += [f"        const {var_types[n]} {n} = p.a_{n}[_presynaptic_idx];" for n in sorted(pre_reads)]
        k += [f"        const {var_types[n]} {n} = p.a_{n}[_postsynaptic_idx];" for n in sorted(post_reads)]
        k += ["        " + x for x in lines]
        k += ["        hb_acc += _synaptic_var;", "    }", "    p.target[j] = hb_acc;", "}"]
        self.source = "\n".join(k) + "\n"


__all__ = ["CodeSpec", "NeuronKernel", "GatherKernel", "PostPathKernel", "SummedKernel", "scalar_values",
           "identifiers", "c_literal"]


# ---------------------------------------------------------------------------
_SHARED_DECL = re.compile(r"^(\s*)__shared__ (?:__align__\((\d+)\) )?"
                          r"((?:unsigned )?(?:long long|int|short|char|double|float|bool)) ([^;]+);(.*)$")
_CT_BYTES = {"long long": 8, "unsigned long long": 8, "double": 8, "int": 4, "unsigned int": 4, "float": 4,
             "short": 2, "unsigned short": 2, "char": 1, "unsigned char": 1, "bool": 1}


def device_body(source, kernel_name, fname):
    """A generated kernel as a __device__ function for a fused launch.

    Returns (declarations, function, lds_bytes): the source's argument struct and
    record declarations, and `fname(const S& p, long long hb_step_off, int hb_blk,
    int hb_nblk, unsigned char* hb_lds)`, in which the workgroup index and count
    are hb_blk and hb_nblk, the step offset is hb_step_off and every __shared__
    variable is a reference into hb_lds (lds_bytes from its start): the bodies of
    one fused kernel then share one LDS allocation instead of adding up."""
    lines = source.split("\n")
    start = next(i for i, x in enumerate(lines) if x.startswith('extern "C" __global__') and f" {kernel_name}(" in x)
    m = re.search(r"\(const (\w+) p\) \{$", lines[start])
    if m is None:
        raise ValueError(f"unexpected signature of {kernel_name}")
    # declarations: after the header (which ends with the prelude), before the kernel
    head = "\n".join(lines[:start])
    decls = head.split(DECL_MARK, 1)[1] if DECL_MARK in head else None
    if decls is None:
        raise ValueError(f"{kernel_name}: declarations are not marked")
    out = [f"__device__ __forceinline__ void {fname}(const {m.group(1)}& p, long long hb_step_off, int hb_blk, "
           "int hb_nblk, unsigned char* hb_lds) {"]
    off = 0
    for x in lines[start + 1:]:
        d = _SHARED_DECL.match(x)
        if d:
            ind, al, ct, names, rest = d.groups()
            nb = _CT_BYTES[ct]
            refs = []
            for nm in (n.strip() for n in names.split(",")):
                a = re.fullmatch(r"(\w+)\[(\d+)\]", nm)
                align = int(al) if al else nb
                off = (off + align - 1) // align * align
                if a:
                    refs.append(f"{ct} (&{a.group(1)})[{a.group(2)}] = *reinterpret_cast<{ct} (*)[{a.group(2)}]>"
                                f"(hb_lds + {off});")
                    off += nb * int(a.group(2))
                else:
                    refs.append(f"{ct}& {nm} = *reinterpret_cast<{ct}*>(hb_lds + {off});")
                    off += nb
            out.append(ind + " ".join(refs) + rest)
            continue
        if "__shared__" in x:
            raise ValueError(f"{kernel_name}: unrecognised LDS declaration: {x.strip()}")
        x = x.replace("blockIdx.x", "hb_blk").replace("gridDim.x", "hb_nblk").replace("p.step_off", "hb_step_off")
        out.append(x)
    return decls, "\n".join(out), (off + 15) // 16 * 16


class FusedStepKernel:
    """One launch per step for a group and the gathers of its incoming pathways.

    Valid when every incoming pathway's delays are at least one step: the gathers
    of step s then read spikes of earlier steps only and write staging buffer
    s % 3, which this step's fold does not read, so they are independent of the
    group update in the same launch (the same independence that lets them run on
    other streams).  Workgroups [0, sum(gather_blocks)) run the gathers, each
    pathway on its own range of workgroups; the rest run the tiles of the group.
    The bodies are the separate kernels' own code (device_body), sharing one LDS
    allocation; the step offset is the fused argument struct's first field."""

    def __init__(self, step, gathers, gather_blocks, n_tiles, tiles_first=False):
        if step.flush or any(g.block != step.T for g in gathers):
            raise ValueError("a fused launch needs a step kernel and gathers of one workgroup size")
        if int(n_tiles) != -(-step.N // (step.T * step.npt)) or len(gather_blocks) != len(gathers):
            raise ValueError("the fused launch's tiles or gather workgroups do not match its kernels")
        self.step, self.gathers = step, list(gathers)
        self.gather_blocks = [int(b) for b in gather_blocks]
        self.n_tiles = int(n_tiles)
        self.tiles_first = bool(tiles_first)
        self.block = step.T
        self.kernel_name = f"fused_{step.name}"
        parts = [(step, f"hbb_step_{step.name}", "s")] + [
            (g, f"hbb_gather_{g.name}", f"g{i}") for i, g in enumerate(self.gathers)]
        decls, funcs, lds = [], [], 0
        for K, fname, _ in parts:
            d, f, nb = device_body(K.source, K.kernel_name, fname)
            decls.append(d)
            funcs.append(f)
            lds = max(lds, nb)
        self.lds_bytes = lds
        if lds > LDS_BYTES:
            raise ValueError("the fused kernel's LDS exceeds a CU's")
        self.struct_name = f"hbf_{step.name}"
        fields = ["    long long step_off;"] + [f"    {K.struct.name} {m};" for K, _, m in parts]
        occ = f" __attribute__((amdgpu_waves_per_eu({step.waves_per_eu})))" if step.waves_per_eu else ""
        k = [source_header(step.T)] + decls + funcs
        k.append(f"struct {self.struct_name} {{\n" + "\n".join(fields) + "\n};")
        k.append(f'extern "C" __global__ void __launch_bounds__({step.T}){occ} {self.kernel_name}'
                 f"(const {self.struct_name} a) {{")
        k.append(f"    __shared__ __align__(16) unsigned char hb_lds[{lds}];")
        k.append("    int b = (int)blockIdx.x;")
        tiles = f"    if (b < {self.n_tiles}) {{ hbb_step_{step.name}(a.s, a.step_off, b, {self.n_tiles}, hb_lds); return; }}"
        if self.tiles_first:
            k += [tiles, f"    b -= {self.n_tiles};"]
        for i, (g, nb) in enumerate(zip(self.gathers, self.gather_blocks)):
            k.append(f"    if (b < {nb}) {{ hbb_gather_{g.name}(a.g{i}, a.step_off, b, {nb}, hb_lds); return; }}")
            k.append(f"    b -= {nb};")
        if not self.tiles_first:
            k.append(tiles)
        k.append("}")
        self.source = "\n".join(k) + "\n"
        self.grid = self.n_tiles + sum(self.gather_blocks)

    def ctypes_struct(self, step_cls, gather_clss):
        import ctypes

        fields = [("step_off", ctypes.c_int64), ("s", step_cls)] + [(f"g{i}", c) for i, c in enumerate(gather_clss)]
        return type(self.struct_name, (ctypes.Structure,), {"_fields_": fields})
