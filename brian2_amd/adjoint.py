"""Differentiable recurrent LIF network: device forward, exact adjoint backward.

The network: `n_in` Poisson sources -> `N` LIF neurons with an exponential
current (dv/dt = (I + bias - v)/tau_m, dI/dt = -I/tau_s, threshold v > theta,
reset v = 0, exact integration, no refractoriness), sparse input and recurrent
weights (`on_pre: I += w`), many independent trials (one workgroup each).  The
neuron update is Brian's own optimised statements for this model
(brian2_amd/specs/lif.json); the pathways run in Brian's schedule order (by
name: syn_in_pre, then syn_rec_pre), each a left fold in (source, synapse)
order, so a trial's forward pass is bit-identical to the cython runtime.

Loss per run (mean over trials b):
    L = 1/B sum_b [ sum_r ((t_br - target_r) / t_ref)^2
                    + c_v / (N T) sum_k sum_j (v_bkj / theta)^2 ]
t_br is the first spike time of readout neuron r (the first `n_readout`
neurons), linearly interpolated inside the step of the threshold crossing
between the state before the step and the updated state; v_bkj the membrane
potential after step k (after reset).  A readout that never spikes
contributes (T dt - target)^2 and no gradient.

Backward: the exact adjoint of the discrete forward map.  Per step
    u  = C + B I + A v          (the statements are linear in v, I: A, B, C and
    I' = beta I + sum_i s_i w     beta are read off them with autograd)
    v' = (1 - s) u,   s = [u > theta]
run in reverse time: lam_u = (1 - s)(lam_v' + dL/dv') (the reset is the jump
term: nothing flows back through a reset), plus the first-spike terms
dt*dt/du, dt*dt/dv at the crossing step; lam_v = A lam_u, lam_I = B lam_u +
beta lam_I'; weight gradients are event-based: dL/dw_ij += lam_I'_j at every
step where source i spiked.  Spikes themselves carry no gradient (a
perturbation that creates or deletes no spike leaves them unchanged).  The
state is checkpointed every `ckpt_every` steps; the backward pass recomputes
each segment forward from its checkpoint, then walks it in reverse.  Weight
gradients accumulate per trial (no atomics) and are summed over trials in a
fixed order.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .codegen import ParamStruct, source_header, translate_block
from .kernels import _rename, scalar_values
from .runtime import Device, Model

L = _lib.lib
SEED = 1234
S_RATE, S_POISSON, S_CONN_IN, S_CONN_REC, S_W_IN, S_W_REC, S_V0_LIF, S_TARGET = 9, 10, 11, 12, 13, 14, 15, 16
LIF = dict(N=1000, NE=800, n_in=200, tau_m=0.02, tau_s=0.005, theta=0.01, bias=0.0105, p=0.1,
           rate_lo=5.0, rate_hi=50.0, w_scale=0.5e-3, n_steps=2000, n_readout=10, t_ref=0.01, c_v=0.1)
BLOCK = 1024


def _torch():
    import torch

    return torch


def _rng_header():
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "csrc", "hb_rng.h")) as f:
        return f.read()


def lif_inputs(trials, seed=SEED, c=LIF):
    """Inputs from the shared generator (csrc/hb_rng.h streams 9-16): input
    rates r_i = 5 + 45 U Hz; Bernoulli(0.1) input and recurrent connectivity
    (CSR by source, targets ascending); weights sigma * z (z standard normal,
    polar method; sigma = 0.5 mV / sqrt(expected fan-in)); v0 = theta * U per
    (trial, neuron); readout targets 20 + 160 U ms.  The Poisson trains are
    drawn on the device: input i spikes at step k of trial b iff
    U(S_POISSON, (b n_in + i) T + k) < r_i dt."""
    from .synth import bernoulli_host, polar_normal, uniform

    n_in, N = c["n_in"], c["N"]
    rates = c["rate_lo"] + (c["rate_hi"] - c["rate_lo"]) * uniform(seed, S_RATE, np.arange(n_in, dtype=np.uint64))
    pre_in, post_in = bernoulli_host(seed, 0, n_in, N, c["p"], stream=S_CONN_IN)
    pre_rec, post_rec = bernoulli_host(seed, 0, N, N, c["p"], stream=S_CONN_REC)
    sigma = c["w_scale"] / np.sqrt(c["p"] * (n_in + N))
    w_in = sigma * polar_normal(seed, len(pre_in), S_W_IN)
    w_rec = sigma * polar_normal(seed, len(pre_rec), S_W_REC)
    v0 = c["theta"] * uniform(seed, S_V0_LIF, np.arange(trials * N, dtype=np.uint64)).reshape(trials, N)
    targets = 0.02 + 0.16 * uniform(seed, S_TARGET, np.arange(c["n_readout"], dtype=np.uint64))
    return dict(rates=rates, pre_in=pre_in, post_in=post_in, pre_rec=pre_rec, post_rec=post_rec,
                w_in=w_in, w_rec=w_rec, v0=v0, targets=targets, sigma=sigma)


def _csr_ptr(pre, n_src):
    ptr = np.zeros(n_src + 1, np.int64)
    np.cumsum(np.bincount(pre, minlength=n_src), out=ptr[1:])
    return ptr.astype(np.uint32)


def linear_map(spec, consts, dt, float_t="double"):
    """The update statements of a two-variable linear model as v' = C + B I + A v,
    I' = beta I, with the coefficients and their derivatives with respect to
    tau_m and tau_s (autograd through the scalar block on the host)."""
    torch = _torch()
    with torch.enable_grad():  # also when called from an autograd backward pass
        return _linear_map(torch, spec, consts, dt)


def _linear_map(torch, spec, consts, dt):
    import math

    tau = {n: torch.tensor(float(consts[n]), dtype=torch.float64, requires_grad=True) for n in ("tau_m", "tau_s")}
    env = {k: torch.tensor(float(v), dtype=torch.float64) for k, v in consts.items() if k not in tau}
    env.update(tau)
    env["dt"] = torch.tensor(dt, dtype=torch.float64)
    fns = {"exp": torch.exp, "expm1": torch.expm1, "log": torch.log, "sqrt": torch.sqrt}
    for st in spec.scalar:
        env[st["var"]] = eval(st["expr"], {"__builtins__": {}}, {**fns, **env})
    v = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    I_ = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    env.update(v=v, I=I_)
    loc = dict(env)
    for st in spec.vector:
        loc[st["var"]] = eval(st["expr"], {"__builtins__": {}}, {**fns, **loc})
    vn, In = loc["v"], loc["I"]
    (A, B) = torch.autograd.grad(vn, (v, I_), create_graph=True)
    (beta_v, beta) = torch.autograd.grad(In, (v, I_), create_graph=True, allow_unused=True)
    if beta_v is not None and float(beta_v) != 0.0:
        raise NotImplementedError("the synaptic current's update depends on v")
    C = vn - (A * v + B * I_)
    coef = {"A": A, "B": B, "C": C, "beta": beta}
    out = {k: float(x.detach()) for k, x in coef.items()}
    for k, x in coef.items():
        g = torch.autograd.grad(x, (tau["tau_m"], tau["tau_s"]), retain_graph=True, allow_unused=True)
        out[f"d{k}"] = tuple(0.0 if gi is None else float(gi) for gi in g)
    # linearity check: the update at (1, 1) must equal C + B + A
    loc2 = dict(env)
    loc2.update(v=torch.tensor(1.0, dtype=torch.float64), I=torch.tensor(1.0, dtype=torch.float64))
    for st in spec.vector:
        loc2[st["var"]] = eval(st["expr"], {"__builtins__": {}}, {**fns, **loc2})
    if not math.isclose(float(loc2["v"].detach()), out["C"] + out["B"] + out["A"], rel_tol=1e-12, abs_tol=1e-15):
        raise NotImplementedError("the neuron update is not linear in (v, I)")
    return out


class LIFNetwork:
    """Trials of the differentiable LIF network on the device (see module doc)."""

    def __init__(self, trials=256, seed=SEED, config=None, weight_dtype="float", ckpt_every=100):
        torch = _torch()
        self.c = dict(LIF if config is None else config)
        c = self.c
        self.B, self.N, self.n_in, self.T = int(trials), c["N"], c["n_in"], c["n_steps"]
        self.R = c["n_readout"]
        if self.N > BLOCK or self.n_in > BLOCK:
            raise NotImplementedError(f"groups of more than {BLOCK} neurons")
        if self.T % ckpt_every:
            raise ValueError("n_steps must be a multiple of ckpt_every")
        self.C = int(ckpt_every)
        self.S = self.T // self.C
        self.seed = int(seed)
        self.dt = 1e-4
        self.wt = weight_dtype
        self._kernels()
        inp = lif_inputs(self.B, seed, c)
        self.inp = inp
        self.D = Device.get()
        dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to("cuda")  # noqa: E731
        self.r_dt = dev(inp["rates"] * self.dt, np.float64)
        self.in_ptr = dev(_csr_ptr(inp["pre_in"], self.n_in).view(np.int32), np.int32)
        self.in_post = dev(inp["post_in"], np.int32)
        self.rec_ptr = dev(_csr_ptr(inp["pre_rec"], self.N).view(np.int32), np.int32)
        self.rec_post = dev(inp["post_rec"], np.int32)
        self.n_syn_in, self.n_syn_rec = len(inp["pre_in"]), len(inp["pre_rec"])
        npw = np.float32 if weight_dtype == "float" else np.float64
        self.w_in0 = dev(inp["w_in"].astype(npw), npw)
        self.w_rec0 = dev(inp["w_rec"].astype(npw), npw)
        self.v0 = dev(inp["v0"], np.float64)
        self.targets = dev(inp["targets"], np.float64)
        self._fwd_fn = self.D.function(self.fwd_source, "lif_forward")
        self._bwd_fn = self.D.function(self.bwd_source, "lif_backward")

    # ------------------------------------------------------------------ kernels
    def _kernels(self):
        """Loads the model's statements and generates both kernel sources."""
        self.model = Model.load("lif")
        self.su, self.th, self.rs = (self.model.group[k] for k in ("stateupdate", "threshold", "reset"))
        th = [(st["var"], st["expr"]) for st in self.th.vector]
        rs = [(st["var"], st["op"], st["expr"]) for st in self.rs.vector]
        if th != [("_cond", "v > theta")] or rs != [("v", "=", "0.0")]:
            raise NotImplementedError("threshold 'v > theta' and reset 'v = 0' expected")
        self.su_code = self._su_lines()
        self.fwd_source = self._fwd_source()
        self.bwd_source = self._bwd_source()

    @classmethod
    def sources(cls, config=None, weight_dtype="float"):
        """The generated kernel sources (no device needed)."""
        self = cls.__new__(cls)
        self.c = dict(LIF if config is None else config)
        self.N, self.dt, self.wt = self.c["N"], 1e-4, weight_dtype
        self._kernels()
        return self.fwd_source, self.bwd_source

    def _su_lines(self):
        types = {"v": "double", "I": "double", "t": "double", "dt": "double", "N": "int", "i": "int", "_idx": "int"}
        names = {}
        for n in ("_lio_1", "_lio_2", "_lio_3", "_lio_4"):
            types[n] = "double"
        sc = scalar_values(self.su, {"dt": self.dt, "N": self.N, **{k: self.c[k] for k in ("tau_m", "tau_s", "bias")}},
                           "double")
        self.su_scalars = sorted(sc)
        for n in self.su_scalars:
            types[n] = sc[n][1]
            names[n] = f"p.su_{n}"
        lines = translate_block(self.su.vector, self.su.conditional_write, types, set(), "double")
        return _rename(lines, names)

    def _struct(self, name, backward):
        WT = self.wt
        ps = ParamStruct(name)
        ps.val("k0", "long long")
        ps.val("n_steps", "int")
        ps.val("T_total", "int")
        ps.val("seed", "unsigned long long")
        ps.val("n_in", "int")
        ps.val("N", "int")
        ps.val("R", "int")
        ps.val("trial0", "int")
        ps.val("theta", "double")
        ps.val("dt", "double")
        for n in self.su_scalars:
            ps.val(f"su_{n}", "double")
        ps.ptr("r_dt", "double", const=True)
        ps.ptr("in_ptr", "unsigned int", const=True)
        ps.ptr("in_post", "int", const=True)
        ps.ptr("rec_ptr", "unsigned int", const=True)
        ps.ptr("rec_post", "int", const=True)
        ps.ptr("hist_v", "double")
        ps.ptr("hist_I", "double")
        ps.ptr("fs_step", "int")
        if not backward:
            ps.ptr("in_w", WT, const=True)
            ps.ptr("rec_w", WT, const=True)
            ps.ptr("v", "double")
            ps.ptr("I", "double")
            ps.ptr("ckpt_v", "double")
            ps.ptr("ckpt_I", "double")
            ps.ptr("fs_vprev", "double")
            ps.ptr("fs_u", "double")
            ps.ptr("l2", "double")
            ps.ptr("spk_i", "int")
            ps.ptr("spk_step", "int")
            ps.ptr("spk_n", "long long")
            ps.val("spk_cap", "long long")
        else:
            ps.val("A", "double")
            ps.val("B", "double")
            ps.val("beta", "double")
            ps.val("coef_fs", "double")
            ps.val("coef_v", "double")
            ps.ptr("targets", "double", const=True)
            ps.ptr("lam_v", "double")
            ps.ptr("lam_I", "double")
            ps.ptr("g_in", "double")
            ps.ptr("g_rec", "double")
            ps.ptr("gacc", "double")
        return ps

    def _common(self, ps):
        NW = BLOCK // 64
        k = [source_header(BLOCK), _rng_header(), ps.c_decl()]
        return k, self.su_code, NW

    def _compaction(self, cond, arr, count):
        """Ordered compaction of `cond` over the threads into LDS `arr`, total in `count`."""
        NW = BLOCK // 64
        return [
            "        {",
            f"            const unsigned long long bal_ = __ballot({cond});",
            "            if (lane == 0) s_w[wid] = __popcll(bal_);",
            "            __syncthreads();",
            "            if (threadIdx.x == 0) {",
            "                int acc_ = 0;",
            f"                for (int w_ = 0; w_ < {NW}; w_++) {{ const int c_ = s_w[w_]; s_w[w_] = acc_; acc_ += c_; }}",
            f"                s_w[{NW}] = acc_;",
            "            }",
            "            __syncthreads();",
            f"            if ({cond}) {arr}[s_w[wid] + __popcll(bal_ & ((1ull << lane) - 1ull))] = (int)threadIdx.x;",
            f"            {count} = s_w[{NW}];",
            "            __syncthreads();",
            "        }",
        ]

    def _fwd_source(self):
        ps = self._struct("hba_fwd", False)
        self.fwd_struct = ps
        k, su, NW = self._common(ps)
        T = BLOCK
        k += [
            f'extern "C" __global__ void __launch_bounds__({T}) lif_forward(const hba_fwd p) {{',
            "    const int b = blockIdx.x;",
            "    const int j = threadIdx.x;",
            "    const int lane = j & 63, wid = j >> 6;",
            "    const int N = p.N;",
            "    const bool act = j < N;",
            "    const double dt = p.dt;",
            f"    __shared__ double s_I[{T}];",
            f"    __shared__ int s_spk[{T}];",
            f"    __shared__ int s_inp[{T}];",
            f"    __shared__ int s_w[{NW + 1}];",
            f"    __shared__ double s_red[{NW}];",
            "    const size_t bo = (size_t)b * N;",
            "    double v = act ? p.v[bo + j] : 0.0;",
            "    double I = act ? p.I[bo + j] : 0.0;",
            "    if (p.ckpt_v && act) {",
            "        const size_t co = ((size_t)(p.k0 / p.n_steps) * gridDim.x + b) * N + j;",
            "        p.ckpt_v[co] = v; p.ckpt_I[co] = I;",
            "    }",
            "    const unsigned long long rng_base = ((unsigned long long)(p.trial0 + b) * p.n_in + j) * (unsigned long long)p.T_total;",
            "    double l2 = 0.0;",
            "    for (int s = 0; s < p.n_steps; s++) {",
            "        const long long k = p.k0 + s;",
            "        const double t = (double)k * dt;",
            "        (void)t;",
            "        if (p.hist_v && act) {",
            "            const size_t ho = ((size_t)b * p.n_steps + s) * N + j;",
            "            p.hist_v[ho] = v; p.hist_I[ho] = I;",
            "        }",
            "        const double v_prev = v;",
            "        const int i = j, _idx = j;",
            "        (void)i; (void)_idx;",
            "        if (act) { // stateupdate (Brian's statements)",
        ]
        k += ["            " + x for x in su]
        k += [
            "        }",
            "        const bool cond = act && (v > p.theta);",
            "        const bool cin = j < p.n_in && hb_uniform01(hb_rand64(p.seed, "
            f"{S_POISSON}ull, rng_base + (unsigned long long)k)) < p.r_dt[j < p.n_in ? j : 0];",
            "        int nspk, ninp;",
        ]
        k += self._compaction("cond", "s_spk", "nspk")
        k += self._compaction("cin", "s_inp", "ninp")
        k += [
            "        if (cond && j < p.R && p.fs_step && p.fs_step[b * p.R + j] < 0) {",
            "            p.fs_step[b * p.R + j] = (int)k;",
            "            if (p.fs_vprev) { p.fs_vprev[b * p.R + j] = v_prev; p.fs_u[b * p.R + j] = v; }",
            "        }",
            "        // propagation, in schedule order: syn_in_pre, then syn_rec_pre; each a left fold",
            "        // in (source, synapse) order",
            "        s_I[j] = I;",
            "        __syncthreads();",
            "        for (int q = 0; q < ninp; q++) {",
            "            const int src = s_inp[q];",
            "            const unsigned int a = p.in_ptr[src], e_ = p.in_ptr[src + 1];",
            f"            for (unsigned int e = a + j; e < e_; e += {T}) s_I[p.in_post[e]] += (double)p.in_w[e];",
            "            __syncthreads();",
            "        }",
            "        for (int q = 0; q < nspk; q++) {",
            "            const int src = s_spk[q];",
            "            const unsigned int a = p.rec_ptr[src], e_ = p.rec_ptr[src + 1];",
            f"            for (unsigned int e = a + j; e < e_; e += {T}) s_I[p.rec_post[e]] += (double)p.rec_w[e];",
            "            __syncthreads();",
            "        }",
            "        I = s_I[j];",
            "        if (cond) v = 0.0; // reset",
            "        l2 += v * v;",
            "        if (p.spk_i && nspk) {",
            "            const long long base = p.spk_n[b];",
            f"            for (int q = j; q < nspk; q += {T})",
            "                if (base + q < p.spk_cap) { p.spk_i[(size_t)b * p.spk_cap + base + q] = s_spk[q]; "
            "p.spk_step[(size_t)b * p.spk_cap + base + q] = (int)k; }",
            "            __syncthreads();",
            "            if (j == 0) p.spk_n[b] = base + nspk;",
            "        }",
            "        __syncthreads();",
            "    }",
            "    if (act) { p.v[bo + j] = v; p.I[bo + j] = I; }",
            "    if (p.l2) {",
            "        // fixed-order reduction: wave shuffles, then the waves in order",
            "        double x = act ? l2 : 0.0;",
            "        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);",
            "        if (lane == 0) s_red[wid] = x;",
            "        __syncthreads();",
            f"        if (j == 0) {{ double acc = 0.0; for (int w_ = 0; w_ < {NW}; w_++) acc += s_red[w_]; p.l2[b] += acc; }}",
            "    }",
            "}",
        ]
        return "\n".join(k) + "\n"

    def _bwd_source(self):
        ps = self._struct("hba_bwd", True)
        self.bwd_struct = ps
        k, su, NW = self._common(ps)
        T = BLOCK
        k += [
            f'extern "C" __global__ void __launch_bounds__({T}) lif_backward(const hba_bwd p) {{',
            "    const int b = blockIdx.x;",
            "    const int j = threadIdx.x;",
            "    const int lane = j & 63, wid = j >> 6;",
            "    const int N = p.N;",
            "    const bool act = j < N;",
            "    const double dt = p.dt;",
            f"    __shared__ double s_lam[{T}];",
            f"    __shared__ int s_spk[{T}];",
            f"    __shared__ int s_inp[{T}];",
            f"    __shared__ int s_w[{NW + 1}];",
            f"    __shared__ double s_red[{NW}][4];",
            "    const size_t bo = (size_t)b * N;",
            "    double lam_v = act ? p.lam_v[bo + j] : 0.0;",
            "    double lam_I = act ? p.lam_I[bo + j] : 0.0;",
            "    double gA = 0.0, gB = 0.0, gC = 0.0, gbeta = 0.0;",
            "    const int fs_k = (j < p.R) ? p.fs_step[b * p.R + j] : -1;",
            "    const unsigned long long rng_base = ((unsigned long long)(p.trial0 + b) * p.n_in + j) * (unsigned long long)p.T_total;",
            "    double* g_in = p.g_in + (size_t)b * p.in_ptr[p.n_in];",
            "    double* g_rec = p.g_rec + (size_t)b * p.rec_ptr[N];",
            "    for (int s = p.n_steps - 1; s >= 0; s--) {",
            "        const long long k = p.k0 + s;",
            "        const double t = (double)k * dt;",
            "        (void)t;",
            "        const size_t ho = ((size_t)b * p.n_steps + s) * N + j;",
            "        const double v_prev = act ? p.hist_v[ho] : 0.0;",
            "        const double I_prev = act ? p.hist_I[ho] : 0.0;",
            "        double v = v_prev, I = I_prev;",
            "        const int i = j, _idx = j;",
            "        (void)i; (void)_idx;",
            "        if (act) { // the step's update again (identical statements and values)",
        ]
        k += ["            " + x for x in su]
        k += [
            "        }",
            "        const double u = v;",
            "        const bool cond = act && (u > p.theta);",
            "        const bool cin = j < p.n_in && hb_uniform01(hb_rand64(p.seed, "
            f"{S_POISSON}ull, rng_base + (unsigned long long)k)) < p.r_dt[j < p.n_in ? j : 0];",
            "        const double v_post = cond ? 0.0 : u;",
            "        double dLdu = cond ? 0.0 : (lam_v + p.coef_v * 2.0 * v_post);",
            "        double extra = 0.0;",
            "        if (fs_k == (int)k) { // first spike of a readout: interpolated crossing time",
            "            const double d = u - v_prev;",
            "            const double tc = dt * ((double)k + (p.theta - v_prev) / d);",
            "            const double dLdt = p.coef_fs * 2.0 * (tc - p.targets[j]);",
            "            dLdu += dLdt * (-dt * (p.theta - v_prev) / (d * d));",
            "            extra = dLdt * (dt * (p.theta - u) / (d * d));",
            "        }",
            "        int nspk, ninp;",
        ]
        k += self._compaction("cond", "s_spk", "nspk")
        k += self._compaction("cin", "s_inp", "ninp")
        k += [
            "        // event-based weight gradients: lam of I after this step at every target of a source",
            "        // that spiked in it (rows of different sources never share a synapse)",
            "        s_lam[j] = lam_I;",
            "        __syncthreads();",
            "        for (int q = 0; q < ninp; q++) {",
            "            const int src = s_inp[q];",
            "            const unsigned int a = p.in_ptr[src], e_ = p.in_ptr[src + 1];",
            f"            for (unsigned int e = a + j; e < e_; e += {T}) g_in[e] += s_lam[p.in_post[e]];",
            "        }",
            "        for (int q = 0; q < nspk; q++) {",
            "            const int src = s_spk[q];",
            "            const unsigned int a = p.rec_ptr[src], e_ = p.rec_ptr[src + 1];",
            f"            for (unsigned int e = a + j; e < e_; e += {T}) g_rec[e] += s_lam[p.rec_post[e]];",
            "        }",
            "        const double lamJ = lam_I;",
            "        gA += dLdu * v_prev;",
            "        gB += dLdu * I_prev;",
            "        gC += dLdu;",
            "        gbeta += lamJ * I_prev;",
            "        lam_v = p.A * dLdu + extra;",
            "        lam_I = p.B * dLdu + p.beta * lamJ;",
            "        __syncthreads();",
            "    }",
            "    if (act) { p.lam_v[bo + j] = lam_v; p.lam_I[bo + j] = lam_I; }",
            "    double x[4] = {act ? gA : 0.0, act ? gB : 0.0, act ? gC : 0.0, act ? gbeta : 0.0};",
            "    for (int c = 0; c < 4; c++) {",
            "        for (int o = 32; o > 0; o >>= 1) x[c] += __shfl_down(x[c], o, 64);",
            "        if (lane == 0) s_red[wid][c] = x[c];",
            "    }",
            "    __syncthreads();",
            "    if (j < 4) {",
            "        double acc = 0.0;",
            f"        for (int w_ = 0; w_ < {NW}; w_++) acc += s_red[w_][j];",
            "        p.gacc[b * 4 + j] += acc;",
            "    }",
            "}",
        ]
        return "\n".join(k) + "\n"

    # ------------------------------------------------------------------ host side
    def _scalars(self, tau_m, tau_s):
        consts = {"tau_m": tau_m, "tau_s": tau_s, "bias": self.c["bias"], "theta": self.c["theta"]}
        sc = scalar_values(self.su, {"dt": self.dt, "N": self.N, **consts}, "double")
        return {n: sc[n][0] for n in self.su_scalars}, consts

    def _launch(self, fn, cls, vals):
        a = cls()
        for name, _ in cls._fields_:
            if name in vals:
                setattr(a, name, vals[name])
        buf = bytes(a)
        stream = ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)
        _lib.check(L.hb_launch(fn, self.B, BLOCK, 0, stream, buf, len(buf)), "lif launch")

    def _base_vals(self, sc):
        v = dict(T_total=self.T, seed=self.seed, n_in=self.n_in, N=self.N, R=self.R, trial0=0,
                 theta=self.c["theta"], dt=self.dt, r_dt=self.r_dt.data_ptr(), in_ptr=self.in_ptr.data_ptr(),
                 in_post=self.in_post.data_ptr(), rec_ptr=self.rec_ptr.data_ptr(), rec_post=self.rec_post.data_ptr())
        for n, x in sc.items():
            v[f"su_{n}"] = x
        return v

    def forward(self, w_in, w_rec, tau_m, tau_s, record_spikes=False, spike_cap=200000):
        """Runs all trials; returns (loss, saved) where saved holds what the backward
        pass needs (checkpoints, first-spike records)."""
        torch = _torch()
        w_in = w_in.detach().to(self.w_in0.dtype).contiguous()
        w_rec = w_rec.detach().to(self.w_rec0.dtype).contiguous()
        sc, _ = self._scalars(tau_m, tau_s)
        B, N, R = self.B, self.N, self.R
        v = self.v0.clone()
        I = torch.zeros_like(v)
        ckpt_v = torch.empty((self.S, B, N), dtype=torch.float64, device="cuda")
        ckpt_I = torch.empty_like(ckpt_v)
        fs_step = torch.full((B, R), -1, dtype=torch.int32, device="cuda")
        fs_vprev = torch.zeros((B, R), dtype=torch.float64, device="cuda")
        fs_u = torch.zeros_like(fs_vprev)
        l2 = torch.zeros(B, dtype=torch.float64, device="cuda")
        vals = self._base_vals(sc)
        vals.update(n_steps=self.C, in_w=w_in.data_ptr(), rec_w=w_rec.data_ptr(), v=v.data_ptr(), I=I.data_ptr(),
                    ckpt_v=ckpt_v.data_ptr(), ckpt_I=ckpt_I.data_ptr(), fs_step=fs_step.data_ptr(),
                    fs_vprev=fs_vprev.data_ptr(), fs_u=fs_u.data_ptr(), l2=l2.data_ptr())
        if record_spikes:
            spk_i = torch.zeros((B, spike_cap), dtype=torch.int32, device="cuda")
            spk_step = torch.zeros_like(spk_i)
            spk_n = torch.zeros(B, dtype=torch.int64, device="cuda")
            vals.update(spk_i=spk_i.data_ptr(), spk_step=spk_step.data_ptr(), spk_n=spk_n.data_ptr(),
                        spk_cap=spike_cap)
        cls = self.fwd_struct.ctypes_struct()
        for sg in range(self.S):
            vals["k0"] = sg * self.C
            self._launch(self._fwd_fn, cls, vals)
        loss, parts = self._loss(fs_step, fs_vprev, fs_u, l2)
        saved = dict(w_in=w_in, w_rec=w_rec, tau_m=tau_m, tau_s=tau_s, ckpt_v=ckpt_v, ckpt_I=ckpt_I,
                     fs_step=fs_step, fs_vprev=fs_vprev, fs_u=fs_u, v=v, I=I, parts=parts)
        if record_spikes:
            n = spk_n.cpu().numpy()
            if (n > spike_cap).any():
                raise _lib.HipBrianError("spike record buffer overflow")
            si, ss = spk_i.cpu().numpy(), spk_step.cpu().numpy()
            saved["spikes"] = [(si[b, :n[b]].copy(), ss[b, :n[b]].astype(np.int64)) for b in range(B)]
        return loss, saved

    def _loss(self, fs_step, fs_vprev, fs_u, l2):
        c, B, N, T = self.c, self.B, self.N, self.T
        k = fs_step.cpu().numpy().astype(np.float64)
        fv, fu = fs_vprev.cpu().numpy(), fs_u.cpu().numpy()
        spiked = k >= 0
        with np.errstate(divide="ignore", invalid="ignore"):
            tc = np.where(spiked, self.dt * (k + (c["theta"] - fv) / (fu - fv)), T * self.dt)
        tgt = self.targets.cpu().numpy()
        fs_term = (((tc - tgt[None, :]) / c["t_ref"]) ** 2).sum() / B
        v_term = c["c_v"] / (N * T * c["theta"] ** 2) * float(l2.sum().item()) / B
        return fs_term + v_term, dict(first_spike=tc, fs_term=fs_term, v_term=v_term)

    def backward(self, saved, grad_out=1.0):
        """Gradients of the loss with respect to (w_in, w_rec, tau_m, tau_s)."""
        torch = _torch()
        c, B, N = self.c, self.B, self.N
        tau_m, tau_s = saved["tau_m"], saved["tau_s"]
        sc, consts = self._scalars(tau_m, tau_s)
        lin = linear_map(self.su, consts, self.dt)
        hist_v = torch.empty((B, self.C, N), dtype=torch.float64, device="cuda")
        hist_I = torch.empty_like(hist_v)
        lam_v = torch.zeros((B, N), dtype=torch.float64, device="cuda")
        lam_I = torch.zeros_like(lam_v)
        g_in = torch.zeros((B, self.n_syn_in), dtype=torch.float64, device="cuda")
        g_rec = torch.zeros((B, self.n_syn_rec), dtype=torch.float64, device="cuda")
        gacc = torch.zeros((B, 4), dtype=torch.float64, device="cuda")
        v = torch.empty((B, N), dtype=torch.float64, device="cuda")
        I = torch.empty_like(v)
        fvals = self._base_vals(sc)
        fvals.update(n_steps=self.C, in_w=saved["w_in"].data_ptr(), rec_w=saved["w_rec"].data_ptr(),
                     v=v.data_ptr(), I=I.data_ptr(), hist_v=hist_v.data_ptr(), hist_I=hist_I.data_ptr())
        bvals = self._base_vals(sc)
        bvals.update(n_steps=self.C, hist_v=hist_v.data_ptr(), hist_I=hist_I.data_ptr(),
                     fs_step=saved["fs_step"].data_ptr(), A=lin["A"], B=lin["B"], beta=lin["beta"],
                     coef_fs=grad_out / (B * c["t_ref"] ** 2),
                     coef_v=grad_out * c["c_v"] / (N * self.T * c["theta"] ** 2) / B,
                     targets=self.targets.data_ptr(), lam_v=lam_v.data_ptr(), lam_I=lam_I.data_ptr(),
                     g_in=g_in.data_ptr(), g_rec=g_rec.data_ptr(), gacc=gacc.data_ptr())
        fcls, bcls = self.fwd_struct.ctypes_struct(), self.bwd_struct.ctypes_struct()
        for sg in reversed(range(self.S)):
            v.copy_(saved["ckpt_v"][sg])
            I.copy_(saved["ckpt_I"][sg])
            fvals["k0"] = bvals["k0"] = sg * self.C
            self._launch(self._fwd_fn, fcls, fvals)
            self._launch(self._bwd_fn, bcls, bvals)
        gw_in = g_in.sum(0)
        gw_rec = g_rec.sum(0)
        ga = gacc.sum(0).cpu().numpy()
        g_coef = {"A": ga[0], "B": ga[1], "C": ga[2], "beta": ga[3]}
        g_tau_m = sum(g_coef[k] * lin[f"d{k}"][0] for k in g_coef)
        g_tau_s = sum(g_coef[k] * lin[f"d{k}"][1] for k in g_coef)
        return gw_in, gw_rec, g_tau_m, g_tau_s


def _make_function():
    torch = _torch()

    class LIFLoss(torch.autograd.Function):
        """loss = LIFLoss.apply(net, w_in, w_rec, tau_m, tau_s): forward on the device,
        backward by the exact adjoint (weights and time constants)."""

        @staticmethod
        def forward(ctx, net, w_in, w_rec, tau_m, tau_s):
            loss, saved = net.forward(w_in, w_rec, float(tau_m), float(tau_s))
            ctx.net, ctx.saved = net, saved
            ctx.dtypes = (w_in.dtype, w_rec.dtype, tau_m.dtype, tau_s.dtype)
            return torch.tensor(loss, dtype=torch.float64, device=w_in.device)

        @staticmethod
        def backward(ctx, grad):
            g_in, g_rec, g_m, g_s = ctx.net.backward(ctx.saved, float(grad))
            d = ctx.dtypes
            return (None, g_in.to(d[0]), g_rec.to(d[1]), torch.tensor(g_m, dtype=d[2], device=grad.device),
                    torch.tensor(g_s, dtype=d[3], device=grad.device))

    return LIFLoss


_FN = None


def lif_loss(net, w_in, w_rec, tau_m, tau_s):
    """Differentiable loss of the LIF network (see module doc)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(net, w_in, w_rec, tau_m, tau_s)
