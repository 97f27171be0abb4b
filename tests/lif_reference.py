"""Unrolled pure-tensor reference of the differentiable LIF network (test
infrastructure): the same discrete forward map and loss as
brian2_amd/adjoint.py, written with dense torch fp64 operations on the CPU, and
differentiated by autograd.  Spikes are masks without gradient (as in the
adjoint); the summation order of the synaptic inputs differs from the device's
sequential fold, so values agree to rounding, not bit for bit.
"""
import numpy as np
import torch


def reference_loss(inp, trains, v0, w_in, w_rec, tau_m, tau_s, c, lio_values=None):
    """inp: the generator's inputs (pre/post of both pathways, targets);
    trains: per trial (input index, step) arrays; v0: [B, N]; w_in, w_rec, tau_m,
    tau_s: torch fp64 tensors (leaves with requires_grad for the gradients);
    lio_values: optional exact values of the loop invariants (the host's hb_exp
    values), substituted without changing their gradient.  Returns (loss,
    first-spike times [B, R], spike raster [T, B, N] as bool)."""
    dt = 1e-4
    B, N = v0.shape
    n_in, T, R = c["n_in"], c["n_steps"], c["n_readout"]
    theta, bias = c["theta"], c["bias"]
    dt_t = torch.tensor(dt, dtype=torch.float64)
    lio1 = torch.exp(-dt_t / tau_s)
    lio2 = bias - bias * torch.exp(-dt_t / tau_m)
    lio3 = (((tau_s * ((-torch.exp(dt_t / tau_m)) + torch.exp(dt_t / tau_s))) * torch.exp(-dt_t / tau_m))
            * torch.exp(-dt_t / tau_s)) / (tau_m - tau_s)
    lio4 = torch.exp(-dt_t / tau_m)
    if lio_values is not None:
        lio1, lio2, lio3, lio4 = (x + (torch.tensor(val, dtype=torch.float64) - x).detach()
                                  for x, val in zip((lio1, lio2, lio3, lio4), lio_values))
    W_in = torch.zeros((n_in, N), dtype=torch.float64).index_put(
        (torch.from_numpy(inp["pre_in"]).long(), torch.from_numpy(inp["post_in"]).long()), w_in)
    W_rec = torch.zeros((N, N), dtype=torch.float64).index_put(
        (torch.from_numpy(inp["pre_rec"]).long(), torch.from_numpy(inp["post_rec"]).long()), w_rec)
    S_in = np.zeros((T, B, n_in))
    for b, (ii, kk) in enumerate(trains):
        S_in[kk, b, ii] = 1.0
    S_in = torch.from_numpy(S_in)
    v = torch.from_numpy(np.ascontiguousarray(v0, np.float64))
    I = torch.zeros((B, N), dtype=torch.float64)
    first = [[None] * R for _ in range(B)]
    l2 = torch.zeros((), dtype=torch.float64)
    raster = np.zeros((T, B, N), bool)
    for k in range(T):
        I_new = lio1 * I
        u = lio2 + ((lio3 * I) + (lio4 * v))
        s = (u > theta).detach()
        raster[k] = s.numpy()
        sd = s.double()
        for b in range(B):
            for r in range(R):
                if first[b][r] is None and bool(s[b, r]):
                    first[b][r] = dt * (k + (theta - v[b, r]) / (u[b, r] - v[b, r]))
        I = I_new + S_in[k] @ W_in + sd @ W_rec
        v = (1.0 - sd) * u
        l2 = l2 + (v * v).sum()
    tgt = torch.from_numpy(inp["targets"])
    fs = torch.stack([torch.stack([first[b][r] if first[b][r] is not None
                                   else torch.tensor(T * dt, dtype=torch.float64) for r in range(R)])
                      for b in range(B)])
    loss = (((fs - tgt[None, :]) / c["t_ref"]) ** 2).sum() / B + c["c_v"] / (N * T * theta ** 2) * l2 / B
    return loss, fs, raster
