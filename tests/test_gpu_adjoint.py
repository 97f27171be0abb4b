"""The differentiable LIF network on the device: forward parity with Brian, the
adjoint against autograd of the unrolled tensor reference, and against fp64
central finite differences at perturbations that create or delete no spike."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _net(trials=2, wt="float"):
    from brian2_amd.adjoint import LIFNetwork

    return LIFNetwork(trials=trials, weight_dtype=wt)


def test_forward_bitexact_vs_brian():
    z, meta = O.load_golden("lif")
    net = _net(meta["trials"])
    loss, saved = net.forward(net.w_in0, net.w_rec0, net.c["tau_m"], net.c["tau_s"], record_spikes=True)
    for b in range(meta["trials"]):
        i, s = saved["spikes"][b]
        sel = z["trial"] == b
        assert len(i) == sel.sum()
        assert np.array_equal(i, z["i"][sel]) and np.array_equal(s, z["step"][sel])
        assert np.array_equal(saved["v"][b].cpu().numpy(), z["final"][b][0])
        assert np.array_equal(saved["I"][b].cpu().numpy(), z["final"][b][1])
    assert np.isfinite(loss)


def _reference(net, trials):
    import torch

    from lif_reference import reference_loss

    inp = O.lif_inputs(trials)
    sc, _ = net._scalars(net.c["tau_m"], net.c["tau_s"])
    lio = tuple(sc[f"_lio_{k}"] for k in (1, 2, 3, 4))
    w_in = torch.tensor(net.w_in0.cpu().numpy().astype(np.float64), requires_grad=True)
    w_rec = torch.tensor(net.w_rec0.cpu().numpy().astype(np.float64), requires_grad=True)
    tm = torch.tensor(net.c["tau_m"], dtype=torch.float64, requires_grad=True)
    ts = torch.tensor(net.c["tau_s"], dtype=torch.float64, requires_grad=True)
    loss, fs, raster = reference_loss(inp, inp["trains"], inp["v0"], w_in, w_rec, tm, ts, net.c, lio_values=lio)
    loss.backward()
    return float(loss), w_in.grad.numpy(), w_rec.grad.numpy(), float(tm.grad), float(ts.grad), raster


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_adjoint_equals_autograd_of_the_tensor_reference():
    import torch

    from brian2_amd.adjoint import lif_loss

    net = _net(2, "double")
    w_in = net.w_in0.clone().requires_grad_(True)
    w_rec = net.w_rec0.clone().requires_grad_(True)
    tm = torch.tensor(net.c["tau_m"], dtype=torch.float64, requires_grad=True)
    ts = torch.tensor(net.c["tau_s"], dtype=torch.float64, requires_grad=True)
    loss = lif_loss(net, w_in, w_rec, tm, ts)
    loss.backward()
    ref_loss, g_in, g_rec, g_m, g_s, _ = _reference(net, 2)
    print(f"loss {float(loss)!r} reference {ref_loss!r}; dL/dtau_m {float(tm.grad)!r} vs {g_m!r}; "
          f"dL/dtau_s {float(ts.grad)!r} vs {g_s!r}", flush=True)
    assert abs(float(loss) - ref_loss) <= 1e-11 * abs(ref_loss)
    assert _rel(w_in.grad.cpu().numpy(), g_in) < 1e-9
    assert _rel(w_rec.grad.cpu().numpy(), g_rec) < 1e-9
    assert abs(float(tm.grad) - g_m) <= 1e-9 * abs(g_m)
    assert abs(float(ts.grad) - g_s) <= 1e-9 * abs(g_s)


def test_adjoint_equals_central_finite_differences():
    """fp64 central differences at perturbations that leave every spike in place:
    for each parameter the largest step of a short list (relative 1e-5 .. 1e-7 for
    the time constants, 1e-2 .. 1e-4 sigma for weights) whose +h and -h runs
    reproduce the unperturbed spike raster exactly."""
    net = _net(2, "double")
    w_in, w_rec = net.w_in0.clone(), net.w_rec0.clone()
    tm, ts = net.c["tau_m"], net.c["tau_s"]
    loss0, saved = net.forward(w_in, w_rec, tm, ts, record_spikes=True)
    g_in, g_rec, g_m, g_s = net.backward(saved)
    g_in, g_rec = g_in.cpu().numpy(), g_rec.cpu().numpy()
    ref = saved["spikes"]

    def L(wi, wr, m, s):
        loss, sv = net.forward(wi, wr, m, s, record_spikes=True)
        same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(sv["spikes"], ref))
        return loss, same

    def fd(make, steps):
        for h in steps:
            (lp, sp), (lm, sm) = L(*make(h)), L(*make(-h))
            if sp and sm:
                return (lp - lm) / (2 * h), h
        return None, None

    checks = []
    checks.append(("tau_m",) + fd(lambda h: (w_in, w_rec, tm + h, ts), [r * tm for r in (1e-5, 1e-6, 1e-7)]) + (g_m,))
    checks.append(("tau_s",) + fd(lambda h: (w_in, w_rec, tm, ts + h), [r * ts for r in (1e-5, 1e-6, 1e-7)]) + (g_s,))
    sig = net.inp["sigma"]
    for name, g, which in (("w_in", g_in, 0), ("w_rec", g_rec, 1)):
        for e in np.argsort(-np.abs(g))[:3]:
            def make(h, e=int(e), which=which):
                w = [w_in.clone(), w_rec.clone()]
                w[which][e] += h
                return w[0], w[1], tm, ts
            checks.append((f"{name}[{e}]",) + fd(make, [r * sig for r in (1e-2, 1e-3, 1e-4)]) + (float(g[e]),))
    for name, d, h, adj in checks:
        print(f"{name}: finite difference {d!r} (step {h!r}), adjoint {adj!r}", flush=True)
        assert d is not None, f"{name}: every step changed a spike"
        assert abs(d - adj) <= 1e-6 * abs(adj), name


def test_full_configuration_is_deterministic():
    """256 trials, fp32 weights: forward + backward twice give identical results."""
    import torch

    net = _net(256, "float")
    outs = []
    for _ in range(2):
        loss, saved = net.forward(net.w_in0, net.w_rec0, net.c["tau_m"], net.c["tau_s"])
        g = net.backward(saved)
        outs.append((loss, g[0].cpu().numpy(), g[1].cpu().numpy(), g[2], g[3]))
        torch.cuda.synchronize()
    a, b = outs
    assert np.isfinite(a[0]) and a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and a[3] == b[3] and a[4] == b[4]
    assert np.isfinite(a[1]).all() and np.abs(a[2]).max() > 0
