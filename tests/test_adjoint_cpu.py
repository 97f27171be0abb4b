"""The differentiable LIF configuration on the CPU: the product's input
generator against the oracle's, the linear-map coefficients against their
closed forms, the generated kernels compile for gfx950, and the unrolled
tensor reference against Brian's golden forward pass."""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_py as O


def test_inputs_match_the_oracle_generator():
    from brian2_amd import adjoint

    a = adjoint.lif_inputs(3)
    o = O.lif_inputs(3)
    for k in ("rates", "pre_in", "post_in", "pre_rec", "post_rec", "v0", "targets"):
        assert np.array_equal(a[k], o[k]), k
    assert np.array_equal(a["w_in"].astype(np.float32), o["w_in"])
    assert np.array_equal(a["w_rec"].astype(np.float32), o["w_rec"])


def test_linear_map_coefficients_and_derivatives():
    from brian2_amd import adjoint
    from brian2_amd.runtime import Model

    su = Model.load("lif").group["stateupdate"]
    c = adjoint.LIF
    tm, ts, bias, dt = c["tau_m"], c["tau_s"], c["bias"], 1e-4
    lin = adjoint.linear_map(su, {"tau_m": tm, "tau_s": ts, "bias": bias, "theta": c["theta"]}, dt)
    A = math.exp(-dt / tm)
    beta = math.exp(-dt / ts)
    Bc = ts / (ts - tm) * (beta - A)
    C = bias * (1 - A)
    for k, x in (("A", A), ("B", Bc), ("C", C), ("beta", beta)):
        assert math.isclose(lin[k], x, rel_tol=1e-13), k
    dA = A * dt / tm ** 2
    dbeta = beta * dt / ts ** 2
    dB_m = ts / (ts - tm) ** 2 * (beta - A) - ts / (ts - tm) * dA
    dB_s = -tm / (ts - tm) ** 2 * (beta - A) + ts / (ts - tm) * dbeta
    exp = {"A": (dA, 0.0), "beta": (0.0, dbeta), "B": (dB_m, dB_s), "C": (-bias * dA, 0.0)}
    for k, (gm, gs) in exp.items():
        assert math.isclose(lin["d" + k][0], gm, rel_tol=1e-10, abs_tol=1e-300), k
        assert math.isclose(lin["d" + k][1], gs, rel_tol=1e-10, abs_tol=1e-300), k


@pytest.mark.parametrize("wt", ["float", "double"])
def test_adjoint_kernels_compile_for_gfx950(wt):
    from brian2_amd.adjoint import LIFNetwork

    for src in LIFNetwork.sources(weight_dtype=wt):
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as d:
            path = os.path.join(d, "k.hip")
            with open(path, "w") as f:
                f.write("#include <hip/hip_runtime.h>\n" + src)
            r = subprocess.run(["hipcc", "--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-c", path, "-o",
                                os.path.join(d, "k.o")], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-3000:]
        assert "atomicAdd(&p.g_" not in src and "atomicAdd(g_" not in src


def test_tensor_reference_forward_equals_brian():
    """The unrolled reference used for the gradient check reproduces Brian's
    forward pass: identical spikes, final state within 1e-12."""
    import torch

    from lif_reference import reference_loss

    z, meta = O.load_golden("lif")
    inp = O.lif_inputs(meta["trials"], meta["seed"])
    c = dict(O.LIF, t_ref=0.01, c_v=0.1)
    f64 = lambda a: torch.tensor(np.asarray(a, np.float64))  # noqa: E731
    L = O.lib()
    dt = 1e-4
    lio = (L.hbr_exp(-dt / c["tau_s"]), c["bias"] - c["bias"] * L.hbr_exp(-dt / c["tau_m"]),
           (((c["tau_s"] * ((-L.hbr_exp(dt / c["tau_m"])) + L.hbr_exp(dt / c["tau_s"]))) * L.hbr_exp(-dt / c["tau_m"]))
            * L.hbr_exp(-dt / c["tau_s"])) / (c["tau_m"] - c["tau_s"]), L.hbr_exp(-dt / c["tau_m"]))
    with torch.no_grad():
        loss, fs, raster = reference_loss(inp, inp["trains"], inp["v0"], f64(inp["w_in"]), f64(inp["w_rec"]),
                                          f64(c["tau_m"]), f64(c["tau_s"]), c, lio_values=lio)
    for b in range(meta["trials"]):
        kk, ii = np.nonzero(raster[:, b, :])
        o = np.lexsort((ii, kk))
        sel = z["trial"] == b
        assert np.array_equal(ii[o], z["i"][sel]) and np.array_equal(kk[o], z["step"][sel])
    assert np.isfinite(float(loss))
