"""Float % on the device equals the cython target's (host restatement codegen.py_mod,
itself checked against CPython's float modulo): the generated ((a % b) + b) % b of
cython_generator.py:86-89, each % as Cython compiles it without cdivision."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_mod_equals_host_mod():
    import torch

    from brian2_amd import _lib, codegen
    from brian2_amd.runtime import Device

    from test_codegen import MOD_CASES

    rng = np.random.default_rng(4)
    pairs = np.array(MOD_CASES + [tuple(x) for x in rng.standard_normal((4000, 2)) * [[50.0, 3.0]]], np.float64)
    n = len(pairs)
    c, _ = codegen.Renderer({"a": "double", "b": "double"}).render("a % b")
    src = codegen.source_header(256) + f"""
struct hbt_mod {{ const double* a; const double* b; double* out; long long n; }};
extern "C" __global__ void __launch_bounds__(256) mod_test(const hbt_mod p) {{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double a = p.a[i], b = p.b[i];
    p.out[i] = {c};
}}
"""
    D = Device.get()
    fn = D.function(src, "mod_test")
    a = torch.from_numpy(pairs[:, 0].copy()).cuda()
    b = torch.from_numpy(pairs[:, 1].copy()).cuda()
    out = torch.empty(n, dtype=torch.float64, device="cuda")

    class P(ctypes.Structure):
        _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("out", ctypes.c_void_p), ("n", ctypes.c_int64)]

    args = bytes(P(a.data_ptr(), b.data_ptr(), out.data_ptr(), n))
    _lib.check(_lib.lib.hb_launch(fn, (n + 255) // 256, 256, 0, D.stream, args, len(args)), "launch")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = np.array([codegen.HostEval({"a": x, "b": y}).eval("a % b") for x, y in pairs])
    assert np.array_equal(got.view(np.int64), want.view(np.int64))  # bit for bit, signs of zero included
