"""The C-ABI library loads and exports every symbol include/hipbrian.h declares (no GPU needed)."""
import ctypes
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(REPO, "include", "hipbrian.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^(?:int|double)\s+(hb_\w+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(REPO, "brian2_amd", "libhipbrian.so"))
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_the_header():
    from brian2_amd import _lib

    assert sorted(_lib.exported_symbols()) == declared_symbols()
    assert _lib.lib.hb_version() == 1


def test_errors_are_reported_without_gpu():
    from brian2_amd import _lib

    # a launch record with a step offset outside its argument buffer is rejected on the host
    rc = _lib.lib.hb_plan_add(None, None, 1, 64, 0, b"\0" * 8, 8, 4)
    assert rc != 0 and "step offset" in _lib.last_error()


def test_host_generator_matches_the_oracles():
    """brian2_amd.synth (numpy SplitMix64) == the oracle's C generator."""
    import numpy as np

    import oracle_py as O
    from brian2_amd import synth

    ctr = np.arange(1000, dtype=np.uint64) * np.uint64(7919)
    L = O.lib()
    assert all(int(synth.rand64(1234, 3, c)) == L.hbr_rand64(1234, 3, int(c)) for c in ctr[:200])
    u = synth.uniform(1234, 2, ctr)
    assert np.array_equal(u, [L.hbr_uniform(1234, 2, int(c)) for c in ctr])
    pre, post = synth.bernoulli_host(1234, 10, 60, 300, 0.05)
    opre, opost = O.bernoulli(1234, 10, 60, 300, 0.05)
    assert np.array_equal(pre, opre) and np.array_equal(post, opost)
    z = synth.polar_normal(1234, 50)
    assert np.array_equal(z, O.polar_normal(1234, 50))


def test_one_hip_runtime_in_the_process():
    """libhipbrian.so binds to the HIP runtime torch loaded, not a second copy."""
    import brian2_amd  # noqa: F401

    with open("/proc/self/maps") as f:
        libs = {ln.split()[-1] for ln in f if "libamdhip64" in ln}
    assert len(libs) == 1, libs


def test_code_object_spill_counts():
    """hb_code_object_spills reads the spill counts a gfx950 code object records: none for
    a small kernel, some for one held to 32 VGPRs with many live values."""
    import subprocess
    import tempfile

    from brian2_amd import _lib

    small = r"""
#include <hip/hip_runtime.h>
extern "C" __global__ void __launch_bounds__(64) small(float* x) { x[threadIdx.x] += 1.0f; }
"""
    big = r"""
#include <hip/hip_runtime.h>
extern "C" __global__ void __launch_bounds__(1024) __attribute__((amdgpu_num_vgpr(32))) big(float* x) {
    float a[96];
    #pragma unroll
    for (int i = 0; i < 96; i++) a[i] = x[threadIdx.x + 1024 * i];
    float s = 0.f;
    #pragma unroll
    for (int i = 0; i < 96; i++) s += a[i] * a[95 - i] + a[(i * 7) % 96];
    #pragma unroll
    for (int i = 0; i < 96; i++) x[threadIdx.x + 1024 * i] = a[i] * s;
}
"""
    counts = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as d:
        for name, src in (("small", small), ("big", big)):
            p = os.path.join(d, f"{name}.hip")
            with open(p, "w") as f:
                f.write(src)
            co = os.path.join(d, f"{name}.co")
            r = subprocess.run(["hipcc", "--offload-arch=gfx950", "-O3", "--genco", p, "-o", co],
                               capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-2000:]
            with open(co, "rb") as f:
                data = f.read()
            v, s = ctypes.c_int(), ctypes.c_int()
            assert _lib.lib.hb_code_object_spills(data, len(data), ctypes.byref(v), ctypes.byref(s)) == 0
            counts[name] = (v.value, s.value)
    assert counts["small"] == (0, 0)
    assert counts["big"][0] > 0
