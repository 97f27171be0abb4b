"""hb_scan_u32_noalloc (the scan that allocates nothing: its block sums live in a
fixed device array of the library) against hb_scan_u32, through the C ABI."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan(fn_name, counts):
    import torch

    from brian2_amd import _lib

    d = torch.from_numpy(np.concatenate([counts, [0xDEADBEEF]]).astype(np.uint32).view(np.int32)).cuda()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fn = getattr(_lib.lib, fn_name)
    _lib.check(fn(ctypes.c_void_p(d.data_ptr()), len(counts), stream), fn_name)
    torch.cuda.synchronize()
    return d.cpu().numpy().view(np.uint32)


def test_noalloc_scan_equals_the_allocating_scan_on_two_calls():
    rng = np.random.default_rng(11)
    # a large call, then a smaller one: the second must not see the first's block sums
    for n in (3_000_001, 100_003):
        print(f"scan of {n} counts", flush=True)
        counts = rng.integers(0, 40, n, dtype=np.uint32)
        want = _scan("hb_scan_u32", counts)
        got = _scan("hb_scan_u32_noalloc", counts)
        assert np.array_equal(got, want)
        ref = np.zeros(n + 1, np.uint64)
        np.cumsum(counts, out=ref[1:])
        assert np.array_equal(got, ref.astype(np.uint32))


def test_noalloc_scan_at_its_maximum_and_empty():
    n = 65536 * 1024
    print(f"scan of {n} counts (the maximum)", flush=True)
    counts = np.ones(n, np.uint32)
    got = _scan("hb_scan_u32_noalloc", counts)
    assert got[0] == 0 and got[n] == n and np.array_equal(got[:1000], np.arange(1000))
    assert np.array_equal(got[-3:], [n - 2, n - 1, n])
    assert _scan("hb_scan_u32_noalloc", np.zeros(0, np.uint32))[0] == 0


def test_noalloc_scan_rejects_sizes_beyond_its_maximum():
    from brian2_amd import _lib

    # rejected on the host before anything is launched (the pointer is never touched)
    rc = _lib.lib.hb_scan_u32_noalloc(None, 65536 * 1024 + 1, None)
    assert rc != 0 and "exceed the maximum" in _lib.last_error()
