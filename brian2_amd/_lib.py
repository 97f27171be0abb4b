"""ctypes binding of libhipbrian.so (C ABI in include/hipbrian.h).

The product path has no fallback: if the library is missing or fails to load,
importing this module raises.
"""
import ctypes
import os

try:
    # torch ships its own libamdhip64 (SONAME libamdhip64.so.7) and loads it by file
    # name; loading torch first makes this library bind to that same HIP runtime
    # instead of a second copy from /opt/rocm (two runtimes in one process do not
    # share devices, streams or allocations).
    import torch  # noqa: F401
except ImportError:
    pass

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhipbrian.so")


class HipBrianError(RuntimeError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python brian2_amd/build.py`"
    )

lib = ctypes.CDLL(LIB_PATH)

_i, _i32, _u32, _i64, _u64, _sz, _dbl = (ctypes.c_int, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64,
                                         ctypes.c_uint64, ctypes.c_size_t, ctypes.c_double)
_vp, _cp = ctypes.c_void_p, ctypes.c_char_p
_P = ctypes.POINTER

_SIGS = {
    "hb_version": ([], _i),
    "hb_last_error": ([ctypes.c_char_p, _sz], _i),
    "hb_ctx_create": ([_i, _P(_vp)], _i),
    "hb_ctx_destroy": ([_vp], _i),
    "hb_device_count": ([_P(_i)], _i),
    "hb_malloc": ([_vp, _sz, _P(_vp)], _i),
    "hb_free": ([_vp, _vp], _i),
    "hb_memcpy": ([_vp, _vp, _sz, _i, _vp], _i),
    "hb_fill_u32": ([_vp, _u32, _sz, _vp], _i),
    "hb_stream_sync": ([_vp], _i),
    "hb_module_load": ([_vp, _cp, _cp, _P(_vp), _cp, _sz], _i),
    "hb_module_unload": ([_vp], _i),
    "hb_module_function": ([_vp, _cp, _P(_vp)], _i),
    "hb_launch": ([_vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _vp, _vp, _sz], _i),
    "hb_function_info": ([_vp, _P(_i), _P(_i64), _P(_i64)], _i),
    "hb_function_spills": ([_vp, _P(_i), _P(_i)], _i),
    "hb_code_object_spills": ([_vp, _sz, _P(_i), _P(_i)], _i),
    "hb_plan_create": ([_vp, _vp, _P(_vp)], _i),
    "hb_plan_add": ([_vp, _vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _vp, _sz, ctypes.c_long], _i),
    "hb_plan_set_args": ([_vp, _i, _vp, _sz], _i),
    "hb_plan_set_lane": ([_vp, _i, _i], _i),
    "hb_plan_run": ([_vp, _vp, _i64, _i], _i),
    "hb_plan_timing": ([_vp, _i, _P(_dbl), _P(_i64)], _i),
    "hb_plan_destroy": ([_vp], _i),
    "hb_plan_profile": ([_vp, _vp, _i64, _P(ctypes.c_float)], _i),
    "hb_spikegen_function": ([_P(_vp)], _i),
    "hb_poisson_function": ([_P(_vp)], _i),
    "hb_csr_count": ([_vp, _vp, _i64, _i32, _i32, _i32, _i32, _vp, _vp], _i),
    "hb_scan_u32": ([_vp, _i64, _vp], _i),
    "hb_scan_u32_noalloc": ([_vp, _i64, _vp], _i),
    "hb_gen_fixed_indegree": ([_u64, _u64, _i32, _i32, _i32, _vp, _vp, _vp], _i),
    "hb_csr_from_indegree": ([_vp, _i32, _i32, _i32, _i32, _u64, _u64, _i32, _i32, _u64, _dbl, _dbl, _dbl, _vp, _vp,
                              _vp, _vp, _vp], _i),
    "hb_gen_bernoulli_rows": ([_u64, _u64, _i32, _i32, _i32, _dbl, _vp, _vp, _vp], _i),
    "hb_host_exp": ([_dbl], _dbl),
    "hb_host_expm1": ([_dbl], _dbl),
    "hb_host_log": ([_dbl], _dbl),
    "hb_host_pow": ([_dbl, _dbl], _dbl),
    "hb_host_exprel": ([_dbl], _dbl),
}

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res


def last_error():
    buf = ctypes.create_string_buffer(8192)
    lib.hb_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise HipBrianError(f"{what}: {last_error()}" if what else last_error())
    return rc


def exported_symbols():
    return sorted(_SIGS)
