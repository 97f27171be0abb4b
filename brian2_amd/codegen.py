"""HIP code generation from Brian's optimised statements.

Brian turns equations into abstract code and then into optimised statement
lists (brian2/codegen/translation.py:181 make_statements,
brian2/codegen/optimisation.py:63-167 optimise_statements): a scalar block of
loop invariants (``_lio_*``) and a vector block, with boolean-split variants of
expressions that use boolean variables.  The C++ target renders them
statement by statement (brian2/codegen/generators/cpp_generator.py:251-424);
this module does the same for gfx950, with three differences:

* the scalar block is evaluated once on the host (in Python, IEEE double, with
  the same elementary functions as the device) and its values become kernel
  arguments;
* the stateupdate, threshold and reset code objects of a group, the fold of the
  synaptic events delivered to it and its monitors are fused into one kernel
  per group and step (one HBM pass over the state);
* expressions keep Brian's association exactly (fully parenthesised), and the
  kernels are compiled with -ffp-contract=off, so fp64 results are bit-identical
  to the cython runtime's.
"""
import ast
import math

import numpy as np

from . import _lib

L = _lib.lib

# ---------------------------------------------------------------------------
# C types
# ---------------------------------------------------------------------------
_RANK = {"bool": 0, "int": 1, "long long": 2, "float": 3, "double": 4}
CLOCK_DOUBLE = {"t", "dt", "lastspike", "_clock_t", "_clock_dt", "_source_t"}


def promote(a, b):
    return a if _RANK[a] >= _RANK[b] else b


def is_int(t):
    return t in ("bool", "int", "long long")


def py_mod(a, b):
    """Float % as Cython compiles it without cdivision (the cython target's scalar
    blocks): fmod, then b added once when the remainder is non-zero and its sign
    differs from b's (the same IEEE operations as the device's hb_pymod)."""
    r = math.fmod(a, b)
    return r + float((r != 0) & ((r < 0) ^ (b < 0))) * b


def ctype_of(dtype, float_t="double", name=None):
    """C type of a Brian dtype name; float64 state follows the precision mode."""
    if dtype in ("float64", "double"):
        return "double" if (float_t == "double" or name in CLOCK_DOUBLE) else float_t
    if dtype in ("float32", "float"):
        return "float"
    if dtype in ("bool", "bool_"):
        return "bool"
    if dtype in ("int32", "int"):
        return "int"
    if dtype in ("int64", "long long"):
        return "long long"
    if dtype == "uint64":
        return "unsigned long long"
    raise NotImplementedError(f"dtype {dtype!r} is not supported by the hip target")


def c_literal(v):
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        v = int(v)
        return f"{v}LL" if not (-(2**31) <= v < 2**31) else str(v)
    v = float(v)
    if math.isinf(v):
        return "(__builtin_inf())" if v > 0 else "(-__builtin_inf())"
    if math.isnan(v):
        return "(__builtin_nan(\"\"))"
    r = repr(v)
    if "e" not in r and "." not in r:
        r += ".0"
    return r


# ---------------------------------------------------------------------------
# expression rendering (Brian syntax -> C), cf. CPPNodeRenderer
# (brian2/parsing/rendering.py:314-366)
# ---------------------------------------------------------------------------
_CMP = {ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">=", ast.Eq: "==", ast.NotEq: "!="}
_BIN = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*"}


class Renderer:
    """Renders a Brian expression string into a C expression and its type.

    ``types`` maps every identifier to its C type.
    """

    def __init__(self, types):
        self.types = types

    def render(self, expr):
        node = ast.parse(expr.strip(), mode="eval").body
        return self._r(node)

    def _r(self, n):
        m = getattr(self, "_" + type(n).__name__, None)
        if m is None:
            raise NotImplementedError(f"expression element {type(n).__name__} is not supported")
        return m(n)

    def _Constant(self, n):
        v = n.value
        if isinstance(v, bool):
            return ("true" if v else "false"), "bool"
        if isinstance(v, int):
            return c_literal(v), ("int" if -(2**31) <= v < 2**31 else "long long")
        if isinstance(v, float):
            return c_literal(v), "double"
        raise NotImplementedError(f"constant {v!r}")

    def _Name(self, n):
        if n.id in ("True", "False"):
            return ("true" if n.id == "True" else "false"), "bool"
        if n.id == "inf":
            return "(__builtin_inf())", "double"
        if n.id not in self.types:
            raise NotImplementedError(f"unknown identifier {n.id!r} in generated code")
        return n.id, self.types[n.id]

    def _UnaryOp(self, n):
        c, t = self._r(n.operand)
        if isinstance(n.op, ast.USub):
            return f"(-{c})", ("int" if t == "bool" else t)
        if isinstance(n.op, ast.UAdd):
            return f"(+{c})", ("int" if t == "bool" else t)
        if isinstance(n.op, ast.Not):
            return f"(!{c})", "bool"
        raise NotImplementedError(type(n.op).__name__)

    def _BoolOp(self, n):
        op = " && " if isinstance(n.op, ast.And) else " || "
        parts = [self._r(v)[0] for v in n.values]
        return "(" + op.join(parts) + ")", "bool"

    def _Compare(self, n):
        if len(n.ops) != 1:
            raise NotImplementedError("chained comparisons")
        a, _ = self._r(n.left)
        b, _ = self._r(n.comparators[0])
        return f"({a} {_CMP[type(n.ops[0])]} {b})", "bool"

    def _BinOp(self, n):
        a, ta = self._r(n.left)
        b, tb = self._r(n.right)
        op = type(n.op)
        if op in _BIN:
            t = promote(ta, tb)
            if t == "bool":
                t = "int"
            return f"({a} {_BIN[op]} {b})", t
        if op is ast.Div:
            if is_int(ta) and is_int(tb):
                return f"((double)({a}) / (double)({b}))", "double"
            return f"({a} / {b})", promote(ta, tb)
        if op is ast.Pow:
            t = promote(ta, tb)
            if t == "float":
                return f"powf({a}, {b})", "float"
            return f"hb_pow((double)({a}), (double)({b}))", "double"
        if op is ast.FloorDiv:
            if is_int(ta) and is_int(tb):
                t = "long long" if "long long" in (ta, tb) else "int"
                return f"hb_floordiv<{t}>({a}, {b})", t
            t = promote(ta, tb)
            return f"floor((double)({a}) / ({b}))", "double" if t != "float" else "float"
        if op is ast.Mod:
            if is_int(ta) and is_int(tb):
                t = "long long" if "long long" in (ta, tb) else "int"
                return f"hb_imod<{t}>({a}, {b})", t
            # the cython target renders a % b as ((a % b) + b) % b (cython_generator.py:86-89),
            # each % with Python's sign convention (cdivision=False)
            t = promote(ta, tb)
            if t == "float":
                return f"hb_pymodf(hb_pymodf({a}, {b}) + ({b}), {b})", "float"
            return f"hb_pymod(hb_pymod((double)({a}), (double)({b})) + (double)({b}), (double)({b}))", "double"
        raise NotImplementedError(f"operator {op.__name__}")

    _F64 = {"exp": "hb_exp", "expm1": "hb_expm1", "exprel": "hb_exprel", "log": "hb_log", "sqrt": "sqrt",
            "floor": "floor", "ceil": "ceil", "sin": "sin", "cos": "cos", "tan": "tan", "sinh": "sinh",
            "cosh": "cosh", "tanh": "tanh", "arcsin": "asin", "arccos": "acos", "arctan": "atan",
            "log10": "log10", "log1p": "log1p"}
    _F32 = {"exp": "expf", "expm1": "expm1f", "exprel": "hb_exprelf", "log": "logf", "sqrt": "sqrtf",
            "floor": "floorf", "ceil": "ceilf", "sin": "sinf", "cos": "cosf", "tan": "tanf", "sinh": "sinhf",
            "cosh": "coshf", "tanh": "tanhf", "arcsin": "asinf", "arccos": "acosf", "arctan": "atanf",
            "log10": "log10f", "log1p": "log1pf"}

    def _Call(self, n):
        if not isinstance(n.func, ast.Name):
            raise NotImplementedError("call of a non-name")
        f = n.func.id
        args = [self._r(a) for a in n.args]
        if f in self._F64:
            (c, t), = args
            if t == "float":
                return f"{self._F32[f]}({c})", "float"
            return f"{self._F64[f]}((double)({c}))", "double"
        if f == "int":
            (c, t), = args
            return f"((int)({c}))", "int"
        if f == "float":
            (c, t), = args
            return f"((double)({c}))", "double"
        if f == "abs":
            (c, t), = args
            if is_int(t):
                return f"(({c}) < 0 ? -({c}) : ({c}))", ("int" if t == "bool" else t)
            return (f"fabsf({c})", "float") if t == "float" else (f"fabs((double)({c}))", "double")
        if f == "timestep":
            (a, _), (b, _) = args
            return f"hb_timestep((double)({a}), (double)({b}))", "long long"
        if f == "clip":
            (a, ta), (b, _), (c, _) = args
            t = "double" if ta != "float" else "float"
            return f"hb_clip<{t}>(({t})({a}), (double)({b}), (double)({c}))", t
        if f == "sign":
            (a, ta), = args
            return f"(((0 < ({a})) - (({a}) < 0)))", "int"
        raise NotImplementedError(f"function {f!r} is not supported by the hip target")


# ---------------------------------------------------------------------------
# host evaluation of the scalar (loop-invariant) block
# ---------------------------------------------------------------------------
def _host_timestep(t, dt):
    return int((t + 1e-3 * dt) / dt)


_HOST_FUNCS = {
    "exp": lambda x: L.hb_host_exp(float(x)),
    "expm1": lambda x: L.hb_host_expm1(float(x)),
    "exprel": lambda x: L.hb_host_exprel(float(x)),
    "log": lambda x: L.hb_host_log(float(x)),
    "sqrt": lambda x: math.sqrt(float(x)),
    "abs": abs,
    "int": lambda x: int(x),
    "float": float,
    "timestep": _host_timestep,
    "floor": lambda x: float(math.floor(x)),
    "ceil": lambda x: float(math.ceil(x)),
}


class HostEval:
    """Evaluates scalar statements in IEEE double with C semantics."""

    def __init__(self, values):
        self.values = dict(values)

    def eval(self, expr):
        return self._e(ast.parse(expr.strip(), mode="eval").body)

    def _e(self, n):
        if isinstance(n, ast.Constant):
            return n.value
        if isinstance(n, ast.Name):
            if n.id in ("True", "False"):
                return n.id == "True"
            if n.id == "inf":
                return float("inf")
            if n.id not in self.values:
                raise NotImplementedError(f"scalar code uses {n.id!r}, which has no host value")
            return self.values[n.id]
        if isinstance(n, ast.UnaryOp):
            v = self._e(n.operand)
            if isinstance(n.op, ast.USub):
                return -v
            if isinstance(n.op, ast.UAdd):
                return +v
            if isinstance(n.op, ast.Not):
                return not v
        if isinstance(n, ast.BoolOp):
            vals = [bool(self._e(v)) for v in n.values]
            return all(vals) if isinstance(n.op, ast.And) else any(vals)
        if isinstance(n, ast.Compare):
            a, b = self._e(n.left), self._e(n.comparators[0])
            return {ast.Lt: a < b, ast.LtE: a <= b, ast.Gt: a > b, ast.GtE: a >= b, ast.Eq: a == b,
                    ast.NotEq: a != b}[type(n.ops[0])]
        if isinstance(n, ast.BinOp):
            a, b = self._e(n.left), self._e(n.right)
            op = type(n.op)
            if op is ast.Add:
                return a + b
            if op is ast.Sub:
                return a - b
            if op is ast.Mult:
                return a * b
            if op is ast.Div:
                return float(a) / float(b)
            if op is ast.Pow:
                return L.hb_host_pow(float(a), float(b))
            if op is ast.FloorDiv:
                return a // b if isinstance(a, int) and isinstance(b, int) else float(math.floor(a / b))
            if op is ast.Mod:
                if isinstance(a, int) and isinstance(b, int):
                    return ((a % b) + b) % b
                return py_mod(py_mod(float(a), float(b)) + float(b), float(b))
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _HOST_FUNCS:
            return _HOST_FUNCS[n.func.id](*[self._e(a) for a in n.args])
        raise NotImplementedError(f"scalar expression element {ast.dump(n)}")


def cast_value(v, ctype):
    if ctype == "double":
        return float(v)
    if ctype == "float":
        return float(np.float32(v))
    if ctype == "bool":
        return bool(v)
    return int(v)


# ---------------------------------------------------------------------------
# statements -> C (CPPCodeGenerator.translate_statement, cpp_generator.py:251-312)
# ---------------------------------------------------------------------------
def translate_statement(st, types, declared, float_t):
    """One statement dict (captured from Brian) -> list of C lines."""
    var, op, dtype = st["var"], st["op"], st["dtype"]
    R = Renderer(types)
    lines = []
    ct = types.get(var) or ctype_of(dtype, float_t, var)
    if op == ":=":
        types[var] = ct
        declared.add(var)
        op = "="
        if st.get("bool_exprs"):
            lines.append(f"{ct} {var};")
        else:
            c, _ = R.render(st["expr"])
            lines.append(f"const {ct} {var} = ({ct})({c});" if st.get("constant") else f"{ct} {var} = ({ct})({c});")
            return lines
    if st.get("bool_exprs"):
        first = True
        nb = len(st["bool_vars"])
        for assigns, e in st["bool_exprs"]:
            conds = " && ".join(v if val else f"!{v}" for v, val in assigns)
            c, _ = R.render(e)
            head = "" if first else "else "
            if first or nb > 1:
                head += f"if ({conds})"
            lines.append(f"{head} {{ {var} {op} ({ct})({c}); }}")
            first = False
        return lines
    c, _ = R.render(st["expr"])
    lines.append(f"{var} {op} ({ct})({c});" if op == "=" else f"{var} {op} {c};")
    return lines


def translate_block(statements, conditional_write, types, declared, float_t):
    out = []
    for st in statements:
        ls = translate_statement(st, types, declared, float_t)
        cw = conditional_write.get(st["var"])
        if cw is not None:
            out.append(f"if ({cw}) {{")
            out.extend("    " + x for x in ls)
            out.append("}")
        else:
            out.extend(ls)
    return out


# ---------------------------------------------------------------------------
# kernel parameter structs (one struct by value per kernel)
# ---------------------------------------------------------------------------
import ctypes  # noqa: E402

_CT = {
    "double": ctypes.c_double, "float": ctypes.c_float, "int": ctypes.c_int32, "long long": ctypes.c_int64,
    "bool": ctypes.c_bool, "unsigned long long": ctypes.c_uint64, "unsigned int": ctypes.c_uint32,
    "ptr": ctypes.c_void_p,
}


class ParamStruct:
    """A kernel's argument struct, rendered to C and to ctypes with the same layout."""

    def __init__(self, name):
        self.name = name
        self.fields = []  # (cname, ctype-decl, kind)

    def ptr(self, name, ctype, const=False):
        self.fields.append((name, f"{'const ' if const else ''}{ctype}*", "ptr"))

    def val(self, name, ctype):
        self.fields.append((name, ctype, ctype))

    def c_decl(self):
        body = "\n".join(f"    {t} {n};" for n, t, _ in self.fields)
        return f"struct {self.name} {{\n{body}\n}};\n"

    def ctypes_struct(self):
        fields = [(n, _CT[k]) for n, _, k in self.fields]
        return type(self.name, (ctypes.Structure,), {"_fields_": fields})


PRELUDE = r"""
__device__ __forceinline__ long long hb_timestep(double t, double dt) { return (long long)((t + 1e-3 * dt) / dt); }
template <typename T> __device__ __forceinline__ T hb_clip(T v, double lo, double hi) {
    if (v < lo) return (T)lo;
    if (v > hi) return (T)hi;
    return v;
}
template <typename T> __device__ __forceinline__ T hb_floordiv(T a, T b) {
    T q = a / b; T r = a - q * b; q -= ((r != 0) & ((r ^ b) < 0)); return q;
}
template <typename T> __device__ __forceinline__ T hb_imod(T x, T y) {
    T r = x % y; r += ((r != 0) & ((r ^ y) < 0)) * y; return r;
}
// float % with Python's sign convention, as Cython compiles it without cdivision: fmod, then
// the divisor added once when the remainder is non-zero and its sign differs from the divisor's
__device__ __forceinline__ double hb_pymod(double a, double b) {
    double r = fmod(a, b); r += (double)((r != 0) & ((r < 0) ^ (b < 0))) * b; return r;
}
__device__ __forceinline__ float hb_pymodf(float a, float b) {
    float r = fmodf(a, b); r += (float)((r != 0) & ((r < 0) ^ (b < 0))) * b; return r;
}
__device__ __forceinline__ float hb_exprelf(float x) {
    if (fabsf(x) < 1e-16f) return 1.0f;
    if (x > 717.0f) return __builtin_inff();
    return expm1f(x) / x;
}
// Staged event records: the 64-bit order key and the pathway value's bits, one record per
// event (key and value of an event share cache lines when written and read).
struct hb_rec4 { unsigned int k0, k1, v; };                                   // 4-byte values, 12 B
struct __attribute__((aligned(16))) hb_rec8 { unsigned long long k, v; };     // 8-byte values, 16 B
__device__ __forceinline__ unsigned long long hb_rkey(const hb_rec4& r) {
    return ((unsigned long long)r.k1 << 32) | r.k0;
}
__device__ __forceinline__ unsigned long long hb_rkey(const hb_rec8& r) { return r.k; }
template <typename V> __device__ __forceinline__ V hb_rval(const hb_rec4& r) { return __builtin_bit_cast(V, r.v); }
template <typename V> __device__ __forceinline__ V hb_rval(const hb_rec8& r) { return __builtin_bit_cast(V, r.v); }
template <typename V> __device__ __forceinline__ void hb_rset(hb_rec4& r, unsigned long long k, V v) {
    r.k0 = (unsigned int)k; r.k1 = (unsigned int)(k >> 32); r.v = __builtin_bit_cast(unsigned int, v);
}
template <typename V> __device__ __forceinline__ void hb_rset(hb_rec8& r, unsigned long long k, V v) {
    r.k = k; r.v = __builtin_bit_cast(unsigned long long, v);
}
// Sorts a run of record indices by the records' keys (one target's events, in global
// memory): insertion sort for short runs, heapsort (O(m log m), in place) for long ones.
template <typename R>
__device__ void hb_sort_idx(unsigned int* ix, const R* recs, int n) {
    if (n <= 24) {
        for (int a = 1; a < n; a++) {
            const unsigned int xx = ix[a]; const unsigned long long kk = hb_rkey(recs[xx]); int b = a;
            while (b > 0 && hb_rkey(recs[ix[b - 1]]) > kk) { ix[b] = ix[b - 1]; b--; }
            ix[b] = xx;
        }
        return;
    }
    for (int start = n / 2 - 1; start >= 0; start--) {
        int r = start;
        while (2 * r + 1 < n) {
            int c = 2 * r + 1;
            if (c + 1 < n && hb_rkey(recs[ix[c]]) < hb_rkey(recs[ix[c + 1]])) c++;
            if (hb_rkey(recs[ix[r]]) < hb_rkey(recs[ix[c]])) { unsigned int t = ix[r]; ix[r] = ix[c]; ix[c] = t; r = c; }
            else break;
        }
    }
    for (int end = n - 1; end > 0; end--) {
        { unsigned int t = ix[0]; ix[0] = ix[end]; ix[end] = t; }
        int r = 0;
        while (2 * r + 1 < end) {
            int c = 2 * r + 1;
            if (c + 1 < end && hb_rkey(recs[ix[c]]) < hb_rkey(recs[ix[c + 1]])) c++;
            if (hb_rkey(recs[ix[r]]) < hb_rkey(recs[ix[c]])) { unsigned int t = ix[r]; ix[r] = ix[c]; ix[c] = t; r = c; }
            else break;
        }
    }
}
// inclusive scan over the 64 lanes of a wave: shifts within rows of 16 lanes, then the
// last lane of row 0 added to row 1 and of row 2 to row 3, then that of row 1 to rows 2-3
// (DPP row_shr and row_bcast; lanes whose source is out of the row contribute 0)
__device__ __forceinline__ int hb_wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return x;
}
// offset of this thread's v items in a workgroup-wide allocation counted in *ctr (LDS):
// a prefix within the wave, and one LDS atomic per wave for the wave's base (waves in
// arbitrary order).  *ctr must be zero before; after a barrier it holds the total.
__device__ __forceinline__ int hb_block_alloc(int v, int* ctr) {
    const int x = hb_wave_incl_scan(v);
    const int tot = __builtin_amdgcn_readlane(x, 63);
    int base = 0;
    if ((threadIdx.x & 63) == 63 && tot != 0) base = atomicAdd(ctr, tot);
    return __builtin_amdgcn_readlane(base, 63) + x - v;
}
// exclusive scan of one int per thread over a workgroup of HB_T threads (HB_T/64 waves):
// scan within each wave, wave totals through LDS; lane l of every wave reads wave l's
// total and a second wave scan gives the totals before this wave (lane wid - 1) and the
// workgroup's total (lane 63), read out as wave-uniform values (sh: HB_T/64 ints)
__device__ __forceinline__ int hb_block_exscan(int v, int* sh, int* total) {
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int x = hb_wave_incl_scan(v);
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    const int w = hb_wave_incl_scan(lane < HB_T / 64 ? sh[lane] : 0);
    const int pre = wid > 0 ? __builtin_amdgcn_readlane(w, wid - 1) : 0;
    *total = __builtin_amdgcn_readlane(w, 63);
    __syncthreads();
    return x - v + pre;
}
"""


def source_header(T):
    import os

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "csrc", "hb_math.h")) as f:
        math_h = f.read()
    return f"#define HB_T {T}\n" + math_h + PRELUDE
