"""Code generation (CPU only): expression rendering, host evaluation of the
loop-invariant block, and compilation of the generated kernels for gfx950."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from brian2_amd import codegen as cg
from brian2_amd.kernels import CarryKernel, GatherKernel, NeuronKernel, scalar_values
from brian2_amd.runtime import Model, NeuronGroup


def R(expr, **types):
    return cg.Renderer(types).render(expr)


def test_render_association_and_types():
    c, t = R("_lio_12 + (((_lio_13 * ge) + (_lio_14 * gi)) + (_lio_15 * v))", _lio_12="double", _lio_13="double",
             _lio_14="double", _lio_15="double", ge="double", gi="double", v="double")
    assert c == "(_lio_12 + (((_lio_13 * ge) + (_lio_14 * gi)) + (_lio_15 * v)))" and t == "double"
    # left association of a chain is kept (a - b) - c
    assert R("a - b - c", a="double", b="double", c="double")[0] == "((a - b) - c)"


def test_render_division_and_powers():
    assert R("a / b", a="int", b="int") == ("((double)(a) / (double)(b))", "double")
    assert R("a / b", a="double", b="int") == ("(a / b)", "double")
    assert R("n ** 4", n="double") == ("hb_pow((double)(n), (double)(4))", "double")
    assert R("x ** 2", x="float")[1] == "float"
    assert R("a // b", a="int", b="int")[0] == "hb_floordiv<int>(a, b)"
    assert R("a % b", a="int", b="int")[0] == "hb_imod<int>(a, b)"


def test_render_booleans_and_functions():
    assert R("(v > Vt) and not_refractory", v="double", Vt="double", not_refractory="bool") == \
        ("((v > Vt) && not_refractory)", "bool")
    assert R("not a or b", a="bool", b="bool")[0] == "((!a) || b)"
    assert R("timestep(t - lastspike, dt) >= _lio_1", t="double", lastspike="double", dt="double",
             _lio_1="long long")[0] == "(hb_timestep((double)((t - lastspike)), (double)(dt)) >= _lio_1)"
    assert R("exprel(x)", x="double")[0] == "hb_exprel((double)(x))"
    assert R("exp(x)", x="float")[0] == "expf(x)"
    assert R("int(b)", b="bool") == ("((int)(b))", "int")
    with pytest.raises(NotImplementedError):
        R("rand()")


def test_literals_round_trip():
    for v in (0.1, 1e-16, 2980.9579870417283, -0.05555555555555555, 1.0, 5e-324):
        assert float(cg.c_literal(v)) == v


def test_host_eval_matches_c_semantics():
    ev = cg.HostEval({"dt": 1e-4, "taum": 0.02})
    assert ev.eval("timestep(0.005, dt)") == 50
    assert ev.eval("(- dt) / taum") == -1e-4 / 0.02
    assert ev.eval("7 / 2") == 3.5
    assert ev.eval("2 ** 10") == 1024.0


def test_translate_statement_bool_split_and_conditional_write():
    st = {"var": "_v", "op": ":=", "expr": "x", "dtype": "float64", "bool_vars": ["nr"],
          "bool_exprs": [[[["nr", False]], "a + v"], [[["nr", True]], "b + v"]]}
    types = {"a": "double", "b": "double", "v": "double", "nr": "bool"}
    lines = cg.translate_statement(st, types, set(), "double")
    assert lines == ["double _v;", "if (!nr) { _v = (double)((a + v)); }", "else  { _v = (double)((b + v)); }"]
    out = cg.translate_block([{"var": "v", "op": "=", "expr": "_v", "dtype": "float64"}], {"v": "nr"},
                             dict(types, _v="double"), set(), "double")
    assert out[0] == "if (nr) {" and out[-1] == "}"


def _scalars(G, dt=1e-4):
    sc = {}
    for tag, spec in (("su", G.su), ("th", G.th), ("rs", G.rs)):
        if spec is None:
            continue
        for n, vc in scalar_values(spec, {"dt": dt, "N": G.N}, G.float_t).items():
            sc[(tag, n)] = vc
    return sc


def test_cuba_loop_invariants_equal_the_oracles():
    """The host-evaluated _lio values are the same doubles the oracle computes."""
    import math

    G = NeuronGroup(10, Model.load("cuba"))
    sc = _scalars(G)
    dt, taum, taue, taui, El = 1e-4, 0.02, 0.005, 0.01, -0.049
    from brian2_amd._lib import lib as L

    e = L.hb_host_exp
    assert sc[("su", "_lio_1")] == (50, "long long")
    assert sc[("su", "_lio_2")][0] == e((-dt) / taue)
    assert sc[("su", "_lio_12")][0] == El - (El * e((-dt) / taum))
    lio4, lio5 = taue * e((-dt) / taue), -e(dt / taue)
    assert sc[("su", "_lio_13")][0] == ((lio4 * (lio5 + e(dt / taum))) * e((-dt) / taum)) / ((0.0 - taum) + taue)
    assert sc[("th", "Vt")][0] == -0.05 and sc[("rs", "Vr")][0] == -0.06
    assert not math.isnan(sc[("su", "_lio_14")][0])


def _compile(source):
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as d:
        src = os.path.join(d, "k.hip")
        with open(src, "w") as f:
            f.write("#include <hip/hip_runtime.h>\n" + source)
        r = subprocess.run(["hipcc", "--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-c", src, "-o",
                            os.path.join(d, "k.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]


def _no_float_atomics(source):
    for m in re.finditer(r"atomic\w+\(([^,]+),", source):
        arg = m.group(1)
        assert "val" not in arg and "a_" not in arg, f"floating-point atomic on {arg}"


@pytest.mark.parametrize("model,float_t,tile", [("cuba", "double", 256), ("cobahh", "double", 256),
                                                ("cuba_synw", "float", 1024), ("cuba_synw", "double", 1024)])
def test_generated_kernels_compile_for_gfx950(model, float_t, tile):
    m = Model.load(model)
    G = NeuronGroup(1000, m, float_t=float_t)
    sc = _scalars(G)
    fold = [(0, "gi", "+="), (1, "ge", "+=")]
    K = NeuronKernel("g", 1000, G.vars, float_t, G.su, G.th, G.rs, "fused" if G.rs else "none", sc, fold, 1024,
                     "double", spikemon=(0, 1000), statemon=["v"], uses_refractory=G.uses_refractory, tile=tile)
    _no_float_atomics(K.source)
    _compile(K.source)
    spec = m.pathways["exc"]
    syn = {"w": "float"} if model == "cuba_synw" else {}
    n_tiles = -(-1000 // tile)
    gk = GatherKernel("p", 1, 50 if model == "cuba_synw" else 1, 1, spec.vector, syn,
                      scalar_values(spec, {"dt": 1e-4}, float_t), "double", float_t, "ge", n_tiles, tile)
    _no_float_atomics(gk.source)
    _compile(gk.source)
    # the kernel that delivers events carried over a change of dt, from the gather's records
    ck = CarryKernel(gk)
    _no_float_atomics(ck.source)
    _compile(ck.source)


def test_gather_kernel_limits():
    spec = Model.load("cuba").pathways["exc"]
    sc = scalar_values(spec, {"dt": 1e-4}, "double")
    with pytest.raises(NotImplementedError):
        GatherKernel("p", 16, 1, 0, spec.vector, {}, sc, "double", "double", "ge", 4, 256)
    with pytest.raises(NotImplementedError):
        GatherKernel("p", 0, 5000, 0, spec.vector, {}, sc, "double", "double", "ge", 4, 256)


def test_plasticity_and_summed_kernels_compile_for_gfx950():
    """The STDP pathway (on_pre with synaptic writes, on_post) and the gap-junction
    summed variable generate kernels that compile, without floating-point atomics,
    and the synapses' analysis finds the target update and the written variables."""
    from brian2_amd.kernels import PostPathKernel, SummedKernel
    from brian2_amd.runtime import SpikeGeneratorGroup, Synapses

    m = Model.load("stdp")
    G = NeuronGroup(20, m)
    S = Synapses(SpikeGeneratorGroup(10, [0], timebins=[0]), G, m.pathways["pre"], on_post=m.pathways["post"])
    assert S.target_var == "ge" and S.op == "+=" and S.spec.vector[S.target_index]["var"] == "ge"
    assert S.plastic == {"Apre", "Apost", "w", "lastupdate"}
    types = {n: "double" for n in S.syn_vars}
    gk = GatherKernel("stdp", 0, 1, 0, S.spec.vector, types, scalar_values(S.spec, {"dt": 1e-4}, "double"),
                      "double", "double", "ge", 1, 256, target_index=S.target_index, plastic=S.plastic)
    assert gk.rec_arrays == {} and gk.rec_words == 1  # plastic variables are not copied into records
    pk = PostPathKernel("stdp", S.post_spec.vector, types, S.plastic,
                        scalar_values(S.post_spec, {"dt": 1e-4}, "double"), "double", 0)
    g = Model.load("gap")
    H = NeuronGroup(30, g)
    T = Synapses(H, H, summed={"Igap": g.summed["Igap"]})
    pre_r, post_r = T.summed_reads["Igap"]
    assert T.syn_vars == ["w"] and pre_r == {"v_pre": "v"} and post_r == {"v_post": "v"}
    sk = SummedKernel("gap_Igap", g.summed["Igap"].vector, {"w": "double"}, pre_r, post_r,
                      {"v_pre": "double", "v_post": "double"}, "double", {}, "double")
    for src in (gk.source, pk.source, sk.source):
        _no_float_atomics(src)
        _compile(src)


def test_step_kernel_two_workgroups_per_cu_build():
    """The 1024-thread step kernel of the 1M-neuron model asks for 8 waves per SIMD and
    takes the LDS share of two workgroups per CU."""
    from brian2_amd.kernels import LDS_BYTES, fold_capacity, fold_lds_bytes

    m = Model.load("cuba_synw")
    G = NeuronGroup(1000, m, float_t="double")
    cap = fold_capacity(1024, "float", 1, wgs_per_cu=2)
    assert fold_lds_bytes(1024, cap, "float") <= LDS_BYTES // 2
    K = NeuronKernel("g", 1000, G.vars, "double", G.su, G.th, G.rs, "fused", _scalars(G),
                     [(0, "gi", "+="), (1, "ge", "+=")], cap, "float", spikemon=(0, 1000), tile=1024,
                     waves_per_eu=8)
    assert "amdgpu_waves_per_eu(8)" in K.source
    _compile(K.source)


def test_fused_step_and_gathers_compile_for_gfx950():
    """A group's step kernel and the gathers into it as one kernel: the bodies share
    one LDS allocation (references into it, none of their own), use the fused
    launch's workgroup ranges and step offset, and the two-per-CU build keeps the
    LDS within half a CU."""
    from brian2_amd.kernels import LDS_BYTES, FusedStepKernel, fold_capacity

    m = Model.load("cuba_synw")
    N = 5000
    G = NeuronGroup(N, m, float_t="double")
    cap = fold_capacity(1024, "float", 1, wgs_per_cu=2)
    K = NeuronKernel("g", N, G.vars, "double", G.su, G.th, G.rs, "fused", _scalars(G),
                     [(0, "ge", "+="), (1, "gi", "+=")], cap, "float", spikemon=(0, N), tile=1024, waves_per_eu=8)
    nt = -(-N // 1024)
    gs = [GatherKernel(f"p{r}", r, 50, 1, m.pathways[pw].vector, {"w": "float"},
                       scalar_values(m.pathways[pw], {"dt": 1e-4}, "double"), "float", "double", tv, nt, 1024,
                       events_per_thread=2)
          for r, (pw, tv) in enumerate((("exc", "ge"), ("inh", "gi")))]
    for tiles_first in (False, True):
        F = FusedStepKernel(K, gs, [7, 3], nt, tiles_first=tiles_first)
        assert F.grid == nt + 10 and F.lds_bytes <= LDS_BYTES // 2
        body = F.source.split("struct hbf_g", 1)[0]
        funcs = body[body.index("__device__ __forceinline__ void hbb_step_g"):]
        assert "__shared__" not in funcs and "blockIdx" not in funcs and "gridDim" not in funcs
        assert "p.step_off" not in funcs
        _no_float_atomics(F.source)
        _compile(F.source)
    with pytest.raises(ValueError):
        FusedStepKernel(K, [GatherKernel("q", 0, 1, 1, m.pathways["exc"].vector, {"w": "float"},
                                         scalar_values(m.pathways["exc"], {"dt": 1e-4}, "double"), "float",
                                         "double", "ge", nt, 1024)], [1], nt + 1)


def test_network_fuse_options():
    """Network(fuse=...) settings are checked on the host (no device needed)."""
    from brian2_amd.runtime import Network

    assert Network(fuse={}).fuse and not Network(fuse=False).fuse and Network().fuse
    assert Network(fuse={"events_per_thread": 3, "gather_blocks": [128, 128]}).fuse_cfg["gather_blocks"] == [128, 128]
    with pytest.raises(TypeError):
        Network(fuse={"blocks": 3})
    with pytest.raises(ValueError):
        Network(fuse={"gather_blocks": 128})
    with pytest.raises(ValueError):
        Network(fuse={"gather_blocks": [0, 4]})


def test_split_spike_ring_kernels_compile_for_gfx950():
    """A group whose pathways read parts of it (exc/inh) keeps one spike list per part of
    its ring: the step kernel compacts spikes per part, each gather enumerates only the
    parts of its source range, and an on_post pathway reads every part."""
    from brian2_amd.kernels import FusedStepKernel, PostPathKernel, fold_capacity

    m = Model.load("cuba_synw")
    N = 5000
    G = NeuronGroup(N, m, float_t="double")
    cap = fold_capacity(1024, "float", 1, wgs_per_cu=2)
    K = NeuronKernel("g", N, G.vars, "double", G.su, G.th, G.rs, "fused", _scalars(G),
                     [(0, "ge", "+="), (1, "gi", "+=")], cap, "float", spikemon=(0, N), tile=1024, waves_per_eu=8,
                     ring_parts=(0, 4000))
    assert "s_npart" in K.source and "ring_cnt[slot * 2 + r]" in K.source
    nt = -(-N // 1024)
    gs = [GatherKernel(f"p{r}", r, 50, 1, m.pathways[pw].vector, {"w": "float"},
                       scalar_values(m.pathways[pw], {"dt": 1e-4}, "double"), "float", "double", tv, nt, 1024,
                       events_per_thread=3, ring_parts=(2, r, ((0, 4000)[r],)))
          for r, (pw, tv) in enumerate((("exc", "ge"), ("inh", "gi")))]
    assert "* 2 + 1 + s % 1]" in gs[1].source
    both = GatherKernel("pall", 0, 5, 1, m.pathways["exc"].vector, {"w": "float"},
                        scalar_values(m.pathways["exc"], {"dt": 1e-4}, "double"), "float", "double", "ge", nt, 1024,
                        ring_parts=(2, 0, (0, 4000)))
    s = Model.load("stdp")
    H = NeuronGroup(20, s)
    from brian2_amd.runtime import SpikeGeneratorGroup, Synapses

    S = Synapses(SpikeGeneratorGroup(10, [0], timebins=[0]), H, s.pathways["pre"], on_post=s.pathways["post"])
    pk = PostPathKernel("stdp", S.post_spec.vector, {n: "double" for n in S.syn_vars}, S.plastic,
                        scalar_values(S.post_spec, {"dt": 1e-4}, "double"), "double", 0, ring_parts=(0, 7, 12))
    assert "hb_c3" in pk.source
    F = FusedStepKernel(K, gs, [7, 3], nt)
    for src in (K.source, gs[0].source, both.source, pk.source, F.source):
        _no_float_atomics(src)
        _compile(src)
    with pytest.raises(ValueError):
        NeuronKernel("g", N, G.vars, "double", G.su, G.th, G.rs, "fused", _scalars(G), [], cap, "float",
                     tile=1024, ring_parts=(10, 4000))


# float % and // as the cython target evaluates them (cython_generator.py:86-89 renders
# a % b as ((a % b) + b) % b; Cython compiles % of doubles without cdivision as fmod plus
# a sign correction, and // of doubles as floor(a / b))
MOD_CASES = [(-7.5, 2.0), (7.5, -2.0), (-7.5, -2.0), (7.5, 2.0), (-1e-17, 1.0), (1e-17, -1.0), (5.0, 2.5),
             (-5.0, 2.5), (0.1, 0.03), (-0.1, 0.03), (1e300, 3.0), (-2.5e-3, 1e-3), (3.0, 1e-300)]


def test_py_mod_is_cpythons_float_mod_where_zero_signs_agree():
    import math

    rng = np.random.default_rng(2)
    pairs = MOD_CASES + [(float(a), float(b)) for a, b in rng.standard_normal((500, 2)) * [[100.0, 3.0]]]
    for a, b in pairs:
        r = cg.py_mod(a, b)
        want = a % b  # CPython's float modulo (the same algorithm up to the sign of a zero result)
        assert r == want or (r == 0 and want == 0), (a, b, r, want)
        assert math.copysign(1.0, r) == math.copysign(1.0, b) or r == 0, (a, b)
    # the known discontinuity: -1e-17 % 1.0 rounds to 1.0; the cython rendering applies % twice
    assert cg.py_mod(-1e-17, 1.0) == 1.0
    assert cg.HostEval({"a": -1e-17, "b": 1.0}).eval("a % b") == 0.0
    assert cg.HostEval({"a": -7.5, "b": 2.0}).eval("a % b") == 0.5
    assert cg.HostEval({"a": -7, "b": 3}).eval("a % b") == 2
    assert cg.HostEval({"a": -7.5, "b": 2.0}).eval("a // b") == -4.0


def test_mod_renders_as_the_cython_target():
    c, t = R("a % b", a="double", b="double")
    assert c.count("hb_pymod(") == 2 and t == "double"
    c, t = R("a % b", a="float", b="float")
    assert c.count("hb_pymodf(") == 2 and t == "float"
    c, t = R("i % j", i="int", j="int")
    assert "hb_imod" in c and t == "int"
    c, t = R("a // b", a="double", b="double")
    assert c.startswith("floor(")
