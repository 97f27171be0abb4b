"""Pins the oracle (oracle/hbref.c) before it is trusted as the checker.

- against Brian's cython runtime (golden fixtures written by tests/golden/make_golden.py);
- against the reference's own CSpikeQueue compiled from brian2/synapses/spikequeue.h
  (oracle/_ref, built by `make -C oracle ref` where the reference tree exists);
- against the reference's known-answer tests (tests/kats.py);
- against what the mathematics fixes (closed-form CUBA solution, the exponential
  Euler formula of the HH equations, libm accuracy of the elementary functions).
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import kats
import oracle_py as O

REPO = O.REPO
REF_CSQ = os.path.join(REPO, "oracle", "_ref", "libcspikequeue.so")


# --------------------------------------------------------------------------- goldens
def test_oracle_cuba4000_bitexact_vs_brian():
    z, meta = O.load_golden("cuba4000")
    G, paths = O.build_cuba4000(meta)
    si, ss, rec = O.run([G], paths, meta["dt"], meta["n_steps"], rec_idx=z["rec"], rec_vars=[0, 1, 2])
    assert np.array_equal(si, z["i"]) and np.array_equal(ss, z["step"])
    assert np.array_equal(G.state, z["final"])
    assert np.array_equal(np.transpose(rec, (1, 2, 0))[:, :, : z["trace"].shape[2]], z["trace"])
    assert np.array_equal(G.lastspike, z["lastspike"])


def test_oracle_delays_weights_bitexact_vs_brian():
    z, meta = O.load_golden("cuba_delay")
    G, paths = O.build_cuba_delay(meta)
    si, ss, rec = O.run([G], paths, meta["dt"], meta["n_steps"], rec_idx=z["rec"], rec_vars=[0, 1, 2])
    assert np.array_equal(si, z["i"]) and np.array_equal(ss, z["step"])
    assert np.array_equal(G.state, z["final"])
    assert np.array_equal(np.transpose(rec, (1, 2, 0))[:, :, : z["trace"].shape[2]], z["trace"])


def test_oracle_cobahh_bitexact_vs_brian_first_100ms():
    z, meta = O.load_golden("cobahh4000")
    G, paths = O.build_cobahh(meta)
    n = 1000
    si, ss, rec = O.run([G], paths, meta["dt"], n, rec_idx=z["rec"], rec_vars=list(range(6)))
    sel = z["step"] < n
    assert np.array_equal(si, z["i"][sel]) and np.array_equal(ss, z["step"][sel])
    assert np.array_equal(np.transpose(rec, (1, 2, 0)), z["trace"][:, :, :n])


def test_golden_cuba_equals_unmodified_brian():
    """The CUBA golden was also produced with Brian's unmodified libm exp: identical."""
    z, meta = O.load_golden("cuba4000")
    assert meta["identical_to_unmodified_libm"] is True


# --------------------------------------------------------------------------- statements
def _spec(name):
    with open(os.path.join(REPO, "tests", "golden", f"spec_{name}.json")) as f:
        return {c["template"] + ":" + c["owner"]: c for c in json.load(f)}


def _vec(c):
    return [(s["var"], s["op"], s["expr"]) for s in c["blocks"]["None"]["vector"]]


def _sca(c):
    return [(s["var"], s["op"], s["expr"]) for s in c["blocks"]["None"]["scalar"]]


def test_oracle_cuba_restates_brians_statements():
    """hbref.c's cuba_scalar/cuba_update restate exactly these statements."""
    sp = _spec("cuba")
    su = sp["stateupdate:neurongroup_1"]
    assert _sca(su) == [
        ("_lio_1", ":=", "timestep(0.005, dt)"), ("_lio_2", ":=", "exp((- dt) / taue)"),
        ("_lio_3", ":=", "exp((- dt) / taui)"), ("_lio_4", ":=", "taue * exp((- dt) / taue)"),
        ("_lio_5", ":=", "- exp(dt / taue)"), ("_lio_6", ":=", "dt / taum"), ("_lio_7", ":=", "(- dt) / taum"),
        ("_lio_8", ":=", "0.0 - taum"), ("_lio_9", ":=", "taui * exp((- dt) / taui)"),
        ("_lio_10", ":=", "- exp(dt / taui)"), ("_lio_11", ":=", "El - El"),
        ("_lio_12", ":=", "El - (El * exp(_lio_7))"),
        ("_lio_13", ":=", "((_lio_4 * (_lio_5 + exp(_lio_6))) * exp(_lio_7)) / (_lio_8 + taue)"),
        ("_lio_14", ":=", "((_lio_9 * (_lio_10 + exp(_lio_6))) * exp(_lio_7)) / (_lio_8 + taui)"),
        ("_lio_15", ":=", "exp(_lio_7)")]
    vec = su["blocks"]["None"]["vector"]
    assert [(s["var"], s["op"]) for s in vec] == [("not_refractory", "="), ("_ge", ":="), ("_gi", ":="),
                                                  ("_v", ":="), ("ge", "="), ("gi", "="), ("v", "=")]
    assert vec[0]["expr"] == "timestep(t - lastspike, dt) >= _lio_1"
    assert vec[1]["expr"] == "_lio_2 * ge" and vec[2]["expr"] == "_lio_3 * gi"
    assert dict((tuple(tuple(a) for a in k), e) for k, e in vec[3]["bool_exprs"]) == {
        (("not_refractory", False),): "_lio_11 + v",
        (("not_refractory", True),): "_lio_12 + (((_lio_13 * ge) + (_lio_14 * gi)) + (_lio_15 * v))"}
    assert su["blocks"]["None"]["conditional_write"] == {"v": "not_refractory"}
    assert _vec(sp["threshold:neurongroup_1"]) == [("_cond", ":=", "(v > Vt) and not_refractory")]
    assert _vec(sp["reset:neurongroup_1"]) == [("v", "=", "Vr")]
    assert _vec(sp["synapses:synapses"]) == [("ge", "+=", "we")]
    assert _vec(sp["synapses:synapses_1"]) == [("gi", "+=", "wi")]


def test_oracle_cobahh_restates_brians_statements():
    sp = _spec("cobahh")
    su = sp["stateupdate:neurongroup"]
    names = [s[0] for s in _sca(su)]
    assert names == [f"_lio_{k}" for k in range(1, 28)]
    vec = _vec(su)
    assert [v for v, _, _ in vec] == ["not_refractory", "_ge", "_gi", "_BA_h", "_h", "_BA_m", "_m", "_BA_n", "_n",
                                      "_BA_v", "_v", "ge", "gi", "h", "m", "n", "v"]
    # the two expressions with exp/exprel/** that hbref.c restates character by character
    assert dict((v, e) for v, _, e in vec)["_BA_v"] == (
        "(_lio_19 + ((((_lio_20 * (n ** 4)) + (_lio_21 * (h * (m ** 3)))) + (_lio_22 * ge)) + (_lio_23 * gi))) / "
        "((_lio_24 + (_lio_25 * (n ** 4))) - (((_lio_26 * (h * (m ** 3))) + (_lio_27 * ge)) + (_lio_27 * gi)))")
    assert dict((v, e) for v, _, e in vec)["_BA_m"] == (
        "_lio_10 / (((_lio_11 / exprel(_lio_12 - (_lio_13 * v))) - (_lio_14 / exprel(_lio_15 + (_lio_7 * v)))) * "
        "exprel(_lio_12 - (_lio_13 * v)))")
    assert _vec(sp["threshold:neurongroup"]) == [("_cond", ":=", "(v > _lio_1) and not_refractory")]


def test_product_specs_are_brians_statements():
    """brian2_amd/specs/*.json hold exactly the captured code objects."""
    for model, golden in (("cuba", "cuba"), ("cobahh", "cobahh"), ("cuba_synw", "cuba_delay")):
        with open(os.path.join(REPO, "brian2_amd", "specs", f"{model}.json")) as f:
            prod = json.load(f)
        with open(os.path.join(REPO, "tests", "golden", f"spec_{golden}.json")) as f:
            cap = json.load(f)
        for c in cap:
            if c["template"] in ("stateupdate", "threshold", "reset"):
                assert prod["group"][c["template"]] == c
            elif c["template"] == "synapses":
                role = {"synapses": "exc", "synapses_1": "inh"}[c["owner"]]
                assert prod["pathways"][role] == c


# --------------------------------------------------------------------------- mathematics
def _cuba_closed_form(v0, ge0, gi0, t, taum, taue, taui, El):
    """Exact solution of dv/dt=(ge+gi-(v-El))/taum, dge/dt=-ge/taue, dgi/dt=-gi/taui."""
    # v(t) = El + (v0-El) e^{-t/taum} + sum_x x0 * tau_x/(tau_x - taum) (e^{-t/tau_x} - e^{-t/taum})
    v = El + (v0 - El) * math.exp(-t / taum)
    for x0, tx in ((ge0, taue), (gi0, taui)):
        v += x0 * tx / (tx - taum) * (math.exp(-t / tx) - math.exp(-t / taum))
    return v, ge0 * math.exp(-t / taue), gi0 * math.exp(-t / taui)


def test_oracle_cuba_update_matches_closed_form():
    """No threshold crossing (Vt raised): 200 steps of the restated exact update vs the
    analytic solution of the linear system (pins the _lio structure to 1e-12)."""
    taum, taue, taui, El = 0.02, 0.005, 0.01, -0.049
    N = 5
    v0 = np.array([-0.06, -0.055, -0.05, -0.045, -0.07])
    ge0 = np.array([0.0, 0.001, 0.002, -0.001, 0.003])
    gi0 = np.array([0.0, -0.002, 0.001, 0.0, -0.004])
    G = O.OracleGroup(O.HBR_CUBA_EXACT, N, np.stack([v0, ge0, gi0]), [taum, taue, taui, El, 1.0, -0.06, 0.005])
    n = 200
    O.run([G], [], 1e-4, n, mon_group=None)
    for k in range(N):
        v, ge, gi = _cuba_closed_form(v0[k], ge0[k], gi0[k], n * 1e-4, taum, taue, taui, El)
        assert abs(G.state[0, k] - v) <= 1e-12 * abs(v)
        assert abs(G.state[1, k] - ge) <= 1e-12 * max(abs(ge), 1e-300)
        assert abs(G.state[2, k] - gi) <= 1e-12 * max(abs(gi), 1e-300)


def test_oracle_cobahh_step_is_exponential_euler():
    """One step of the restated COBAHH update equals x' = (x + B/A) e^{A dt} - B/A with
    A, B taken from the HH equations directly (examples/COBAHH.py), to 1e-10."""
    z, meta = O.load_golden("cobahh4000")
    m = meta
    rng = np.random.default_rng(0)
    N = 64
    v = rng.uniform(-0.08, 0.03, N)
    gm, gn, gh = rng.uniform(0.01, 0.99, (3, N))
    ge, gi = rng.uniform(0, 1e-7, (2, N))
    G = O.OracleGroup(O.HBR_COBAHH_EE, N, np.stack([v, gm, gn, gh, ge, gi]), O.cobahh_params(m))
    O.run([G], [], m["dt"], 1, mon_group=None)
    dt = m["dt"]
    mV, ms, VT = 1e-3, 1e-3, m["VT"]

    def exprel(x):
        return 1.0 if abs(x) < 1e-16 else math.expm1(x) / x

    for k in range(N):
        vv = v[k]
        am = 0.32 * 4 / exprel((13 * mV - vv + VT) / (4 * mV)) / ms
        bm = 0.28 * 5 / exprel((vv - VT - 40 * mV) / (5 * mV)) / ms
        ah = 0.128 * math.exp((17 * mV - vv + VT) / (18 * mV)) / ms
        bh = 4.0 / (1 + math.exp((40 * mV - vv + VT) / (5 * mV))) / ms
        an = 0.032 * 5 / exprel((15 * mV - vv + VT) / (5 * mV)) / ms
        bn = 0.5 * math.exp((10 * mV - vv + VT) / (40 * mV)) / ms

        def ee(x, A, B):
            return (x + B / A) * math.exp(A * dt) - B / A

        exp_m = ee(gm[k], -(am + bm), am)
        exp_n = ee(gn[k], -(an + bn), an)
        exp_h = ee(gh[k], -(ah + bh), ah)
        gna = m["g_na"] * gm[k] ** 3 * gh[k]
        gkd = m["g_kd"] * gn[k] ** 4
        A = -(m["gl"] + ge[k] + gi[k] + gna + gkd) / m["Cm"]
        B = (m["gl"] * m["El"] + ge[k] * m["Ee"] + gi[k] * m["Ei"] + gna * m["ENa"] + gkd * m["EK"]) / m["Cm"]
        exp_v = ee(vv, A, B)
        for got, want in ((G.state[0, k], exp_v), (G.state[1, k], exp_m), (G.state[2, k], exp_n),
                          (G.state[3, k], exp_h)):
            assert abs(got - want) <= 1e-10 * max(abs(want), 1e-12)
        assert G.state[4, k] == ge[k] * math.exp(-dt / m["taue"]) or abs(
            G.state[4, k] - ge[k] * math.exp(-dt / m["taue"])) <= 1e-15 * ge[k]


def _ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-2**63) - ia, ia)
    ib = np.where(ib < 0, np.int64(-2**63) - ib, ib)
    return np.abs(ia - ib)


def test_shared_math_accuracy_vs_libm():
    """hb_math (used by device and oracle) is within 1 ulp of glibc on random arguments."""
    L = O.lib()
    rng = np.random.default_rng(1)
    x = rng.uniform(-700, 700, 20000)
    assert max(_ulps([L.hbr_exp(t) for t in x], np.exp(x))) <= 1
    y = rng.uniform(-3, 3, 20000)
    assert max(_ulps([L.hbr_expm1(t) for t in y], np.expm1(y))) <= 1
    p = np.exp(rng.uniform(-200, 200, 20000))
    assert max(_ulps([L.hbr_log(t) for t in p], np.log(p))) <= 1
    q = rng.uniform(-1, 1, 20000)
    assert max(_ulps([L.hbr_pow(a, b) for a, b in zip(p[:20000], q)], [math.pow(a, b) for a, b in zip(p, q)])) <= 1
    # exact cases
    assert L.hbr_exp(0.0) == 1.0 and L.hbr_log(1.0) == 0.0 and L.hbr_pow(2.0, 10.0) == 1024.0
    assert math.isinf(L.hbr_exp(710.0)) and L.hbr_exp(-746.0) == 0.0


def test_shared_math_is_the_products():
    """The product's host exports (libhipbrian.so) compute the same bits as the oracle's."""
    from brian2_amd import _lib

    L = O.lib()
    rng = np.random.default_rng(2)
    for x in rng.uniform(-50, 50, 2000):
        assert _lib.lib.hb_host_exp(x) == L.hbr_exp(x)
        assert _lib.lib.hb_host_expm1(x / 10) == L.hbr_expm1(x / 10)
        ax = abs(x) + 1e-3
        assert _lib.lib.hb_host_log(ax) == L.hbr_log(ax)
        assert _lib.lib.hb_host_pow(ax, x / 7) == L.hbr_pow(ax, x / 7)


# --------------------------------------------------------------------------- queue
def test_oracle_queue_kats():
    """test_spikequeue.py:36-61: one-to-one and all-to-all with delays i*dt."""
    N = 100
    src, d = kats.queue_one_to_one(N)
    q = O.Queue(0, N)
    q.prepare(d, src, kats.DT)
    q.push(np.arange(N, dtype=np.int32))
    for i in range(N):
        assert np.array_equal(q.peek(), [i])
        q.advance()
    for i in range(N):
        assert len(q.peek()) == 0
        q.advance()
    src, d = kats.queue_all_to_all(N)
    q = O.Queue(0, N)
    q.prepare(d, src, kats.DT)
    q.push(np.arange(N, dtype=np.int32))
    for i in range(N):
        assert np.array_equal(q.peek(), i * N + np.arange(N))
        q.advance()


@pytest.mark.skipif(not os.path.exists(REF_CSQ), reason="oracle/_ref not built (needs the reference tree)")
def test_oracle_queue_vs_reference_cspikequeue():
    """Random sequences against the reference's CSpikeQueue (spikequeue.h) itself."""
    R = ctypes.CDLL(REF_CSQ)
    R.csq_create.restype = ctypes.c_void_p
    R.csq_create.argtypes = [ctypes.c_int, ctypes.c_int]
    R.csq_destroy.argtypes = [ctypes.c_void_p]
    R.csq_prepare.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                              ctypes.POINTER(ctypes.c_int32), ctypes.c_int, ctypes.c_double]
    R.csq_push.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
    R.csq_peek.restype = ctypes.c_long
    R.csq_peek.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_long]
    R.csq_advance.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(3)
    for trial in range(12):
        n_src = int(rng.integers(1, 40))
        start = int(rng.integers(0, 5))
        S = int(rng.integers(0, 400))
        sources = rng.integers(start, start + n_src, S).astype(np.int32)
        scalar = trial % 3 == 0
        dt = 1e-4
        delays = (np.array([rng.integers(0, 30)]) if scalar else rng.integers(0, 30, S)) * dt
        delays = delays.astype(np.float64) + rng.uniform(-0.3, 0.3, len(delays)) * dt
        delays = np.maximum(delays, 0)
        rq = R.csq_create(start, start + n_src)
        oq = O.Queue(start, start + n_src)
        R.csq_prepare(rq, delays.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(delays),
                      sources.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), S, dt)
        oq.prepare(delays, sources, dt)
        for step in range(80):
            spk = np.sort(rng.choice(np.arange(start - 2, start + n_src + 2), int(rng.integers(0, 8)),
                                     replace=False)).astype(np.int32)
            R.csq_push(rq, spk.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(spk))
            oq.push(spk)
            buf = np.empty(100000, np.int32)
            n = R.csq_peek(rq, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(buf))
            assert np.array_equal(buf[:n], oq.peek()), (trial, step)
            R.csq_advance(rq)
            oq.advance()
        R.csq_destroy(rq)


@pytest.mark.skipif(not os.path.exists(REF_CSQ), reason="oracle/_ref not built (needs the reference tree)")
def test_oracle_queue_dt_changes_vs_reference_cspikequeue():
    """Random push sequences with dt changed between runs (re-preparing with the same
    delays in seconds), against the reference's CSpikeQueue: the re-binning of queued
    spikes (spikequeue.h:58-74), including factors below 1 where later slots replace
    earlier ones."""
    R = ctypes.CDLL(REF_CSQ)
    R.csq_create.restype = ctypes.c_void_p
    R.csq_create.argtypes = [ctypes.c_int, ctypes.c_int]
    R.csq_destroy.argtypes = [ctypes.c_void_p]
    R.csq_prepare.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                              ctypes.POINTER(ctypes.c_int32), ctypes.c_int, ctypes.c_double]
    R.csq_push.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
    R.csq_peek.restype = ctypes.c_long
    R.csq_peek.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_long]
    R.csq_advance.argtypes = [ctypes.c_void_p]
    R.csq_size.restype = ctypes.c_long
    R.csq_size.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(5)
    P = ctypes.POINTER
    for trial in range(16):
        n_src, start = int(rng.integers(1, 30)), int(rng.integers(0, 4))
        S = int(rng.integers(1, 300))
        sources = rng.integers(start, start + n_src, S).astype(np.int32)
        delays = (rng.integers(0, 40, S) * 1e-4).astype(np.float64)
        rq, oq = R.csq_create(start, start + n_src), O.Queue(start, start + n_src)
        dt = 1e-4
        for run in range(4):
            R.csq_prepare(rq, delays.ctypes.data_as(P(ctypes.c_double)), S, sources.ctypes.data_as(P(ctypes.c_int32)),
                          S, dt)
            oq.prepare(delays, sources, dt)
            assert R.csq_size(rq) == oq.size(), (trial, run)
            for step in range(int(rng.integers(3, 30))):
                spk = np.sort(rng.choice(np.arange(start, start + n_src), min(n_src, int(rng.integers(0, 5))),
                                         replace=False)).astype(np.int32)
                R.csq_push(rq, spk.ctypes.data_as(P(ctypes.c_int32)), len(spk))
                oq.push(spk)
                buf = np.empty(100000, np.int32)
                n = R.csq_peek(rq, buf.ctypes.data_as(P(ctypes.c_int32)), len(buf))
                assert np.array_equal(buf[:n], oq.peek()), (trial, run, step)
                R.csq_advance(rq)
                oq.advance()
            # factors old/new dt of at least 0.5 (collisions of old slots included); below 0.5 the
            # reference can index past its vector (see hbr_queue_prepare)
            dt = dt * float(rng.choice([0.1, 0.25, 0.5, 2.0 / 3.0, 1.0, 1.5, 2.0]))
        R.csq_destroy(rq)


def test_oracle_dt_changes_bitexact_vs_brian():
    """The CUBA-with-delays model run in four pieces with dt changed between them
    (0.1, 0.05, 0.1, 0.2 ms; spikes in flight at every change), against Brian's
    cython runtime: spike indices and times, final state and lastspike."""
    z, meta = O.load_golden("cuba_dt_change")
    i, t, G = O.run_dt_changes(meta)
    assert len(i) > 1000
    assert np.array_equal(i, z["i"]) and np.array_equal(t, z["t"])
    assert np.array_equal(G.state, z["final"])
    assert np.array_equal(G.lastspike, z["lastspike"])


def test_oracle_queue_dt_change_rebins():
    """spikequeue.h:58-74: queued spikes are re-binned when dt changes between runs."""
    q = O.Queue(0, 3)
    src = np.array([0, 1, 2], np.int32)
    q.prepare(np.array([0.0, 2e-4, 4e-4]), src, 1e-4)
    q.push(np.array([0, 1, 2], np.int32))
    q.advance()
    # now halve dt: the remaining entries at delays 1 and 3 (in old steps) move to 2 and 6
    q.prepare(np.array([0.0, 2e-4, 4e-4]), src, 0.5e-4)
    seen = {}
    for k in range(10):
        for s in q.peek():
            seen[int(s)] = k
        q.advance()
    assert seen == {1: 2, 2: 6}


# --------------------------------------------------------------------------- transmission KATs
def _oracle_kat(case):
    n_src, n_tgt = case["n_src"], case["n_tgt"]
    idx, steps = (np.asarray(a, np.int32) for a in case["spikes"])
    o = np.lexsort((idx, steps))
    src = O.OracleGroup(O.HBR_SPIKEGEN, n_src, np.zeros((1, n_src)), [0.0], gen=(idx[o], steps[o]))
    tgt = O.OracleGroup(O.HBR_PASSIVE, n_tgt, np.zeros((1, n_tgt)), [0.0])
    w = None if case["w"] is None else np.asarray(case["w"], np.float64)
    d = case["delay"]
    path = O.OraclePathway(src, tgt, case["i"], case["j"], 0, w=w, w_scalar=1.0,
                           delay=None if np.ndim(d) == 0 else np.asarray(d, np.float64) * kats.DT,
                           delay_scalar=float(d) * kats.DT if np.ndim(d) == 0 else 0.0)
    n = case["steps"]
    _, _, rec = O.run([src, tgt], [path], kats.DT, n, mon_group=None, rec_group=1,
                      rec_idx=np.arange(n_tgt), rec_vars=[0])
    # when='end' values: the 'start' value of the next step, and the final state
    start = rec[:, 0, :].T
    return np.concatenate([start[:, 1:], tgt.state[0][:, None]], axis=1)


@pytest.mark.parametrize("name", sorted(kats.CASES))
def test_oracle_transmission_kats(name):
    case = kats.CASES[name]
    assert np.array_equal(_oracle_kat(case), kats.expected_matrix(case))


def test_oracle_subgroup_propagation_kat():
    """test_subgroup.py:718-729: neurons 1, 3, 5 of G1 spike at step 0 (v = 1.1 > 1);
    Synapses(G1[1:6], G2[10:]) connected i == j within the subgroups (absolute
    indices 1..5 -> 10..14), on_pre v += 1: G2.v is 1 at 10, 12, 14 after one step."""
    src = O.OracleGroup(O.HBR_SPIKEGEN, 10, np.zeros((1, 10)), [0.0],
                        gen=(np.array([1, 3, 5], np.int32), np.zeros(3, np.int32)))
    tgt = O.OracleGroup(O.HBR_PASSIVE, 20, np.zeros((1, 20)), [0.0])
    path = O.OraclePathway(src, tgt, np.arange(1, 6), np.arange(10, 15), 0, w_scalar=1.0, delay_scalar=0.0,
                           source_range=(1, 6))
    O.run([src, tgt], [path], kats.DT, 1, mon_group=None)
    expected = np.zeros(20)
    expected[[10, 12, 14]] = 1
    assert np.array_equal(tgt.state[0], expected)


def test_burst_fold_restatement_equals_oracle():
    """The sequential fold of tests/bursts.py (used as the device expectation) is
    the oracle's spike queue + passive target, value for value."""
    import bursts as B

    pre, post, w, d = B.burst_case()
    exp = B.expected_trace(pre, post, w, d)
    n_tgt = B.N_TGT + B.N_HUB
    idx = np.tile(np.arange(B.N_SRC, dtype=np.int32), len(B.SPIKE_STEPS))
    steps = np.repeat(np.array(B.SPIKE_STEPS, np.int32), B.N_SRC)
    o = np.lexsort((idx, steps))
    src = O.OracleGroup(O.HBR_SPIKEGEN, B.N_SRC, np.zeros((1, B.N_SRC)), [0.0], gen=(idx[o], steps[o]))
    tgt = O.OracleGroup(O.HBR_PASSIVE, n_tgt, np.zeros((1, n_tgt)), [0.0])
    path = O.OraclePathway(src, tgt, pre, post, 0, w=w.astype(np.float64), delay=d.astype(np.float64) * kats.DT)
    _, _, rec = O.run([src, tgt], [path], kats.DT, B.STEPS, mon_group=None, rec_group=1,
                      rec_idx=np.arange(n_tgt), rec_vars=[0])
    start = rec[:, 0, :].T
    got = np.concatenate([start[:, 1:], tgt.state[0][:, None]], axis=1)
    assert np.array_equal(got, exp)
    # a different order changes some value: sources reversed (every source spikes at the
    # same steps), or the synapse order reversed (only multapses share source and delay)
    assert not np.array_equal(exp, B.expected_trace(B.N_SRC - 1 - pre, post, w, d))
    assert not np.array_equal(exp, B.expected_trace(pre[::-1], post[::-1], w[::-1], d[::-1]))


def test_oracle_lif_equals_brian_golden():
    """The differentiable LIF configuration's forward pass (Poisson trains, two
    current pathways, exact integration, no refractoriness): the oracle
    reproduces Brian's cython run bit for bit (tests/golden/golden_lif.npz)."""
    z, meta = O.load_golden("lif")
    inp = O.lif_inputs(meta["trials"], meta["seed"])
    for b in range(meta["trials"]):
        src, G, paths = O.build_lif(inp, b)
        si, ss, rec = O.run([src, G], paths, meta["dt"], meta["n_steps"], mon_group=1, rec_group=1,
                            rec_idx=[0, 1, 2, 3], rec_vars=[0, 1])
        sel = z["trial"] == b
        assert len(si) == sel.sum() > 1000
        assert np.array_equal(si, z["i"][sel]) and np.array_equal(ss, z["step"][sel])
        assert np.array_equal(G.state, z["final"][b])
        assert np.array_equal(np.transpose(rec, (1, 2, 0)), z["trace"][b])


def test_poisson_oracle_rates_and_independence():
    """The oracle's PoissonGroup threshold (rand() < rates * dt, poissongroup.py:106) against
    what the mathematics fixes: per-neuron counts within binomial bounds of n * rates * dt,
    z-scores of mean 0 and spread 1 over the population, and no excess of spikes in
    consecutive steps (independent draws); a rates/dt mix-up or a reversed comparison
    fails these bounds, and a run started later must give the same trains (the draw
    of a neuron depends on the seed and the absolute step only)."""
    N, steps, seed, dt = 2000, 20000, 31, 1e-4
    rates = 5.0 + 195.0 * O.uniform(5, 3, N)
    g = O.OracleGroup(O.HBR_POISSON, N, np.zeros(N), [], refractory=False, poisson=(seed, rates * dt))
    i, s = O.run([g], [], dt, steps, mon_group=0)[:2]
    p = rates * dt
    cnt = np.bincount(i, minlength=N)
    z = (cnt - steps * p) / np.sqrt(steps * p * (1 - p))
    assert np.all(np.abs(z) < 5.5)
    assert abs(z.mean()) < 0.15 and 0.9 < z.std() < 1.1
    order = np.lexsort((s, i))
    i2, s2 = i[order], s[order]
    consecutive = np.count_nonzero((i2[1:] == i2[:-1]) & (s2[1:] == s2[:-1] + 1))
    expect = np.sum((steps - 1) * p * p)
    assert abs(consecutive - expect) < 5 * np.sqrt(expect)
    # the same trains from a later start step: draws depend on (seed, step, neuron) only
    g2 = O.OracleGroup(O.HBR_POISSON, N, np.zeros(N), [], refractory=False, poisson=(seed, rates * dt))
    i3, s3 = O.run([g2], [], dt, 500, start_step=7000, mon_group=0)[:2]
    sel = (s >= 7000) & (s < 7500)
    assert np.array_equal(i3, i[sel]) and np.array_equal(s3, s[sel])


def test_oracle_openmp_equals_one_thread():
    """The OpenMP leg of the oracle (the CPU baseline's multi-core figure) folds each
    target's events in queue order: state, spikes and per-synapse counts are
    identical to one thread's."""
    z, meta = O.load_golden("cuba_delay")
    out = []
    for threads in (1, 4):
        G, paths = O.build_cuba_delay(meta)
        for p in paths:
            p.counts = np.zeros(len(p.pre), np.int32)
        i, s, _ = O.run([G], paths, meta["dt"], 1500, threads=threads)
        out.append((i, s, G.state.copy(), [p.counts.copy() for p in paths]))
    (i1, s1, st1, c1), (i4, s4, st4, c4) = out
    assert np.array_equal(i1, z["i"][z["step"] < 1500]) and np.array_equal(i1, i4) and np.array_equal(s1, s4)
    assert np.array_equal(st1, st4)
    assert all(np.array_equal(a, b) for a, b in zip(c1, c4))
    assert sum(int(c.sum()) for c in c1) > 100_000


def test_oracle_queues_kept_across_runs():
    """With the queues kept across runs (a pathway keeps its SpikeQueue), runs in
    pieces equal one run, spikes in flight included; fresh queues per run lose them."""
    z, meta = O.load_golden("cuba_delay")
    G1, p1 = O.build_cuba_delay(meta)
    O.run([G1], p1, meta["dt"], 600)
    G2, p2 = O.build_cuba_delay(meta)
    q = O.OracleQueues(len(p2))
    a = O.run([G2], p2, meta["dt"], 250, queues=q)
    b = O.run([G2], p2, meta["dt"], 350, start_step=250, queues=q)
    assert np.array_equal(G1.state, G2.state)
    i = np.concatenate([a[0], b[0]])
    assert np.array_equal(i, z["i"][z["step"] < 600])
    G3, p3 = O.build_cuba_delay(meta)
    O.run([G3], p3, meta["dt"], 250)
    O.run([G3], p3, meta["dt"], 350, start_step=250)
    assert not np.array_equal(G1.state, G3.state)
