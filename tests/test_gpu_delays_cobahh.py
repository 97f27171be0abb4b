"""Delays + per-synapse fp32 weights, and COBAHH (exponential_euler), on the device
vs Brian's cython runtime (golden fixtures) and the oracle."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def test_synw_device_csr_matches_oracle_generator():
    """Device fixed in-degree sampling + (source, delay) CSR + fp32 weights == the oracle's generator."""
    from brian2_amd import synth

    z, meta = O.load_golden("cuba_delay")
    c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                ginh=meta["ginh"])
    pre, post, d, w32, exc = O.cuba_delay_synapses(meta)
    n_seg = meta["dmax"] - meta["dmin"] + 1
    for S, sel in ((c["Ce"], exc), (c["Ci"], ~exc)):
        a = S.src_range[0]
        seg = S.csr["seg_ptr"].cpu().numpy().view(np.uint32).astype(np.int64)
        dpost = S.csr["post"].cpu().numpy()
        dw = S.csr["syn"]["w"].cpu().numpy()[: len(dpost)]
        counts = np.diff(seg)
        rows = np.repeat(np.arange(len(counts)), counts)
        dpre = a + rows // n_seg
        ddel = meta["dmin"] + rows % n_seg
        # oracle order: (pre, delay, post)
        o = np.lexsort((post[sel], d[sel], pre[sel]))
        assert np.array_equal(dpre, pre[sel][o])
        assert np.array_equal(ddel, d[sel][o])
        assert np.array_equal(dpost, post[sel][o])
        assert np.array_equal(dw, w32[sel][o])


@pytest.mark.parametrize("shape", [
    {},
    {"tile": 64, "bucket_cap": 128, "gather_pairs": 32},
    {"tile": 1024, "gather_pairs": 1024},
    {"tile": 256, "npt": 3},
], ids=["default", "tile64-cap128-pairs32", "tile1024-pairs1024", "tile256-npt3"])
def test_synw_delays_bitexact_vs_brian(shape):
    """Launch shapes force multi-pass folds, per-target global sorting (in-degree
    1000 above half of an LDS capacity of 128) and multi-window gather units."""
    from brian2_amd import synth

    z, meta = O.load_golden("cuba_delay")
    print(f"cuba_delay golden, launch shape {shape}", flush=True)
    c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                ginh=meta["ginh"], record=[0, 1, 2, 3], network_options=shape)
    c["net"].run(meta["n_steps"])
    assert np.array_equal(c["mon"].i, z["i"]) and np.array_equal(c["mon"].step, z["step"])
    final = np.stack([c["G"].get(n) for n in ("v", "ge", "gi")])
    assert np.array_equal(final, z["final"])
    tr = c["state"].values()[:, :, : z["trace"].shape[2]]
    assert np.array_equal(tr, z["trace"])


def test_synw_determinism_run_twice():
    from brian2_amd import synth

    z, meta = O.load_golden("cuba_delay")
    outs = []
    for _ in range(2):
        c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], seed=meta["seed"], ginh=meta["ginh"])
        c["net"].run(600)
        outs.append((c["mon"].i, c["mon"].step, c["G"].get("v"), c["G"].get("ge"), c["G"].get("gi")))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_cobahh_bitexact_vs_brian():
    from brian2_amd import synth

    z, meta = O.load_golden("cobahh4000")
    c = synth.cobahh_network(N=meta["N"], p=meta["p"], seed=meta["seed"], record=[0, 1, 2, 3])
    assert np.array_equal(c["G"].get("v")[[0, 1, 2, 3]], z["trace"][0, :, 0])
    c["net"].run(meta["n_steps"])
    tr = c["state"].values()[:, :, : z["trace"].shape[2]]
    assert np.array_equal(tr, z["trace"])
    assert np.array_equal(c["mon"].i, z["i"]) and np.array_equal(c["mon"].step, z["step"])
    names = ["v", "m", "n", "h", "ge", "gi"]
    final = np.stack([c["G"].get(n) for n in names])
    assert np.array_equal(final, z["final"])


def test_cobahh_fp32_error_vs_fp64():
    """fp32 mode: error against the fp64 oracle over 100 ms (reported, bounded)."""
    from brian2_amd import synth

    z, meta = O.load_golden("cobahh4000")
    c = synth.cobahh_network(N=meta["N"], p=meta["p"], seed=meta["seed"], float_t="float", record=[0, 1, 2, 3])
    c["net"].run(1000)
    v32 = c["state"].values()[0]
    v64 = z["trace"][0, :, :1000]
    err = np.max(np.abs(v32 - v64))
    n64 = int((z["step"] < 1000).sum())
    n32 = len(c["mon"].i)
    print(f"COBAHH fp32 vs fp64 over 1000 steps: max |dv| = {err:.3e} V, spikes {n32} vs {n64}")
    assert abs(n32 - n64) <= 0.05 * n64 + 5


@pytest.mark.parametrize("shape", [{}, {"tile": 1024}], ids=["default", "tile1024-fused"])
def test_synw_delays_with_a_whole_group_pathway(shape):
    """A third pathway from the whole group (zero weights into ge, delays 1-3 steps,
    last in the fold order: ge + 0.0 == ge) reads both parts of the group's spike ring
    (the exc and inh pathways read one each); results equal Brian's."""
    from brian2_amd import Synapses, synth
    from test_gpu_stdp import _zero_ge_spec

    z, meta = O.load_golden("cuba_delay")
    c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                ginh=meta["ginh"], record=[0, 1, 2, 3], network_options=shape)
    G = c["G"]
    rng = np.random.default_rng(3)
    n = 20000
    Z = Synapses(G, G, _zero_ge_spec(), name="zzz")
    Z.connect(rng.integers(0, G.N, n), rng.integers(0, G.N, n))
    Z.set("w", 0.0)
    Z.set_delay(rng.integers(1, 4, n) * 1e-4)
    c["net"].add(Z)
    print(f"cuba_delay golden + whole-group pathway, launch shape {shape}", flush=True)
    c["net"].run(meta["n_steps"])
    assert len(G.ring_parts) == 2 and len(Z.kernel.ring_parts[2]) == 2
    assert np.array_equal(c["mon"].i, z["i"]) and np.array_equal(c["mon"].step, z["step"])
    final = np.stack([G.get(n_) for n_ in ("v", "ge", "gi")])
    assert np.array_equal(final, z["final"])


def test_fused_and_separate_launches_agree_in_fp32():
    """fp32 state (the reported-error mode): the fused launch and separate launches run
    the same bodies in the same fold order, so their spikes and final state are
    identical bit for bit."""
    from brian2_amd import synth

    z, meta = O.load_golden("cuba_delay")
    out = []
    for opt in ({"tile": 1024}, {"tile": 1024, "fuse": False}):
        c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                    ginh=meta["ginh"], float_t="float", network_options=opt)
        c["net"].run(3000)
        assert (c["G"].fused is not None) == opt.get("fuse", True)
        out.append((c["mon"].i.copy(), c["mon"].step.copy(), [c["G"].get(n) for n in ("v", "ge", "gi")]))
    (i0, s0, f0), (i1, s1, f1) = out
    assert len(i0) > 100
    assert np.array_equal(i0, i1) and np.array_equal(s0, s1)
    assert all(np.array_equal(a, b) for a, b in zip(f0, f1))
