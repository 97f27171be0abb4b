"""Plastic synapses (STDP: on_pre and on_post pathways, event-driven traces) on the
device against Brian's cython runtime (golden_stdp.npz, written by
tests/golden/make_golden.py from examples/synapses/STDP.py with generator inputs)."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

SYN_VARS = ("w", "Apre", "Apost", "lastupdate")


def _stdp_network(meta, **network_options):
    from brian2_amd import Model, Network, NeuronGroup, SpikeGeneratorGroup, SpikeMonitor, Synapses

    inp = O.stdp_inputs(meta["seed"])
    m = Model.load("stdp")
    src = SpikeGeneratorGroup(meta["n_in"], inp["idx"], timebins=inp["steps"])
    G = NeuronGroup(meta["n_out"], m, name="neurons")
    G.v = inp["v0"]
    S = Synapses(src, G, m.pathways["pre"], on_post=m.pathways["post"], name="stdp")
    S.connect(inp["pre"], inp["post"])
    S.set("w", inp["w"])
    mon = SpikeMonitor(G)
    net = Network(src, G, S, mon, dt=meta["dt"], **network_options)
    return net, G, S, mon


@pytest.mark.parametrize("shape", [{}, {"tile": 64, "gather_pairs": 16}], ids=["default", "tile64-pairs16"])
def test_stdp_bitexact_vs_brian(shape):
    z, meta = O.load_golden("stdp")
    net, G, S, mon = _stdp_network(meta, **shape)
    net.run(meta["n_steps"])
    assert len(mon.i) == len(z["i"])
    assert np.array_equal(mon.i, z["i"]) and np.array_equal(mon.step, z["step"])
    assert np.array_equal(np.stack([G.get("v"), G.get("ge")]), z["final"])
    for k, n in enumerate(SYN_VARS):
        assert np.array_equal(S.get(n), z["syn"][k]), n


def test_stdp_store_restore_and_pieces():
    z, meta = O.load_golden("stdp")
    net, G, S, mon = _stdp_network(meta)
    net.run(3000)
    net.store()
    net.run(2000)
    a = [S.get(n) for n in SYN_VARS] + [G.get("v")]
    net.restore()
    for k in (700, 1, 1299):
        net.run(k)
    b = [S.get(n) for n in SYN_VARS] + [G.get("v")]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    sel = z["step"] < 5000
    assert np.array_equal(mon.i, z["i"][sel]) and np.array_equal(mon.step, z["step"][sel])
