"""Summed variables on the device against Brian's cython runtime (golden_gap.npz:
examples/synapses/gapjunctions.py at 300 neurons, all-to-all, per-synapse weights)."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _gap_network(meta):
    from brian2_amd import Model, Network, NeuronGroup, SpikeMonitor, Synapses

    inp = O.gap_inputs(meta["seed"])
    m = Model.load("gap")
    G = NeuronGroup(meta["n"], m, name="neurons")
    G.v = inp["v"]
    S = Synapses(G, G, summed={"Igap": m.summed["Igap"]}, name="gap")
    S.connect(inp["pre"], inp["post"])
    S.set("w", inp["w"])
    mon = SpikeMonitor(G)
    return Network(G, S, mon, dt=meta["dt"]), G, mon


def test_gap_junctions_bitexact_vs_brian():
    z, meta = O.load_golden("gap")
    net, G, mon = _gap_network(meta)
    net.run(meta["n_steps"])
    assert len(mon.i) == len(z["i"])
    assert np.array_equal(mon.i, z["i"]) and np.array_equal(mon.step, z["step"])
    assert np.array_equal(np.stack([G.get("v"), G.get("Igap")]), z["final"])


def test_gap_junctions_in_pieces():
    z, meta = O.load_golden("gap")
    net, G, mon = _gap_network(meta)
    for k in (1, 999, 1000, 3000):
        net.run(k)
    assert np.array_equal(mon.i, z["i"]) and np.array_equal(mon.step, z["step"])
    assert np.array_equal(G.get("v"), z["final"][0])
