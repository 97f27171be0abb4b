"""Per-synapse delivered-event counts on the device vs the oracle.

The reference's way of counting events per synapse is a synaptic integer
incremented by on_pre (brian2/tests/test_synapses.py:1063-1091): one increment
per synapse id that the spike queue delivers (templates/synapses.pyx:42-50).
The device counts them in the gather (Synapses(count_events=True)), one thread
per event with a plain increment: correct because a synapse is delivered at most
once per step and pathway, which these tests hold to account on the
CUBA-with-delays configuration (fixed in-degree 1000, delays 1-50 steps)."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _oracle_counts(meta, n_steps):
    G, paths = O.build_cuba_delay(meta)
    for p in paths:
        p.counts = np.zeros(len(p.pre), np.int32)
    i, s, _ = O.run([G], paths, meta["dt"], n_steps)
    Ci, Ce = paths
    return {"synapses": Ce.counts, "synapses_1": Ci.counts}, i, s


@pytest.mark.parametrize("shape", [{}, {"fuse": False}, {"tile": 64, "gather_pairs": 64}],
                         ids=["fused", "separate", "tile64"])
def test_event_counts_equal_the_oracles(shape):
    from brian2_amd import synth

    z, meta = O.load_golden("cuba_delay")
    n_steps = meta["n_steps"]
    print(f"event counts, cuba_delay golden config, launch shape {shape}", flush=True)
    c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                ginh=meta["ginh"], network_options=shape, count_events=True)
    c["net"].run(n_steps)
    assert np.array_equal(c["mon"].i, z["i"]) and np.array_equal(c["mon"].step, z["step"])
    want, _, _ = _oracle_counts(meta, n_steps)
    pre, post, d, w32, exc = O.cuba_delay_synapses(meta)
    total = 0
    for S, sel in ((c["Ce"], exc), (c["Ci"], ~exc)):
        got = S.event_counts()
        # device synapses are stored in (source, delay, target) order, the oracle's in (source, target)
        o = np.lexsort((post[sel], d[sel], pre[sel]))
        assert np.array_equal(got, want[S.name][o]), S.name
        assert int(got.sum()) == S.n_events
        total += int(got.sum())
    assert total > 1_000_000
    # at most one delivery per synapse and step: no count exceeds the spikes of its source
    spikes_of = np.bincount(z["i"], minlength=meta["N"])
    for S, sel in ((c["Ce"], exc), (c["Ci"], ~exc)):
        o = np.lexsort((post[sel], d[sel], pre[sel]))
        assert np.all(S.event_counts() <= spikes_of[pre[sel][o]])


def test_event_counts_accumulate_over_runs():
    """Counts are totals since the network was built (runs in pieces add up)."""
    from brian2_amd import synth

    z, meta = O.load_golden("cuba_delay")
    c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], dmin=meta["dmin"], dmax=meta["dmax"], seed=meta["seed"],
                                ginh=meta["ginh"], count_events=True)
    for k in (1, 299, 700):
        c["net"].run(k)
    want, _, _ = _oracle_counts(meta, 1000)
    pre, post, d, w32, exc = O.cuba_delay_synapses(meta)
    for S, sel in ((c["Ce"], exc), (c["Ci"], ~exc)):
        o = np.lexsort((post[sel], d[sel], pre[sel]))
        assert np.array_equal(S.event_counts(), want[S.name][o]), S.name
