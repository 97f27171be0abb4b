"""CUBA on the device vs Brian's cython runtime (golden fixture) and the oracle."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _run_cuba(n_steps, record=None, **kw):
    from brian2_amd import synth

    z, meta = O.load_golden("cuba4000")
    c = synth.cuba_network(N=meta["N"], NE=meta["NE"], p=meta["p"], seed=meta["seed"], record=record, **kw)
    c["net"].run(n_steps)
    return c, z, meta


def test_connectivity_matches_oracle_generator():
    from brian2_amd import synth

    z, meta = O.load_golden("cuba4000")
    c = synth.cuba_network(N=meta["N"], NE=meta["NE"], p=meta["p"], seed=meta["seed"])
    for S, (a, b) in ((c["Ce"], (0, meta["NE"])), (c["Ci"], (meta["NE"], meta["N"]))):
        seg = S.csr["seg_ptr"].cpu().numpy().view(np.uint32).astype(np.int64)
        post = S.csr["post"].cpu().numpy()
        pre_o, post_o = O.bernoulli(meta["seed"], a, b, meta["N"], meta["p"])
        pre_d = np.repeat(np.arange(meta["N"]), np.diff(seg))
        assert np.array_equal(pre_d, pre_o) and np.array_equal(post, post_o)


def test_cuba4000_1s_bitexact_vs_brian():
    c, z, meta = _run_cuba(10000, record=[0, 1, 2, 3])
    i, step = c["mon"].i, c["mon"].step
    assert len(i) == len(z["i"])
    assert np.array_equal(i, z["i"]) and np.array_equal(step, z["step"])
    final = np.stack([c["G"].get(n) for n in ("v", "ge", "gi")])
    assert np.array_equal(final, z["final"])
    tr = c["state"].values()[:, :, : z["trace"].shape[2]]
    assert np.array_equal(tr, z["trace"])
    assert np.array_equal(c["G"].get("lastspike"), z["lastspike"])
    cnt = np.bincount(z["i"], minlength=meta["N"])
    assert np.array_equal(c["mon"].count, cnt)


def test_cuba_runs_in_pieces_equal_one_run():
    c1, z, meta = _run_cuba(700)
    from brian2_amd import synth

    c2 = synth.cuba_network(N=meta["N"], NE=meta["NE"], p=meta["p"], seed=meta["seed"])
    for k in (1, 99, 100, 250, 250):
        c2["net"].run(k)
    assert np.array_equal(c1["mon"].i, c2["mon"].i) and np.array_equal(c1["mon"].step, c2["mon"].step)
    for n in ("v", "ge", "gi"):
        assert np.array_equal(c1["G"].get(n), c2["G"].get(n))


def test_cuba_host_connectivity_equals_device_connectivity():
    from brian2_amd import synth

    z, meta = O.load_golden("cuba4000")
    a = synth.cuba_network(N=meta["N"], NE=meta["NE"], p=meta["p"], seed=meta["seed"], device_conn=False)
    a["net"].run(500)
    sel = z["step"] < 500
    assert np.array_equal(a["mon"].i, z["i"][sel]) and np.array_equal(a["mon"].step, z["step"][sel])


def test_cuba_fp32_mode_error_reported():
    c, z, meta = _run_cuba(2000, float_t="float")
    sel = z["step"] < 2000
    n64 = int(sel.sum())
    n32 = len(c["mon"].i)
    # fp32 mode is not bit-exact; the spike count stays close to the fp64 oracle's
    assert abs(n32 - n64) <= 0.05 * n64


@pytest.mark.parametrize("shape", [{}, {"tile": 1024}], ids=["default", "tile1024-fused"])
def test_store_restore_replays_bit_identically(shape):
    """Network.store/restore (brian2/core/network.py:600-740): a restored network
    repeats the same spikes, state and event counts; profile() (one launch at a
    time) gives the same results as graph replay."""
    from brian2_amd import synth

    z, meta = O.load_golden("cuba_delay")
    c = synth.cuba_synw_network(N=meta["N"], K=meta["K"], seed=meta["seed"], ginh=meta["ginh"],
                                network_options=shape)
    net, mon = c["net"], c["mon"]
    net.run(300)
    net.store()
    n0 = len(mon.i)
    ev0 = net.n_events
    net.run(400)
    a_i, a_s = mon.i[n0:], mon.step[n0:]
    a_state = [c["G"].get(v) for v in ("v", "ge", "gi")]
    a_ev = net.n_events - ev0
    assert len(a_i) > 0
    net.restore()
    assert net.t_step == 300 and len(mon.i) == n0 and net.n_events == ev0
    net.run(400)
    assert np.array_equal(mon.i[n0:], a_i) and np.array_equal(mon.step[n0:], a_s)
    assert all(np.array_equal(c["G"].get(v), x) for v, x in zip(("v", "ge", "gi"), a_state))
    assert net.n_events - ev0 == a_ev
    net.restore()
    prof = net.profile(400)
    assert [k for k, _, _ in prof] == (["fused"] if shape else ["step", "gather", "gather"])
    net.run(0)  # folds the events of the last profiled step
    assert np.array_equal(mon.i[n0:], a_i) and np.array_equal(mon.step[n0:], a_s)
    assert all(np.array_equal(c["G"].get(v), x) for v, x in zip(("v", "ge", "gi"), a_state))
