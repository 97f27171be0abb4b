"""Synchronous bursts (tests/bursts.py) through every path of the gather and the
fold, against the sequential fp64 fold in the cython runtime's delivery order
(itself checked against the oracle on the CPU).

Launch shapes force multi-pass folds (LDS capacity 128 events), per-target
global sorting (hub targets above half that capacity), more events per tile than
the fold's registers hold (the passes read the rest again), gathers of a few
pairs per unit (many windows per step) and tiles of 64 and 1024.
"""
import numpy as np
import pytest

from bursts import N_HUB, N_SRC, N_TGT, SPIKE_STEPS, STEPS, burst_case, expected_trace

pytestmark = pytest.mark.gpu


def device_trace(pre, post, w, d, **shape):
    from brian2_amd import Network, NeuronGroup, SpikeGeneratorGroup, StateMonitor, Synapses
    from test_gpu_kats import onpre_spec

    idx = np.tile(np.arange(N_SRC), len(SPIKE_STEPS))
    steps = np.repeat(np.array(SPIKE_STEPS), N_SRC)
    src = SpikeGeneratorGroup(N_SRC, idx, timebins=steps)
    tgt = NeuronGroup(N_TGT + N_HUB, variables={"v": "float64"})
    S = Synapses(src, tgt, onpre_spec("w", ("w",)))
    S.connect(pre, post)
    S.set("w", w)
    S.set_delay(d.astype(np.float64) * 1e-4)
    mon = StateMonitor(tgt, ["v"], record=np.arange(N_TGT + N_HUB), when="end")
    net = Network(src, tgt, S, mon, dt=1e-4, **shape)
    net.run(STEPS)
    return mon.values()[0]


@pytest.fixture(scope="module")
def case():
    pre, post, w, d = burst_case()
    return pre, post, w, d, expected_trace(pre, post, w, d)


@pytest.mark.parametrize("shape", [
    {},
    {"tile": 64, "bucket_cap": 128, "gather_pairs": 32},
    {"tile": 1024, "gather_pairs": 1024},
    {"tile": 256, "bucket_cap": 512, "gather_pairs": 64},
    {"tile": 128, "npt": 4, "bucket_cap": 256, "gather_pairs": 48},
    {"tile": 1024, "npt": 2},
], ids=["default", "tile64-cap128-pairs32", "tile1024-pairs1024", "tile256-cap512-pairs64", "tile128-npt4-cap256",
        "tile1024-npt2"])
def test_burst_trace_bitexact(case, shape):
    pre, post, w, d, exp = case
    print(f"burst case, launch shape {shape}", flush=True)
    got = device_trace(pre, post, w, d, **shape)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("shape", [
    {},
    {"tile": 64, "bucket_cap": 128, "gather_pairs": 32},
    {"overlap": False},
    {"tile": 1024},
    {"tile": 1024, "fuse": False},
    {"tile": 1024, "fuse": {"allow_spills": True, "events_per_thread": 2}},
    {"tile": 1024, "bucket_cap": 128, "gather_pairs": 32, "fuse": {"tiles_first": True, "gather_blocks": [3]}},
], ids=["default", "tile64-cap128-pairs32", "no-overlap", "tile1024-fused", "tile1024-lanes",
        "tile1024-fused-2percu", "tile1024-fused-tilesfirst-cap128"])
def test_burst_trace_bitexact_with_concurrent_gathers(case, shape):
    """Delays 1-21 steps: the gathers of a step run concurrently with the group
    update of that step (triple-buffered event staging), on other streams or, for
    tiles of 1024 threads, in the group's launch."""
    pre, post, w, d, _ = case
    d1 = d + 1
    exp = expected_trace(pre, post, w, d1)
    print(f"burst case with delays >= 1, launch shape {shape}", flush=True)
    got = device_trace(pre, post, w, d1, **shape)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("shape", [
    {"tile": 64, "bucket_cap": 128, "gather_pairs": 32},
    {"tile": 1024, "fuse": False},
], ids=["tile64-cap128-pairs32", "tile1024-separate"])
def test_burst_trace_bitexact_unsorted_gather_windows(case, shape, monkeypatch):
    """The gathers' other write-out, for windows whose tile sort does not fit the LDS
    (many target tiles and large windows): each event takes its place in its tile's
    run with an LDS atomic and is written from registers.  Forced here by switching
    the sort off; delays 1-21 steps."""
    from brian2_amd import kernels

    monkeypatch.setattr(kernels, "GATHER_STAGE_RUNS", False)
    pre, post, w, d, _ = case
    d1 = d + 1
    exp = expected_trace(pre, post, w, d1)
    print(f"burst case with delays >= 1, unsorted gather windows, launch shape {shape}", flush=True)
    got = device_trace(pre, post, w, d1, **shape)
    assert np.array_equal(got, exp)
