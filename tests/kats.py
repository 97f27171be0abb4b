"""Known-answer tests of the reference for the spike queue and synaptic transmission,
restated with explicit inputs (brian2/tests/test_spikequeue.py:36-61,
brian2/tests/test_synapses.py:1026-1218).  Shared by the oracle tests (CPU)
and the device tests (GPU).

Each case: a spike source (neuron indices and time steps), a passive target
("v : 1" or "v : integer"), synapses (i, j, w, delay in steps), on_pre
"v += w" (or "v += 1"), the number of steps, and the expected values of a
StateMonitor(when='end') as {target index: [v at step 0, 1, ...]}.
"""
import numpy as np

DT = 1e-4

CASES = {
    # test_transmission_all_to_one_heterogeneous_delays (test_synapses.py:1155-1178)
    "all_to_one_heterogeneous_delays": dict(
        n_src=6, n_tgt=1,
        spikes=([0, 1, 4, 5, 2, 3, 4, 5, 0, 1, 2, 3, 4, 5], [0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2]),
        i=[0, 1, 2, 3, 4, 5], j=[0, 0, 0, 0, 0, 0], w=[1, 2, 3, 4, 5, 6], delay=[0, 0, 0, 1, 2, 1],
        dtype="float64", steps=4, expect={0: [3, 12, 33, 48]}),
    # test_transmission_one_to_all_heterogeneous_delays (test_synapses.py:1180-1203), v : integer
    "one_to_all_heterogeneous_delays": dict(
        n_src=1, n_tgt=6, spikes=([0, 0], [0, 2]),
        i=[0, 0, 0, 0, 0, 0], j=[0, 1, 2, 3, 4, 5], w=None, delay=[1, 1, 2, 4, 3, 2],
        dtype="int32", steps=5,
        expect={0: [0, 1, 1, 2, 2], 1: [0, 1, 1, 2, 2], 2: [0, 0, 1, 1, 2], 3: [0, 0, 0, 0, 1],
                4: [0, 0, 0, 1, 1], 5: [0, 0, 1, 1, 2]}),
    # test_transmission_simple (test_synapses.py:1026-1037): spikes at 2 ms and 1 ms
    "simple": dict(
        n_src=2, n_tgt=2, spikes=([0, 1], [20, 10]), i=[0, 1], j=[0, 1], w=None, delay=[0, 0],
        dtype="float64", steps=25,
        expect={0: [0] * 20 + [1] * 5, 1: [0] * 10 + [1] * 15}),
    # test_transmission_scalar_delay (test_synapses.py:1205-1218): delay 0.5 ms
    "scalar_delay": dict(
        n_src=2, n_tgt=2, spikes=([0, 1], [0, 10]), i=[0, 1], j=[0, 1], w=None, delay=5,
        dtype="float64", steps=20,
        expect={0: [0] * 5 + [1] * 15, 1: [0] * 15 + [1] * 5}),
}


def expected_matrix(case):
    n = case["n_tgt"]
    out = np.zeros((n, case["steps"]))
    for k, seq in case["expect"].items():
        out[k] = seq
    return out


# test_spikequeue.py:36-61
def queue_one_to_one(N=100, dt=DT):
    sources = np.arange(N, dtype=np.int32)
    return sources, sources * dt


def queue_all_to_all(N=100, dt=DT):
    sources = np.repeat(np.arange(N, dtype=np.int32), N)
    return sources, sources * dt
