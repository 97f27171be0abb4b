"""Batched sweep (one persistent kernel for many networks) vs the plain path and the oracle."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _sweep(n_nets, **kw):
    from brian2_amd import batched, synth

    nets = synth.sweep_networks(n_nets=n_nets, nmin=500, nmax=2000)
    return nets, batched.BatchedCUBA(nets, **kw)


def test_batched_equals_plain_path_bitexact():
    from brian2_amd import synth

    nets, B = _sweep(6)
    B.run(600)
    for k, net in enumerate(nets):
        c = synth.cuba_network(N=net.N, NE=net.NE, seed=net.seed, we=net.namespace["we"])
        c["net"].run(600)
        i, s = B.spikes(k)
        o = np.lexsort((i, s))
        assert np.array_equal(i[o], c["mon"].i) and np.array_equal(s[o], c["mon"].step), k
        for v in ("v", "ge", "gi"):
            assert np.array_equal(B.get(v, k), c["G"].get(v)), (k, v)


def test_batched_equals_oracle():
    nets, B = _sweep(3)
    B.run(300)
    for k, net in enumerate(nets):
        meta = dict(N=net.N, NE=net.NE, taum=0.02, taue=0.005, taui=0.01, El=-0.049, Vt=-0.05, Vr=-0.06,
                    refractory=0.005, we=net.namespace["we"], wi=-0.009000000000000001, p=0.02, seed=net.seed)
        G, paths = O.build_cuba4000(meta)
        si, ss, _ = O.run([G], paths, 1e-4, 300)
        i, s = B.spikes(k)
        o = np.lexsort((i, s))
        assert np.array_equal(i[o], si) and np.array_equal(s[o], ss)
        assert np.array_equal(B.get("v", k), G.state[0])


def test_batched_results_do_not_depend_on_tuning():
    outs = []
    for block, spl in ((128, 10), (512, 50), (1024, 100)):
        nets, B = _sweep(4)
        B.block, B.steps_per_launch = block, spl
        B.run(300)
        outs.append([(B.spikes(k), B.get("ge", k)) for k in range(4)])
    for o in outs[1:]:
        for (a, ga), (b, gb) in zip(outs[0], o):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(ga, gb)


def test_batched_autotuner_keeps_state():
    nets, B = _sweep(4)
    B.run(100)
    ref = [B.get("v", k) for k in range(4)]
    best, res = B.autotune(blocks=(128, 256), steps=(10, 50), trial_steps=100)
    assert best in res
    assert all(np.array_equal(B.get("v", k), ref[k]) for k in range(4))
