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


def _full_range_sample():
    """Networks of the 1024-network sweep itself (N up to 8000): the largest ones and a
    few small ones, so that the kernel's neuron loop runs over several chunks."""
    from brian2_amd import synth

    nets = synth.sweep_networks(n_nets=1024)
    Ns = np.array([n.N for n in nets])
    big = [int(k) for k in np.argsort(-Ns)[:2]]
    small = [int(k) for k in np.argsort(Ns)[:1]]
    mid = [int(np.argmin(np.abs(Ns - 4100)))]
    pick = big + mid + small
    assert Ns[big[0]] > 7000
    return [nets[k] for k in pick]


def test_batched_full_size_range_equals_oracle():
    from brian2_amd import batched

    nets = _full_range_sample()
    print("batched vs oracle, N =", [n.N for n in nets], flush=True)
    B = batched.BatchedCUBA(nets)
    B.run(300)
    for k, net in enumerate(nets):
        meta = dict(N=net.N, NE=net.NE, taum=0.02, taue=0.005, taui=0.01, El=-0.049, Vt=-0.05, Vr=-0.06,
                    refractory=0.005, we=net.namespace["we"], wi=-0.009000000000000001, p=0.02, seed=net.seed)
        G, paths = O.build_cuba4000(meta)
        si, ss, _ = O.run([G], paths, 1e-4, 300)
        i, s = B.spikes(k)
        o = np.lexsort((i, s))
        assert len(si) > 0
        assert np.array_equal(i[o], si) and np.array_equal(s[o], ss), net.N
        for j, v in enumerate(("v", "ge", "gi")):
            assert np.array_equal(B.get(v, k), G.state[j]), (net.N, v)


def test_batched_fp32_equals_the_plain_fp32_path():
    """fp32 mode: the batched kernel and the plain path evaluate the same statements
    in the same order, so their fp32 results agree bit for bit as well."""
    from brian2_amd import batched, synth

    nets = _full_range_sample()[1:3]
    print("batched fp32 vs plain fp32, N =", [n.N for n in nets], flush=True)
    B = batched.BatchedCUBA(nets, float_t="float")
    B.run(400)
    for k, net in enumerate(nets):
        c = synth.cuba_network(N=net.N, NE=net.NE, seed=net.seed, we=net.namespace["we"], float_t="float")
        c["net"].run(400)
        i, s = B.spikes(k)
        o = np.lexsort((i, s))
        assert np.array_equal(i[o], c["mon"].i) and np.array_equal(s[o], c["mon"].step), net.N
        assert np.array_equal(B.get("v", k), c["G"].get("v")), net.N
