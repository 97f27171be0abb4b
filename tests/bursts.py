"""A synchronous-burst transmission case (shared by the oracle and device tests).

2000 sources all spike at steps 0-4; 300 passive targets with in-degree 150 and
3 hub targets with in-degree 1000 receive "v += w" with fp64 weights spread
over six decades (so any change of summation order changes the result), delays
0-20 steps, and 2000 multapses (same source, target and delay as another
synapse).  expected_trace() is the sequential fp64 left fold of each target's
events in the cython runtime's delivery order: delivery step, then push step,
then source, then synapse index (brian2/synapses/spikequeue.h:151-200,
brian2/codegen/runtime/cython_rt/templates/synapses.pyx:42-51).
"""
import numpy as np

N_SRC, N_TGT, N_HUB, K, K_HUB, DMAX, STEPS = 2000, 300, 3, 150, 1000, 20, 30
SPIKE_STEPS = (0, 1, 2, 3, 4)


def burst_case(seed=7):
    rng = np.random.default_rng(seed)
    pre, post = [], []
    for j in range(N_TGT + N_HUB):
        k = K if j < N_TGT else K_HUB
        pre.append(rng.choice(N_SRC, size=k, replace=False))
        post.append(np.full(k, j))
    pre = np.concatenate(pre)
    post = np.concatenate(post)
    d = rng.integers(0, DMAX + 1, len(pre))
    # 2000 multapses: a second synapse on the same (source, target) with the same delay,
    # whose order against the first is the synapse index
    dup = rng.choice(len(pre), size=2000, replace=False)
    pre, post, d = np.concatenate([pre, pre[dup]]), np.concatenate([post, post[dup]]), np.concatenate([d, d[dup]])
    perm = rng.permutation(len(pre))  # synapse order unrelated to (source, target)
    pre, post, d = pre[perm], post[perm], d[perm]
    w = rng.uniform(-1, 1, len(pre)) * 10.0 ** rng.uniform(-3, 3, len(pre))
    return pre, post, w, d


def expected_trace(pre, post, w, d):
    """v of every target at the end of each step: sequential fp64 folds."""
    n_syn = len(pre)
    syn = np.tile(np.arange(n_syn), len(SPIKE_STEPS))
    t_spike = np.repeat(np.array(SPIKE_STEPS), n_syn)
    deliver = t_spike + d[syn]
    order = np.lexsort((syn, pre[syn], t_spike, deliver))
    v = [0.0] * (N_TGT + N_HUB)
    out = np.zeros((N_TGT + N_HUB, STEPS))
    k = 0
    ev_deliver = deliver[order]
    ev_syn = syn[order]
    for s in range(STEPS):
        while k < len(order) and ev_deliver[k] == s:
            e = ev_syn[k]
            v[post[e]] = v[post[e]] + float(w[e])
            k += 1
        out[:, s] = v
    return out
