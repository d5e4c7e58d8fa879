"""Batches of independent systems analysed as one block-diagonal union
(host analysis only: no GPU needed)."""
import numpy as np
import scipy.sparse as sps

from pastix_amd import pastix, spm


def _systems(ks, seed):
    rng = np.random.default_rng(seed)
    return [spm.laplacian5_2d(k) + rng.uniform(0.0, 1.0) * sps.eye(k * k) for k in ks]


def _alone_flops(A, k):
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    ip[pastix.iparm.ordering] = pastix.order.Personal
    h = pastix.init(ip, dp)
    S = spm.Spm(A, symmetric=True)
    o = pastix.orderGrid(k, k, 1)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    sym = pastix.symbol(h)
    fl = pastix.flops(h)[0]
    pastix.finalize(h)
    return fl, sym["cblknbr"]


def test_batch_union_equals_sum_of_systems():
    ks = [8, 12, 9, 16, 5]
    mats = _systems(ks, 1)
    S = [spm.Spm(A, symmetric=True) for A in mats]
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    h = pastix.init(ip, dp)
    pastix.batch_analyze(h, S, [pastix.orderGrid(k, k, 1) for k in ks])
    off = pastix.batch_offsets(h)
    assert list(off) == list(np.concatenate([[0], np.cumsum([k * k for k in ks])]))
    alone = [_alone_flops(A, k) for A, k in zip(mats, ks)]
    # the union's flop count and supernodes are those of the systems together
    assert pastix.flops(h)[0] == sum(a[0] for a in alone)
    assert pastix.symbol(h)["cblknbr"] == sum(a[1] for a in alone)
    # every system keeps its own range of the permuted numbering
    perm = pastix.symbol(h)["permtab"]
    for s in range(len(ks)):
        blk = perm[off[s]:off[s + 1]]
        assert blk.min() == off[s] and blk.max() == off[s + 1] - 1
    pastix.finalize(h)


def test_batch_default_orderings_stay_per_system():
    ks = [7, 10, 6]
    S = [spm.Spm(A, symmetric=True) for A in _systems(ks, 2)]
    h = pastix.init(*pastix.initParam())
    pastix.batch_analyze(h, S, None)
    off = pastix.batch_offsets(h)
    perm = pastix.symbol(h)["permtab"]
    for s in range(len(ks)):
        blk = np.sort(perm[off[s]:off[s + 1]])
        assert np.array_equal(blk, np.arange(off[s], off[s + 1]))
    pastix.finalize(h)
