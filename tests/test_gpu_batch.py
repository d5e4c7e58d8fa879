"""Batches of independent SPD systems on the GPU: one block-diagonal factor,
multi-RHS solves, FP32 factors refined by batched FP64 CG.

Bars: FP64 batch solutions within 1e-12 (relative, max norm) of dense
per-system solves; after FP32 factorisation + batched CG every system and
right-hand side reaches berr = max|b-Ax| / (||A||_inf max|x| + max|b|)
<= 1e-12, and agrees with plain per-system FP64 solves to cond(A) * 1e-12
(<= 1e-8 relative here, cond(A) < 1e4 for these shifted 2D Laplacians).
"""
import numpy as np
import pytest
import scipy.sparse as sps

from pastix_amd import pastix, spm

pytestmark = pytest.mark.gpu


def _systems(ks, seed):
    rng = np.random.default_rng(seed)
    mats = [spm.laplacian5_2d(k) + rng.uniform(0.0, 1.0) * sps.eye(k * k) for k in ks]
    nrhs = rng.integers(1, 9, size=len(ks))
    bs = [rng.uniform(-1.0, 1.0, size=(k * k, r)) for k, r in zip(ks, nrhs)]
    return mats, bs


def _berr(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


def _batch(mats, ks, prec=None, graph=False):
    S = [spm.Spm(A, symmetric=True) for A in mats]
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    ip[pastix.iparm.refinement] = pastix.refine.CG
    h = pastix.init(ip, dp)
    if prec is not None:
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    if graph:
        pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 1)
    pastix.batch_analyze(h, S, [pastix.orderGrid(k, k, 1) for k in ks])
    pastix.batch_numfact(h, S)
    return h, S


def test_batch_fp64_solve_matches_dense():
    ks = [16, 23, 9, 31, 12, 40]
    mats, bs = _systems(ks, 4)
    h, S = _batch(mats, ks)
    xs = pastix.batch_solve(h, bs)
    for A, b, x in zip(mats, bs, xs):
        xd = np.linalg.solve(A.toarray(), b)
        assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max() * 10
    pastix.finalize(h)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_fp32_factor_fp64_cg(graph):
    ks = [16, 20, 33, 9, 27, 18, 40, 24]
    mats, bs = _systems(ks, 9)
    h, S = _batch(mats, ks, prec=spm.SpmFloat, graph=graph)
    x0 = pastix.batch_solve(h, bs)
    xs = pastix.batch_refine(h, bs, x0)
    assert h.iparm[pastix.iparm.nbiter] >= 1
    # the combined call runs the same solve and recurrence
    xc = pastix.batch_solve_refine(h, bs)
    for x, y in zip(xs, xc):
        assert np.abs(x - y).max() <= 1e-13 * np.abs(x).max()
    for k, A, b, x in zip(ks, mats, bs, xs):
        for c in range(b.shape[1]):
            assert _berr(A, x[:, c], b[:, c]) <= 1e-12
        # plain FP64 path, one system at a time
        ip, dp = pastix.initParam()
        ip[pastix.iparm.factorization] = pastix.factotype.LLT
        ip[pastix.iparm.ordering] = pastix.order.Personal
        h1 = pastix.init(ip, dp)
        S1 = spm.Spm(A, symmetric=True)
        o = pastix.orderGrid(k, k, 1)
        h1.order = o
        pastix.subtask_order(h1, S1, o)
        pastix.subtask_symbfact(h1)
        pastix.subtask_reordering(h1)
        pastix.subtask_blend(h1)
        pastix.task_numfact(h1, S1)
        xp = b.copy(order="F")
        pastix.task_solve(h1, S1, xp)
        assert np.abs(x - xp).max() <= 1e-8 * np.abs(xp).max()
        pastix.finalize(h1)
    pastix.finalize(h)


def test_c4_batch_512_systems_full_size():
    """C4 at its size: 512 shifted 2D Laplacians on k x k grids, k uniform in
    [16, 128], 1-8 right-hand sides each; one FP32 factor of the union (graph
    replay) + batched FP64 CG: every column of every system reaches berr <=
    1e-12, and a sample of systems agrees with plain per-system FP64 solves
    to cond(A) x 1e-12 (<= 1e-8 relative)."""
    rng = np.random.default_rng(7)
    ks = [int(k) for k in rng.integers(16, 129, size=512)]
    mats, bs = _systems(ks, 77)
    h, S = _batch(mats, ks, prec=spm.SpmFloat, graph=True)
    xs = pastix.batch_solve_refine(h, bs)
    worst, cols = 0.0, 0
    for A, b, x in zip(mats, bs, xs):
        R = b - A @ x
        anorm = abs(A).sum(axis=1).max()
        for c in range(b.shape[1]):
            worst = max(worst, float(np.abs(R[:, c]).max() / (anorm * np.abs(x[:, c]).max() + np.abs(b[:, c]).max())))
            cols += 1
    iters = int(h.iparm[pastix.iparm.nbiter])
    print(f"C4: 512 systems, {cols} right-hand sides, worst berr {worst:.2e}, CG iterations {iters}, "
          f"static pivots {h.iparm[pastix.iparm.static_pivoting]}")
    assert worst <= 1e-12
    # an FP32 factor of these systems (cond < 1e4) leaves CG a contraction of
    # about cond x 6e-8 per step: a few steps; many more mean a perturbed
    # factor (the union's norm once raised every system's static-pivot
    # threshold above its pivots)
    assert h.iparm[pastix.iparm.static_pivoting] == 0
    assert iters <= 4
    for i in range(0, 512, 64):
        xd = np.linalg.solve(mats[i].toarray(), bs[i]) if ks[i] <= 64 else None
        if xd is None:
            import scipy.sparse.linalg as spla
            xd = spla.spsolve(mats[i].tocsc(), bs[i]).reshape(bs[i].shape)
        assert np.abs(xs[i] - xd).max() <= 1e-8 * np.abs(xd).max()
    pastix.finalize(h)
