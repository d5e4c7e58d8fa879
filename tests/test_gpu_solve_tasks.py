"""Dependency-driven solve (one launch, device counters) against the
level-by-level kernels: same diagonal inverses and products, so the results
agree to rounding, for LL^t and LU, 1 and several right-hand sides, FP64 and
FP32 factors; and it is deterministic from run to run."""
import numpy as np
import pytest

from pastix_amd import pastix, spm
from tests.helpers import analyse

pytestmark = pytest.mark.gpu

LLT, LU = pastix.factotype.LLT, pastix.factotype.LU


def _solve(h, b, n, nrhs, tasks):
    pastix.set_amd_iparm(h, pastix.amd_iparm.task_solve, tasks)
    db = pastix.DeviceArray(b)
    pastix.solve_device(h, db, n, nrhs)
    x = db.download()
    db.free()
    return x


@pytest.mark.parametrize("facto", [LLT, LU])
@pytest.mark.parametrize("nrhs", [1, 5, 11])
def test_task_solve_matches_level_solve(facto, nrhs):
    grid = (14, 12, 11)
    if facto == LLT:
        A = spm.laplacian7(*grid, diag=6.0)
    else:
        A = spm.convection_diffusion27(*grid, seed=4)
    h, S, sym = analyse(A, facto == LLT, facto, grid=grid, min_bs=16, max_bs=96)
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, nrhs, seed=3)
    n = A.shape[0]
    xl = _solve(h, b, n, nrhs, 0)
    xt = _solve(h, b, n, nrhs, 1)
    xt2 = _solve(h, b, n, nrhs, 1)
    assert np.array_equal(xt, xt2), "task solve not deterministic"
    assert np.linalg.norm(xt - xl) <= 1e-13 * np.linalg.norm(xl)
    assert np.linalg.norm(xt - x0) <= 1e-9 * np.linalg.norm(x0)
    pastix.finalize(h)


def test_task_solve_fp32_factor_and_graph():
    grid = (13, 12, 10)
    A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = LLT
    ip[pastix.iparm.ordering] = pastix.order.Personal
    ip[pastix.iparm.refinement] = pastix.refine.CG
    h = pastix.init(ip, dp)
    pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, spm.SpmFloat)
    o = pastix.orderGrid(*grid)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_blend(h)
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, 3, seed=8)
    n = A.shape[0]
    b32 = b.astype(np.float32)
    xl = _solve(h, b32, n, 3, 0)
    xt = _solve(h, b32, n, 3, 1)
    assert np.linalg.norm(xt - xl) <= 1e-5 * np.linalg.norm(xl)
    # captured solve graph, replayed with eager work in between
    pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 1)
    db = pastix.DeviceArray(b32)
    for rep in range(3):
        db.upload(b32)
        pastix.solve_device(h, db, n, 3)
        assert np.array_equal(db.download(), xt), rep
        pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 0)
        assert np.array_equal(_solve(h, b32, n, 3, 1), xt)
        pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 1)
    db.free()
    # refinement on top of the task solve reaches the FP64 target
    dB = pastix.DeviceArray(b)
    dX = pastix.DeviceArray(np.zeros_like(b))
    pastix.solve_refine_device(h, dB, dX, n, 3)
    X = dX.download()
    for c in range(3):
        r = b[:, c] - A @ X[:, c]
        berr = np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(X[:, c]).max() + np.abs(b[:, c]).max())
        assert berr <= 1e-12
    dB.free()
    dX.free()
    pastix.finalize(h)
