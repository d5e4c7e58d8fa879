"""Size-independent properties at the benchmark sizes, where the oracle would
take too long: backward errors of the FP64 solve and of the FP32 factor with
FP64 refinement (the targets of the hot path), and agreement of the FP32 +
refinement solution with the FP64 one to cond(A) x 1e-12."""
import numpy as np
import pytest

from pastix_amd import pastix, spm

pytestmark = pytest.mark.gpu


def _berr(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


def _instance(A, grid, facto, prec=None):
    S = spm.Spm(A, symmetric=facto == pastix.factotype.LLT)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = facto
    ip[pastix.iparm.ordering] = pastix.order.Personal
    ip[pastix.iparm.refinement] = pastix.refine.CG if facto == pastix.factotype.LLT else pastix.refine.GMRES
    if facto == pastix.factotype.LU:
        dp[pastix.dparm.epsilon_magn_ctrl] = np.ldexp(1.0, -52)
    h = pastix.init(ip, dp)
    if prec is not None:
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    o = pastix.orderGrid(*grid)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    pastix.task_numfact(h, S)
    return h, S


@pytest.mark.parametrize("facto,grid", [(pastix.factotype.LLT, (80, 80, 80)), (pastix.factotype.LU, (64, 64, 64))])
def test_fullsize_fp64_and_fp32_refined(facto, grid):
    if facto == pastix.factotype.LLT:
        A = spm.laplacian7(*grid, diag=6.0)
    else:
        A = spm.convection_diffusion27(*grid, seed=5)
    n = A.shape[0]
    x0, b = spm.rhs_from_x(A, 1, seed=2024)
    h, S = _instance(A, grid, facto)
    assert h.iparm[pastix.iparm.static_pivoting] == 0
    db = pastix.DeviceArray(b)
    pastix.solve_device(h, db, n, 1)
    x64 = db.download()[:, 0]
    db.free()
    e64 = _berr(A, x64, b[:, 0])
    pastix.finalize(h)
    h, S = _instance(A, grid, facto, prec=spm.SpmFloat)
    db = pastix.DeviceArray(b)
    dx = pastix.DeviceArray(np.zeros_like(b))
    pastix.solve_refine_device(h, db, dx, n, 1)
    x32 = dx.download()[:, 0]
    db.free()
    dx.free()
    e32 = _berr(A, x32, b[:, 0])
    print(grid, "berr fp64", e64, "fp32+refinement", e32, "iters", h.iparm[pastix.iparm.nbiter])
    assert e64 <= 1e-15
    assert e32 <= 1e-12
    assert np.linalg.norm(x32 - x64) <= 1e-8 * np.linalg.norm(x64)
    pastix.finalize(h)
