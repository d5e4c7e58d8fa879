"""Schur complement API on the GPU (reference sopalin/schur.c,
kernels/cpucblk_zschur.c, the interface sweeps of kernels/solve_ztrsmsp.c:
383-400,513-546), checked as the reference's example/schur.c uses it:

- pastixGetSchur returns S = A22 - A21 A11^{-1} A12 for the listed unknowns
  (increasing index order), pinned by the dense Schur complement of the
  permuted-free matrix at 1e-12 relative Frobenius (LL^t: lower triangle);
- the example's solve: forward sweep in PastixSolvModeInterface, the caller
  solves S x2 = b2 on the last rows, backward sweep in Interface mode: the
  solution of A x = b (berr <= 1e-13);
- PastixSolvModeLocal (the default) solves A11 x1 = b1 and leaves the Schur
  rows of b as they are.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from pastix_amd import pastix, spm

pytestmark = pytest.mark.gpu
F = pastix.factotype


def _berr(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


def _setup(facto, grid, nschur, mode, prec=None):
    if facto == F.LU:
        A = spm.convection_diffusion27(*grid, seed=3)
    else:
        A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.3))
    n = A.shape[0]
    S = spm.Spm(A, symmetric=facto != F.LU)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = facto
    ip[pastix.iparm.min_blocksize] = 16
    ip[pastix.iparm.max_blocksize] = 32
    h = pastix.init(ip, dp)
    if prec is not None:
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    # the example's list: the first unknowns (here a given count)
    lst = np.arange(nschur, dtype=np.int64)
    pastix.setSchurUnknownList(h, lst)
    ip[pastix.iparm.schur_solv_mode] = mode
    pastix.task_analyze(h, S)
    pastix.task_numfact(h, S)
    return A, S, h, lst


@pytest.mark.parametrize("facto", [F.LLT, F.LU])
def test_schur_complement_and_interface_solve(facto):
    grid = (9, 8, 7)
    n = 9 * 8 * 7
    ns = n // 3
    A, S, h, lst = _setup(facto, grid, ns, pastix.solv_mode.Interface)
    Sd = np.zeros((ns, ns), order="F")
    pastix.getSchur(h, Sd)
    Ad = A.toarray()
    I = lst
    J = np.setdiff1d(np.arange(n), I)
    Sref = Ad[np.ix_(I, I)] - Ad[np.ix_(I, J)] @ np.linalg.solve(Ad[np.ix_(J, J)], Ad[np.ix_(J, I)])
    if facto == F.LU:
        err = np.linalg.norm(Sd - Sref) / np.linalg.norm(Sref)
    else:
        err = np.linalg.norm(np.tril(Sd) - np.tril(Sref)) / np.linalg.norm(np.tril(Sref))
        assert not np.triu(Sd, 1).any()
    assert err <= 1e-12, err
    # example/schur.c: applyorder, forward (Interface), the caller's Schur
    # solve on the last ns rows, backward (Interface), applyorder back
    rng = np.random.default_rng(5)
    x0 = rng.uniform(-1, 1, (n, 2))
    b = A @ x0
    x = np.asfortranarray(b.copy())
    pastix.subtask_applyorder(h, pastix.dir.Forward, x)
    if facto == F.LLT:
        pastix.subtask_trsm(h, pastix.side.Left, pastix.uplo.Lower, pastix.trans.NoTrans, pastix.diag.NonUnit, x)
        Lc = np.linalg.cholesky(np.tril(Sd) + np.tril(Sd, -1).T)
        x[n - ns:] = sla.cho_solve((Lc, True), x[n - ns:])
        pastix.subtask_trsm(h, pastix.side.Left, pastix.uplo.Lower, pastix.trans.ConjTrans, pastix.diag.NonUnit, x)
    else:
        pastix.subtask_trsm(h, pastix.side.Left, pastix.uplo.Lower, pastix.trans.NoTrans, pastix.diag.Unit, x)
        x[n - ns:] = sla.lu_solve(sla.lu_factor(Sd), x[n - ns:])
        pastix.subtask_trsm(h, pastix.side.Left, pastix.uplo.Upper, pastix.trans.NoTrans, pastix.diag.NonUnit, x)
    pastix.subtask_applyorder(h, pastix.dir.Backward, x)
    for c in range(2):
        assert _berr(A, x[:, c], b[:, c]) <= 1e-13
    assert np.abs(x - x0).max() <= 1e-10
    # the Schur mode would solve with the unfactored Schur part
    h.iparm[pastix.iparm.schur_solv_mode] = pastix.solv_mode.Schur
    with pytest.raises(pastix.PastixError):
        pastix.task_solve(h, S, np.asfortranarray(b.copy()))
    pastix.finalize(h)


@pytest.mark.parametrize("facto", [F.LLT, F.LU])
def test_schur_local_mode_solves_the_factored_part(facto):
    grid = (8, 7, 6)
    n = 8 * 7 * 6
    ns = 50
    A, S, h, lst = _setup(facto, grid, ns, pastix.solv_mode.Local)
    Ad = A.toarray()
    I = lst
    J = np.setdiff1d(np.arange(n), I)
    rng = np.random.default_rng(6)
    b = rng.uniform(-1, 1, (n, 3))
    x = np.asfortranarray(b.copy())
    pastix.task_solve(h, S, x)
    # Local: A11 x1 = b1 on the factored unknowns, the Schur rows untouched
    x1 = np.linalg.solve(Ad[np.ix_(J, J)], b[J])
    assert np.abs(x[J] - x1).max() <= 1e-12 * np.abs(x1).max()
    assert np.array_equal(x[I], b[I])
    pastix.finalize(h)


def test_schur_fp32_factor():
    """FP32 factors of the non-Schur part: the Schur complement in the
    factor precision (about 1e-6 relative to the FP64 one)."""
    grid = (8, 7, 6)
    n = 8 * 7 * 6
    ns = 40
    A, S, h, lst = _setup(F.LLT, grid, ns, pastix.solv_mode.Interface, prec=spm.SpmFloat)
    Sd = np.zeros((ns, ns), order="F")
    pastix.getSchur(h, Sd)
    Ad = A.toarray()
    J = np.setdiff1d(np.arange(n), lst)
    Sref = Ad[np.ix_(lst, lst)] - Ad[np.ix_(lst, J)] @ np.linalg.solve(Ad[np.ix_(J, J)], Ad[np.ix_(J, lst)])
    err = np.linalg.norm(np.tril(Sd) - np.tril(Sref)) / np.linalg.norm(np.tril(Sref))
    assert err <= 1e-5, err
    pastix.finalize(h)
