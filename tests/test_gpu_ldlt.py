"""LDL^t (PastixFactLDLT / SYTRF) on the device against the oracle and the
reference's own LDL^t and refinement tests.

The engine runs LDL^t as LU without pivoting of the symmetric matrix: L is
the same unit factor, D = diag(U) and U = D L^t, so the L panel holds the
reference's layout (kernels/core_zsytrfsp.c:162-168).  Pinned by the LU
oracle on the same matrix (itself pinned by dense Doolittle elimination) and
by the LDL^t identities U = D L^t on every off-diagonal tile.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sps

from pastix_amd import pastix, spm
from tests.helpers import analyse, oracle_factor, rel_frob, read_rua

pytestmark = pytest.mark.gpu

LDLT = pastix.factotype.LDLT
HERE = os.path.dirname(os.path.abspath(__file__))


def _check_axb(A, x, b):
    # spmCheckAxb (spm/src/z_spm_genrhs.c:415-425): ||b-Ax||_1 / (||A||_1 ||x||_1 eps)
    r = b - A @ x
    return np.abs(r).sum() / (abs(A).sum(axis=0).max() * np.abs(x).sum() * np.finfo(float).eps)


def _diag_and_offdiag(sym, L, U):
    """D from the diagonal blocks of L, and (L_off, U_off) of every cblk."""
    out = []
    for c in range(sym["cblknbr"]):
        fc, lc, st = sym["cblk_fcol"][c], sym["cblk_fcol"][c + 1], sym["cblk_stride"][c]
        K = lc - fc
        P = L[sym["cblk_coef"][c]:sym["cblk_coef"][c + 1]].reshape((K, st)).T
        Q = U[sym["cblk_coef"][c]:sym["cblk_coef"][c + 1]].reshape((K, st)).T
        out.append((np.diag(P[:K, :K]).copy(), P[K:, :].copy(), Q[K:, :].copy(), P[:K, :K].copy()))
    return out


def test_ldlt_indefinite_parity():
    grid = (10, 9, 8)
    # eigenvalues of the 7-point operator lie in (0, 12): shifted by 6.5 it is indefinite
    A = (spm.laplacian7(*grid, diag=6.0) - 6.5 * sps.identity(np.prod(grid))).tocsr()
    h, S, sym = analyse(A, True, LDLT, grid=grid, min_bs=16, max_bs=48)
    pastix.task_numfact(h, S)
    Lg, Ug = pastix.coeftab(h, lu=True)
    crit = np.sqrt(np.ldexp(1.0, -53)) * np.linalg.norm(A.toarray())
    O, bc, Lo, Uo = oracle_factor(A, sym, lu=True, crit=crit)
    err = rel_frob(Lg, Lo)
    # Without pivoting an indefinite LDL^t grows its entries (growth is
    # printed: about 3.5e3 with this ordering), and rounding differences grow
    # with them.  The bars:
    # - the backward error of the computed factorisation, ||L U - P A P^T||_F
    #   / ||A||_F (L unit lower from the L panels, U from the diagonal squares
    #   and the transposed U panels), meets the textbook bound of LU without
    #   pivoting, gamma_n || |L| |U| ||_F / ||A||_F (Higham, Accuracy and
    #   Stability, Thm 9.3, gamma_n = n eps / (1 - n eps)), for the oracle and
    #   for the engine (whose panel TRSMs multiply by explicit inverses of the
    #   diagonal blocks, so its error is larger but still within the bound);
    # - the factors agree to 1e-7 relative Frobenius and the two sides of U =
    #   D L^t to 1e-7 of their largest entry (rounding times growth; a wrong
    #   term or sign gives O(1));
    # the refined solve meets the reference's spmCheckAxb below.
    B = bc["B"].toarray()
    nn = B.shape[0]
    gamma = nn * np.finfo(float).eps / (1 - nn * np.finfo(float).eps)
    back, bounds = [], []
    for Lp, Up in ((Lg, Ug), (Lo, Uo)):
        Ld = np.tril(O.to_dense(Lp, lower_only=True), -1) + np.eye(nn)
        Ud = np.triu(O.to_dense(Up, lower_only=True).T)
        back.append(np.linalg.norm(Ld @ Ud - B) / np.linalg.norm(B))
        bounds.append(gamma * np.linalg.norm(np.abs(Ld) @ np.abs(Ud)) / np.linalg.norm(B))
    growth = np.abs(Lg).max() / np.abs(A.data).max()
    print("LDL^t rel frob vs LU oracle", err, "backward errors", back, "textbook bounds", bounds, "growth", growth,
          "static pivots", h.iparm[pastix.iparm.static_pivoting])
    assert back[1] <= bounds[1] and back[0] <= bounds[0]
    assert err <= 1e-7
    negs = 0
    for d, Loff, Uoff, sq in _diag_and_offdiag(sym, Lg, Ug):
        negs += int((d < 0).sum())
        # U = D L^t: the transposed U tiles are L D (column j of L scaled by
        # d_j); the two sides come from different products
        if Loff.size:
            assert np.abs(Uoff - Loff * d[None, :]).max() <= 1e-7 * max(1.0, np.abs(Uoff).max())
        # the diagonal block's upper triangle holds D L^t
        K = len(d)
        Lsq = np.tril(sq, -1) + np.eye(K)
        assert np.abs(np.triu(sq, 1) - np.triu((Lsq * d[None, :]).T, 1)).max() <= 1e-7 * max(1.0, np.abs(sq).max())
    assert negs > 0, "the shifted operator must give negative pivots"
    x0, b = spm.rhs_from_x(A, 1, seed=11)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    # ||r|| / ||b|| <= 1e-12 (the default) does not bound the normwise
    # spmCheckAxb criterion on this indefinite matrix (growth ~3.5e3): refine
    # further
    h.dparm[pastix.dparm.epsilon_refinement] = 1e-14
    pastix.task_refine(h, S, b, x)
    assert _check_axb(A, x[:, 0], b[:, 0]) <= 1e2
    pastix.finalize(h)


@pytest.mark.parametrize("method", ["GMRES", "CG"])
def test_ldlt_small_rsa_reference_check(method):
    """The reference's LDL^t test matrix (test/matrix/small.rsa, `-f 1`, static
    pivoting threshold 1e-14 ||A||, example/CTestLists.txt:106,121) with its
    pass criterion spmCheckAxb <= 1e2; the factor matches the LU oracle's on
    the same matrix and threshold."""
    A = read_rua(os.path.join(HERE, "golden", "small.rsa"))
    h, S, sym = analyse(A, True, LDLT)
    h.iparm[pastix.iparm.refinement] = getattr(pastix.refine, method)
    h.dparm[pastix.dparm.epsilon_magn_ctrl] = 1e-14
    pastix.task_numfact(h, S)
    Lg, _ = pastix.coeftab(h, lu=True)
    crit = 1e-14 * np.linalg.norm(A.data)
    O, bc, Lo, Uo = oracle_factor(A, sym, lu=True, crit=crit)
    err = rel_frob(Lg, Lo)
    print("small.rsa", method, "rel frob", err, "static pivots", h.iparm[pastix.iparm.static_pivoting])
    assert err <= 1e-12
    x0, b = spm.rhs_from_x(A, 1, seed=42)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    pastix.task_refine(h, S, b, x)
    crit_axb = _check_axb(A, x[:, 0], b[:, 0])
    print("spmCheckAxb", crit_axb, "iters", h.iparm[pastix.iparm.nbiter])
    assert crit_axb <= 1e2
    pastix.finalize(h)


def test_orsirr_lu_bicgstab_reference_check():
    """The reference's BiCGSTAB refinement test (orsirr.rua, LU,
    example/CTestLists.txt:123) with spmCheckAxb <= 1e2."""
    A = read_rua(os.path.join(HERE, "golden", "orsirr.rua"))
    h, S, sym = analyse(A, False, pastix.factotype.LU)
    h.iparm[pastix.iparm.refinement] = pastix.refine.BiCGSTAB
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, 1, seed=24356)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    pastix.task_refine(h, S, b, x)
    assert _check_axb(A, x[:, 0], b[:, 0]) <= 1e2
    pastix.finalize(h)


def test_ldlt_fp32_factor_fp64_refinement():
    grid = (12, 11, 10)
    A = (spm.laplacian7(*grid, diag=6.0) - 2.5 * sps.identity(np.prod(grid))).tocsr()
    h, S, sym = analyse(A, True, LDLT, grid=grid)
    pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, spm.SpmFloat)
    h.iparm[pastix.iparm.refinement] = pastix.refine.GMRES
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, 1, seed=5)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    pastix.task_refine(h, S, b, x)
    r = b[:, 0] - A @ x[:, 0]
    berr = np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    print("LDL^t FP32 + GMRES berr", berr, "iters", h.iparm[pastix.iparm.nbiter])
    assert berr <= 1e-12
    pastix.finalize(h)
