"""Partial solves (pastix_subtask_trsm / pastix_subtask_diag) chained as the
reference's solve chains them (sopalin/pastix_task_solve.c:433-516), on b in
the permuted ordering.  For LL^t and LU the chain runs the same kernels as the
whole solve and must reproduce it bit for bit; the forward sweep alone is
checked against the dense factor (L z = b)."""
import numpy as np
import pytest
import scipy.sparse as sps

from pastix_amd import pastix, spm
from tests.helpers import analyse
import oracle

pytestmark = pytest.mark.gpu

F = pastix.factotype
L_, U_ = pastix.uplo.Lower, pastix.uplo.Upper
N_, T_ = pastix.trans.NoTrans, pastix.trans.Trans
NU, UN = pastix.diag.NonUnit, pastix.diag.Unit
LEFT = pastix.side.Left


def _case(facto):
    grid = (9, 8, 7)
    if facto == F.LU:
        A = spm.convection_diffusion27(*grid, seed=3)
    elif facto == F.LDLT:
        A = (spm.laplacian7(*grid, diag=6.0) - 2.5 * sps.identity(np.prod(grid))).tocsr()
    else:
        A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, facto != F.LU, facto, grid=grid, min_bs=12, max_bs=40)
    pastix.task_numfact(h, S)
    return A, h, S, sym


@pytest.mark.parametrize("facto", [F.LLT, F.LU, F.LDLT])
def test_subtask_chain_matches_solve(facto):
    A, h, S, sym = _case(facto)
    x0, b = spm.rhs_from_x(A, 3, seed=8)
    whole = b.copy(order="F")
    pastix.subtask_applyorder(h, pastix.dir.Forward, whole)
    pastix.subtask_solve(h, whole)
    x = b.copy(order="F")
    pastix.subtask_applyorder(h, pastix.dir.Forward, x)
    bp = x.copy(order="F")
    if facto == F.LLT:
        pastix.subtask_trsm(h, LEFT, L_, N_, NU, x)
        z = x.copy(order="F")
        pastix.subtask_trsm(h, LEFT, L_, T_, NU, x)
    elif facto == F.LU:
        pastix.subtask_trsm(h, LEFT, L_, N_, UN, x)
        z = x.copy(order="F")
        pastix.subtask_trsm(h, LEFT, U_, N_, NU, x)
    else:
        pastix.subtask_trsm(h, LEFT, L_, N_, UN, x)
        z = x.copy(order="F")
        pastix.subtask_diag(h, x)
        pastix.subtask_trsm(h, LEFT, L_, T_, UN, x)
    if facto == F.LDLT:
        # the chain divides by D and multiplies back before the U^{-1} sweep
        # (L^{-t} = U^{-1} D): rounding, amplified by the indefinite matrix's
        # conditioning (measured 3e-13)
        assert np.abs(x - whole).max() <= 1e-11 * np.abs(whole).max()
    else:
        assert np.array_equal(x, whole)
    # the forward sweep: L z = b with the factor's L (unit for LU and LDL^t)
    Lg, _ = pastix.coeftab(h, lu=facto != F.LLT)
    O = oracle.Oracle(sym)
    Dn = O.to_dense(Lg, lower_only=False)
    Ld = np.tril(Dn) if facto == F.LLT else np.tril(Dn, -1) + np.eye(A.shape[0])
    # residual of a triangular solve, scaled by |L| |z| (L of the indefinite
    # LDL^t carries element growth)
    assert np.abs(Ld @ z - bp).max() <= 1e-13 * (np.abs(Ld) @ np.abs(z)).max()
    pastix.subtask_applyorder(h, pastix.dir.Backward, x)
    assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max()
    pastix.finalize(h)


def test_subtask_diag_divides_by_d():
    A, h, S, sym = _case(F.LDLT)
    Lg, _ = pastix.coeftab(h, lu=True)
    d = np.diag(oracle.Oracle(sym).to_dense(Lg, lower_only=False)).copy()
    v = np.random.default_rng(2).uniform(-1, 1, (A.shape[0], 2))
    w = v.copy(order="F")
    pastix.subtask_diag(h, w)
    assert np.allclose(w, v / d[:, None], rtol=1e-15, atol=0)
    pastix.finalize(h)


def test_subtask_trsm_rejects_other_sweeps():
    A, h, S, sym = _case(F.LLT)
    x = np.zeros((A.shape[0], 1), order="F")
    with pytest.raises(Exception):
        pastix.subtask_trsm(h, LEFT, U_, N_, NU, x)
    with pytest.raises(Exception):
        pastix.subtask_trsm(h, pastix.side.Right, L_, N_, NU, x)
    pastix.finalize(h)


@pytest.mark.parametrize("facto", [F.LLT, F.LU, F.LDLT])
def test_get_diag_matches_factor(facto):
    """pastixGetDiag returns the diagonal of the L panels (reference
    sopalin/coeftab_z.c:462-492) in the permuted ordering, at stride incD."""
    A, h, S, sym = _case(facto)
    L, _ = pastix.coeftab(h, lu=facto != F.LLT)
    want = np.empty(sym["n"])
    for c in range(sym["cblknbr"]):
        f0, f1 = sym["cblk_fcol"][c], sym["cblk_fcol"][c + 1]
        j = np.arange(f1 - f0)
        want[f0:f1] = L[sym["cblk_coef"][c] + j * sym["cblk_stride"][c] + j]
    D = pastix.getDiag(h, sym["n"], incD=3)
    assert np.array_equal(D[0::3], want)
    assert not D[1::3].any() and not D[2::3].any()
    # the diagonal fixes the determinant: det A = prod(L_ii)^2 for LL^t,
    # prod(U_ii) for LU, prod(D_i) for LDL^t
    sign, logdet = np.linalg.slogdet(A.toarray())
    ld = np.log(np.abs(want)).sum() * (2 if facto == F.LLT else 1)
    assert abs(ld - logdet) < 1e-9 * abs(logdet)
    if facto != F.LLT:
        assert np.prod(np.sign(want)) == sign
    pastix.finalize(h)
