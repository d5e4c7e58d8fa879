"""Edge cases through the whole path on the GPU: the smallest systems, a
purely diagonal matrix (no Schur update, no off-diagonal panel row) and a 1D
chain (one cblk per level), for LL^t and LU, against dense solves."""
import numpy as np
import pytest
import scipy.sparse as sps

from pastix_amd import pastix, spm

pytestmark = pytest.mark.gpu

LLT, LU = pastix.factotype.LLT, pastix.factotype.LU


def _run(A, facto, nrhs=3, seed=1):
    A = sps.csc_matrix(A)
    S = spm.Spm(A, symmetric=facto == LLT)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = facto
    if facto == LU:
        dp[pastix.dparm.epsilon_magn_ctrl] = np.ldexp(1.0, -52)
    h = pastix.init(ip, dp)
    pastix.task_analyze(h, S)
    pastix.task_numfact(h, S)
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, size=(A.shape[0], nrhs))
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    xd = np.linalg.solve(A.toarray(), b)
    pastix.finalize(h)
    return x, xd


@pytest.mark.parametrize("facto", [LLT, LU])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_smallest_systems(facto, n):
    rng = np.random.default_rng(n)
    M = rng.uniform(-1.0, 1.0, size=(n, n))
    A = M @ M.T + n * np.eye(n) if facto == LLT else M + (n + 1) * np.eye(n)
    x, xd = _run(A, facto)
    assert np.abs(x - xd).max() <= 1e-13 * max(1.0, np.abs(xd).max())


@pytest.mark.parametrize("facto", [LLT, LU])
def test_diagonal_matrix(facto):
    n = 300
    d = np.linspace(1.0, 4.0, n)
    x, xd = _run(sps.diags(d), facto)
    assert np.abs(x - xd).max() <= 1e-15 * np.abs(xd).max()


@pytest.mark.parametrize("facto", [LLT, LU])
def test_chain(facto):
    A = spm.laplacian7(2000, 1, 1, diag=2.5)
    if facto == LU:
        A = A + sps.diags(np.full(1999, 0.3), 1)  # unsymmetric values, same pattern up to one side
        A = A + sps.diags(np.full(1999, -0.3), -1)
    x, xd = _run(A, facto)
    assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max()
