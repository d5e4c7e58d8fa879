"""Host side of a numfact with new values (pastix_subtask_spm2bcsc, no GPU).

The first call with a pattern builds the permuted CSR through triplets; later
calls with the same index arrays take the values by one gather.  DPARM_A_NORM
is the Frobenius norm of the permuted matrix computed from those values, so
numpy's norm of the input pins the gather: a value read from the wrong input
index, a repeated entry dropped or counted twice, or a stale pattern all
change it."""
import numpy as np
import pytest
import scipy.sparse as sps

from pastix_amd import pastix, spm
from tests.helpers import RawSpm, analyse

F = pastix.factotype


def _norm(h):
    return float(h.dparm[pastix.dparm.a_norm])


@pytest.mark.parametrize("symmetric", [True, False])
def test_gather_matches_frobenius_norm(symmetric):
    A = spm.laplacian7(6, 5, 4, diag=6.0) if symmetric else spm.convection_diffusion27(5, 5, 4, seed=1)
    h, S, _ = analyse(A, symmetric, F.LLT if symmetric else F.LU, min_bs=8, max_bs=24)
    pastix.subtask_spm2bcsc(h, S)
    assert abs(_norm(h) - sps.linalg.norm(A)) < 1e-13 * sps.linalg.norm(A)
    rng = np.random.default_rng(5)
    for _ in range(3):
        C = sps.coo_matrix(A)
        f = rng.uniform(0.5, 2.0, size=C.nnz)
        if symmetric:
            key = np.minimum(C.row, C.col) * A.shape[0] + np.maximum(C.row, C.col)
            _, inv = np.unique(key, return_inverse=True)
            f = rng.uniform(0.5, 2.0, size=inv.max() + 1)[inv]
        A2 = sps.csc_matrix((C.data * f, (C.row, C.col)), shape=A.shape)
        S.set_values(A2)
        pastix.subtask_spm2bcsc(h, S)
        assert abs(_norm(h) - sps.linalg.norm(A2)) < 1e-13 * sps.linalg.norm(A2)
    pastix.finalize(h)


def test_gather_sums_repeated_ijv_entries_and_detects_new_patterns():
    A = spm.convection_diffusion27(5, 4, 4, seed=3)
    n = A.shape[0]
    h, _, _ = analyse(A, False, F.LU, min_bs=8, max_bs=24)
    C = sps.coo_matrix(A)
    rng = np.random.default_rng(9)
    rep = rng.permutation(C.nnz)[: C.nnz // 3]
    rows = np.concatenate([C.row, C.row[rep]])
    cols = np.concatenate([C.col, C.col[rep]])
    w = rng.uniform(0.1, 0.9, size=len(rep))
    vals = np.concatenate([C.data, C.data[rep] * 0])
    vals[rep] = C.data[rep] * w
    vals[C.nnz:] = C.data[rep] * (1 - w)
    S = RawSpm(spm.SpmIJV, cols, rows, vals, n)
    pastix.subtask_spm2bcsc(h, S)
    assert abs(_norm(h) - sps.linalg.norm(A)) < 1e-13 * sps.linalg.norm(A)
    # new values at the same positions: the gather
    S.values[:] = 2.0 * vals
    pastix.subtask_spm2bcsc(h, S)
    assert abs(_norm(h) - 2.0 * sps.linalg.norm(A)) < 1e-13 * sps.linalg.norm(A)
    # a different pattern (entries dropped) must not be read through the gather
    keep = (C.row == C.col) | (np.arange(C.nnz) % 4 != 0)
    A3 = sps.csc_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=A.shape)
    S3 = spm.Spm(A3)
    pastix.subtask_spm2bcsc(h, S3)
    assert abs(_norm(h) - sps.linalg.norm(A3)) < 1e-13 * sps.linalg.norm(A3)
    # and back
    pastix.subtask_spm2bcsc(h, S)
    assert abs(_norm(h) - 2.0 * sps.linalg.norm(A)) < 1e-13 * sps.linalg.norm(A)
    pastix.finalize(h)
