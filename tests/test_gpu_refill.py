"""Numerical factorisation of a new matrix with the analysed pattern.

The reference rebuilds its permuted matrix from the spm on every numfact
(bcsc/bcsc_zinit.c:485-541) and refills the coefficient tables
(kernels/cpucblk_zinit.c:229-309).  Here a later input whose format, type
and index arrays equal the analysed one takes its permuted values by one
gather, and the device fills the tables.  The factor must be the same, bit
for bit, as the one of a fresh instance that reads the same matrix through
the whole triplet path; an input whose pattern differs must be detected and
take the whole path."""
import numpy as np
import pytest
import scipy.sparse as sps

from pastix_amd import pastix, spm
from tests.helpers import RawSpm, analyse

pytestmark = pytest.mark.gpu

F = pastix.factotype
GRID = (9, 8, 7)


class IJV(RawSpm):
    """An spmatrix_t in IJV format whose entries may repeat (repeats add)."""

    def __init__(self, rows, cols, vals, n, symmetric=False):
        super().__init__(spm.SpmIJV, cols, rows, vals, n, symmetric)


def _perturbed(A, seed, symmetric):
    """Same pattern, new values: off-diagonals scaled by factors in [0.5, 1]
    (symmetrically for a symmetric A), diagonal raised by 0.5."""
    rng = np.random.default_rng(seed)
    C = sps.coo_matrix(A)
    f = rng.uniform(0.5, 1.0, size=C.nnz)
    if symmetric:
        key = np.minimum(C.row, C.col) * A.shape[0] + np.maximum(C.row, C.col)
        _, inv = np.unique(key, return_inverse=True)
        f = rng.uniform(0.5, 1.0, size=inv.max() + 1)[inv]
    v = np.where(C.row == C.col, C.data + 0.5, C.data * f)
    return sps.csc_matrix((v, (C.row, C.col)), shape=A.shape)


def _matrix(facto):
    if facto == F.LU:
        return spm.convection_diffusion27(*GRID, seed=5), False
    return spm.laplacian7(*GRID, diag=6.0), True


def _fresh(A, symmetric, facto, prec=None, S=None):
    h, S0, _ = analyse(A, symmetric, facto, grid=GRID, min_bs=12, max_bs=40)
    if prec is not None:
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    pastix.task_numfact(h, S0 if S is None else S)
    return h


def _tables(h, facto, prec):
    dt = np.float32 if prec == spm.SpmFloat else np.float64
    return pastix.coeftab(h, dtype=dt, lu=facto != F.LLT)


@pytest.mark.parametrize("facto,prec", [(F.LLT, None), (F.LLT, spm.SpmFloat), (F.LU, None), (F.LU, spm.SpmFloat),
                                        (F.LDLT, None)])
def test_refill_matches_fresh_instance(facto, prec):
    A1, symmetric = _matrix(facto)
    A2 = _perturbed(A1, 11, symmetric)
    h, S, _ = analyse(A1, symmetric, facto, grid=GRID, min_bs=12, max_bs=40)
    if prec is not None:
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    pastix.task_numfact(h, S)
    L1, _ = _tables(h, facto, prec)
    S.set_values(A2)
    pastix.task_numfact(h, S)  # same pattern: the gather
    L, U = _tables(h, facto, prec)
    h2 = _fresh(A2, symmetric, facto, prec)
    L2, U2 = _tables(h2, facto, prec)
    assert not np.array_equal(L, L1), "the new values did not reach the factor"
    assert np.array_equal(L, L2)
    if U is not None:
        assert np.array_equal(U, U2)
    assert h.dparm[pastix.dparm.a_norm] == h2.dparm[pastix.dparm.a_norm]
    # and the factor solves the new system
    x0, b = spm.rhs_from_x(A2, 2, seed=4)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    tol = 1e-4 if prec == spm.SpmFloat else 1e-10
    assert np.abs(x - x0).max() / np.abs(x0).max() < tol
    pastix.finalize(h)
    pastix.finalize(h2)


@pytest.mark.parametrize("symmetric", [False, True])
def test_refill_ijv_with_repeated_entries(symmetric):
    """IJV input whose entries repeat (summed in input order) and come in
    shuffled order: the gather reproduces the triplet path's sums."""
    facto = F.LLT if symmetric else F.LU
    A1, _ = _matrix(facto)
    rng = np.random.default_rng(7)

    def ijv(A):
        A = sps.csc_matrix(A)
        A.sort_indices()
        C = (sps.tril(A, format="csc") if symmetric else A).tocoo()
        r, c, v = C.row, C.col, C.data
        # every third entry split into two parts
        k = np.arange(len(v)) % 3 == 0
        w = rng.uniform(0.2, 0.8, size=k.sum())
        vv = np.concatenate([v, v[k] * (1 - w)])
        vv[np.flatnonzero(k)] = v[k] * w
        return np.concatenate([r, r[k]]), np.concatenate([c, c[k]]), vv

    r, c, v = ijv(A1)
    order = rng.permutation(len(v))
    r, c = r[order], c[order]
    n = A1.shape[0]
    S = IJV(r, c, v[order], n, symmetric)
    h, _, _ = analyse(A1, symmetric, facto, grid=GRID, min_bs=12, max_bs=40)
    pastix.task_numfact(h, S)
    A2 = _perturbed(A1, 12, symmetric)
    # same positions in the same order, new values (repeats split afresh)
    r2, c2, v2 = ijv(A2)
    assert np.array_equal(r2[order], r) and np.array_equal(c2[order], c)
    S.values[:] = v2[order]
    pastix.task_numfact(h, S)
    L, U = _tables(h, facto, None)
    S2 = IJV(r, c, v2[order], n, symmetric)
    h2 = _fresh(A1, symmetric, facto, S=S2)
    L2, U2 = _tables(h2, facto, None)
    assert np.array_equal(L, L2)
    if U is not None:
        assert np.array_equal(U, U2)
    x0, b = spm.rhs_from_x(A2, 1, seed=5)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    # the summed entries differ from A2's by rounding only
    assert np.abs(x - x0).max() / np.abs(x0).max() < 1e-9
    pastix.finalize(h)
    pastix.finalize(h2)


def test_changed_pattern_takes_whole_path():
    """An input with fewer entries than the analysed one (inside the same
    symbol) must not be read through the analysed pattern's gather."""
    A1, _ = _matrix(F.LU)
    h, S, _ = analyse(A1, False, F.LU, grid=GRID, min_bs=12, max_bs=40)
    pastix.task_numfact(h, S)
    C = sps.coo_matrix(A1)
    keep = (C.row == C.col) | (np.arange(C.nnz) % 5 != 0)
    A3 = sps.csc_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=A1.shape)
    S3 = spm.Spm(A3)
    pastix.task_numfact(h, S3)
    L, U = _tables(h, F.LU, None)
    h2 = _fresh(A1, False, F.LU, S=spm.Spm(A3))
    L2, U2 = _tables(h2, F.LU, None)
    assert np.array_equal(L, L2) and np.array_equal(U, U2)
    # back to the analysed pattern
    pastix.task_numfact(h, S)
    L4, _ = _tables(h, F.LU, None)
    h3 = _fresh(A1, False, F.LU)
    L5, _ = _tables(h3, F.LU, None)
    assert np.array_equal(L4, L5)
    for hh in (h, h2, h3):
        pastix.finalize(hh)


def test_alternating_patterns_keep_device_memory_constant():
    """An instance whose numfact inputs alternate between two patterns (the
    second drops some off-diagonal pairs, so the analysed symbol still holds
    it) frees the first pattern's device buffers before uploading the second:
    the device footprint stays constant, and each factor equals a fresh
    instance's."""
    A, sym = _matrix(F.LLT)
    C = sps.coo_matrix(A)
    keep = ~((C.row != C.col) & ((np.minimum(C.row, C.col) % 5) == 0) & (np.abs(C.row - C.col) == 1))
    B = sps.csc_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=A.shape)
    assert B.nnz < A.nnz
    h, SA, _ = analyse(A, True, F.LLT, grid=GRID, min_bs=12, max_bs=40)
    SB = spm.Spm(B, symmetric=True)
    sizes = []
    for S in (SA, SB, SA, SB, SA, SB):
        pastix.task_numfact(h, S)
        sizes.append(pastix.device_bytes(h))
    assert sizes[2:] == [sizes[0], sizes[1]] * 2, sizes
    LB, _ = pastix.coeftab(h)
    hB = _fresh(A, True, F.LLT, S=SB)
    LB0, _ = pastix.coeftab(hB)
    assert np.array_equal(LB, LB0)
    pastix.finalize(hB)
    pastix.finalize(h)
