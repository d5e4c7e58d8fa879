"""Host analysis (ordering, symbolic factorisation, splitting) through the C-ABI.

Pinned by the hand-derived 3x3x3 grid ordering of the reference rules
(tests/golden/grid_order_3x3x3.json), by the nested-dissection invariants, and
by the fill of a dense Cholesky (the symbol must contain every nonzero of L).
"""
import json
import os

import numpy as np
import pytest

from pastix_amd import pastix, spm
from tests.helpers import RawSpm, analyse

HERE = os.path.dirname(os.path.abspath(__file__))
LLT = pastix.factotype.LLT
LU = pastix.factotype.LU


def test_grid_order_golden_3x3x3():
    g = json.load(open(os.path.join(HERE, "golden", "grid_order_3x3x3.json")))
    o = pastix.orderGrid(3, 3, 3)
    perm, peri, rang, tree = pastix.order_arrays(o)
    assert perm.tolist() == g["permtab"]
    assert rang.tolist() == g["rangtab"]
    assert tree.tolist() == g["treetab"]
    assert np.array_equal(peri[perm], np.arange(27))
    pastix.orderExit(o)


@pytest.mark.parametrize("grid", [(5, 5, 5), (20, 20, 20), (7, 3, 11), (16, 9, 1), (1, 1, 40)])
def test_grid_order_invariants(grid):
    nx, ny, nz = grid
    n = nx * ny * nz
    o = pastix.orderGrid(nx, ny, nz)
    perm, peri, rang, tree = pastix.order_arrays(o)
    cb = o.cblknbr
    assert sorted(perm.tolist()) == list(range(n))
    assert rang[0] == 0 and rang[-1] == n and np.all(np.diff(rang) > 0)
    assert tree[-1] == -1 and np.all((tree[:-1] > np.arange(cb - 1)))
    # a supernode and every node of its subtree are separated from the other
    # subtrees: no grid edge joins two sibling subtrees
    A = spm.laplacian7(nx, ny, nz, diag=6.0).tocoo()
    col2sn = np.searchsorted(rang, np.arange(n), side="right") - 1
    anc = [set() for _ in range(cb)]
    for s in range(cb):
        a, t = s, s
        while t != -1:
            anc[s].add(t)
            t = tree[t]
    for i, j in zip(A.row, A.col):
        si, sj = col2sn[perm[i]], col2sn[perm[j]]
        assert si in anc[sj] or sj in anc[si], "edge between unrelated subtrees"
    pastix.orderExit(o)


@pytest.mark.parametrize("grid,bs,generic", [((6, 6, 6), (160, 320), False), ((9, 8, 7), (8, 16), False),
                                             ((8, 8, 8), (4, 8), True)])
def test_symbol_contains_factor_fill(grid, bs, generic):
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=None if generic else grid, min_bs=bs[0], max_bs=bs[1])
    n = A.shape[0]
    p = sym["permtab"]
    B = np.zeros((n, n))
    Ac = A.tocoo()
    B[p[Ac.row], p[Ac.col]] = Ac.data
    L = np.linalg.cholesky(B)
    # symbol pattern
    pat = np.zeros((n, n), dtype=bool)
    for c in range(sym["cblknbr"]):
        fc, lc = sym["cblk_fcol"][c], sym["cblk_fcol"][c + 1]
        for b in range(sym["cblk_bloknum"][c], sym["cblk_bloknum"][c + 1]):
            pat[sym["blok_frow"][b]:sym["blok_lrow"][b] + 1, fc:lc] = True
    fill = np.abs(np.tril(L)) > 1e-14 * np.abs(L).max()
    assert not np.any(fill & ~pat), "a nonzero of L falls outside the symbol"
    # blocks are sorted, inside one facing cblk, and split widths obey max
    widths = np.diff(sym["cblk_fcol"])
    assert widths.max() <= bs[1]
    for c in range(sym["cblknbr"]):
        b0, b1 = sym["cblk_bloknum"][c], sym["cblk_bloknum"][c + 1]
        assert sym["blok_frow"][b0] == sym["cblk_fcol"][c]
        for b in range(b0 + 1, b1):
            f = sym["blok_fcblk"][b]
            assert sym["blok_frow"][b] > sym["blok_lrow"][b - 1]
            assert sym["cblk_fcol"][f] <= sym["blok_frow"][b] and sym["blok_lrow"][b] < sym["cblk_fcol"][f + 1]
    # IPARM_NNZEROS: strictly-lower entries of the symbol (reference symbol_cost.c:327-361)
    nnz_sym = int(np.tril(pat, -1).sum())
    pastix.finalize(h)


def test_nnzeros_and_thflops_reported():
    grid = (10, 10, 10)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=1000, max_bs=2000)
    n = A.shape[0]
    # strictly-lower symbol entries, before splitting (no split here)
    cnt = 0
    for c in range(sym["cblknbr"]):
        w = sym["cblk_fcol"][c + 1] - sym["cblk_fcol"][c]
        cnt += w * (w - 1) // 2
        for b in range(sym["cblk_bloknum"][c] + 1, sym["cblk_bloknum"][c + 1]):
            cnt += (sym["blok_lrow"][b] - sym["blok_frow"][b] + 1) * w
    assert h.iparm[pastix.iparm.nnzeros] == cnt
    assert h.dparm[pastix.dparm.fact_thflops] > 0
    pastix.finalize(h)


def test_split_widths_follow_reference_rule():
    # computeNbSplit (reference blend/splitsymbol.c:79-100): nseq = ceil(w/max),
    # minus one if w/nseq < min
    for w, bmin, bmax, expect in [(400, 160, 320, 2), (320, 160, 320, 1), (330, 160, 320, 2), (1000, 160, 320, 4),
                                  (500, 200, 300, 2)]:
        grid = (w, 1, 1)
        A = spm.laplacian7(*grid, diag=6.0)
        # identity ordering, one supernode spanning the whole chain
        ip, dp = pastix.initParam()
        ip[pastix.iparm.factorization] = LLT
        ip[pastix.iparm.ordering] = pastix.order.Personal
        ip[pastix.iparm.min_blocksize] = bmin
        ip[pastix.iparm.max_blocksize] = bmax
        h = pastix.init(ip, dp)
        o = pastix.pastix_order_t()
        perm = np.arange(w, dtype=np.int64)
        rang = np.array([0, w], dtype=np.int64)
        tree = np.array([-1], dtype=np.int64)
        import ctypes as ct
        o.vertnbr, o.cblknbr, o.baseval = w, 1, 0
        o.permtab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
        o.peritab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
        o.rangtab = rang.ctypes.data_as(ct.POINTER(ct.c_int64))
        o.treetab = tree.ctypes.data_as(ct.POINTER(ct.c_int64))
        S = spm.Spm(A, symmetric=True)
        pastix.subtask_order(h, S, o)
        pastix.subtask_symbfact(h)
        pastix.subtask_blend(h)
        sym = pastix.symbol(h)
        assert sym["cblknbr"] == expect, (w, bmin, bmax)
        widths = np.diff(sym["cblk_fcol"])
        assert widths.sum() == w
        pastix.finalize(h)


def test_levels_respect_dependencies():
    grid = (12, 12, 12)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=20, max_bs=40)
    # every block of cblk c faces a later cblk: the schedule levels are a topological order
    for c in range(sym["cblknbr"]):
        for b in range(sym["cblk_bloknum"][c] + 1, sym["cblk_bloknum"][c + 1]):
            assert sym["blok_fcblk"][b] > c
    assert sym["nlevels"] >= 1
    pastix.finalize(h)


@pytest.mark.parametrize("grid", [(12, 12, 12), (20, 20, 20)])
def test_update_flops_partition_the_reference_count(grid):
    # the per-kernel split of the Schur-update flops (single-source tiles,
    # two-phase multi-source tiles, single-pass multi-source tiles) covers
    # every (p, q) pair of the reference's per-block GEMMs exactly once
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=20, max_bs=40)
    uf = pastix.update_flops(h)
    fl = pastix.flops(h)
    total = uf["direct"] + uf["multi"] + uf["single"]
    assert total == pytest.approx(fl[3], rel=1e-12)
    # the U side of LU drops the pairs of a source's diagonal-facing block
    assert 0 < uf["direct_u"] + uf["multi_u"] + uf["single_u"] < total
    # product tiles cover at least the two-phase pairs they serve (one side)
    assert uf["wtile"] >= uf["multi"] * (1 - 1e-12)
    pastix.finalize(h)


def test_ldlt_thflops_follow_reference_formulas():
    # DPARM_FACT_THFLOPS for LDL^t: sum over cblks of FLOPS_DSYTRF(N) +
    # FLOPS_DTRSM(Right, M, N) + N M + FLOPS_DSYRK(N, M) + M^2
    # (symbol/symbol_cost_flops.c:85-88,142-144,199-201; common/flops.h:101-102)
    grid = (9, 8, 7)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, pastix.factotype.LDLT, grid=grid, min_bs=12, max_bs=40)
    tot = 0.0
    for c in range(sym["cblknbr"]):
        N = float(sym["cblk_fcol"][c + 1] - sym["cblk_fcol"][c])
        M = float(sum(sym["blok_lrow"][b] - sym["blok_frow"][b] + 1
                      for b in range(sym["cblk_bloknum"][c] + 1, sym["cblk_bloknum"][c + 1])))
        fm = N * ((N / 6 + 0.5) * N + 10 / 3)
        fa = N * ((N / 6) * N - 1 / 6)
        tot += fm + fa
        if M > 0:
            tot += M * N * (N + 1) + N * M + N * M * (M + 1) + M * M
    assert h.dparm[pastix.dparm.fact_thflops] == pytest.approx(tot, rel=1e-12)
    pastix.finalize(h)


@pytest.mark.parametrize("symmetric", [True, False])
def test_same_analysis_for_every_input_format(symmetric):
    """The adjacency graph is read from CSC and CSR arrays in place and from
    IJV through triplets: the nested-dissection ordering and the partition
    built on it must not depend on the format (nor on IJV entry order and
    repeats)."""
    import scipy.sparse as sps
    A = spm.laplacian7(7, 6, 5, diag=6.0) if symmetric else spm.convection_diffusion27(6, 5, 5, seed=2)
    n = A.shape[0]
    T = sps.tril(A, format="csc") if symmetric else sps.csc_matrix(A)
    T.sort_indices()
    R = sps.csr_matrix(T)
    R.sort_indices()
    C = T.tocoo()
    rng = np.random.default_rng(3)
    rep = rng.permutation(C.nnz)[: C.nnz // 4]
    order = rng.permutation(C.nnz + len(rep))
    rows = np.concatenate([C.row, C.row[rep]])[order]
    cols = np.concatenate([C.col, C.col[rep]])[order]
    vals = np.concatenate([C.data, C.data[rep]])[order]
    inputs = [RawSpm(spm.SpmCSC, T.indptr, T.indices, T.data, n, symmetric),
              RawSpm(spm.SpmCSR, R.indices, R.indptr, R.data, n, symmetric),
              RawSpm(spm.SpmIJV, cols, rows, vals, n, symmetric)]
    syms = []
    for S in inputs:
        ip, dp = pastix.initParam()
        ip[pastix.iparm.factorization] = LLT if symmetric else pastix.factotype.LU
        ip[pastix.iparm.min_blocksize] = 8
        ip[pastix.iparm.max_blocksize] = 24
        h = pastix.init(ip, dp)
        pastix.task_analyze(h, S)
        syms.append(pastix.symbol(h))
        pastix.finalize(h)
    for sym in syms[1:]:
        for k in ("permtab", "cblk_fcol", "cblk_bloknum", "blok_frow", "blok_lrow", "blok_fcblk"):
            assert np.array_equal(sym[k], syms[0][k]), k


@pytest.mark.parametrize("case", ["llt_split", "llt_nosplit", "lu_split", "ldlt"])
def test_thflops_and_nnzeros_equal_the_oracle_counts(case):
    """DPARM_FACT_THFLOPS (the metric's denominator) equals the oracle's
    reference count (symbol/symbol_cost.c:53-88 with common/flops.h, pinned
    by brute-force operation counts in test_oracle.py) on the analysed,
    split symbol, and IPARM_NNZEROS equals the strictly-lower entries of the
    symbol (symbol/symbol_cost.c:327-361), doubled for LU
    (blend/pastix_subtask_blend.c:535-539); DPARM_FILL_IN is their ratio to
    the input's stored entries."""
    import oracle
    if case.startswith("lu"):
        grid = (12, 11, 10)
        A = spm.convection_diffusion27(*grid, seed=3)
        facto, sym_in = LU, False
    elif case == "ldlt":
        grid = (9, 8, 7)
        A = spm.laplacian7(*grid, diag=6.0)
        facto, sym_in = pastix.factotype.LDLT, True
    else:
        grid = (16, 15, 14)
        A = spm.laplacian7(*grid, diag=6.0)
        facto, sym_in = LLT, True
    bs = (1000, 2000) if case == "llt_nosplit" else (12, 24)
    h, S, sym = analyse(A, sym_in, facto, grid=grid, min_bs=bs[0], max_bs=bs[1])
    O = oracle.Oracle(sym)
    th = float(h.dparm[pastix.dparm.fact_thflops])
    if case != "ldlt":
        assert th == pytest.approx(O.thflops(lu=facto == LU), rel=1e-14, abs=0)
        assert th == float(pastix.flops(h)[0])
    else:
        # SYTRF counts (symbol_cost_flops.c): below the LU count of the same symbol
        assert 0 < th < O.thflops(lu=True)
    cnt = 0
    for c in range(sym["cblknbr"]):
        w = sym["cblk_fcol"][c + 1] - sym["cblk_fcol"][c]
        cnt += w * (w - 1) // 2
        for b in range(sym["cblk_bloknum"][c] + 1, sym["cblk_bloknum"][c + 1]):
            cnt += (sym["blok_lrow"][b] - sym["blok_frow"][b] + 1) * w
    nnz = int(h.iparm[pastix.iparm.nnzeros])
    assert nnz == (2 * cnt if facto == LU else cnt)
    stored = len(S.values)
    assert float(h.dparm[pastix.dparm.fill_in]) == pytest.approx(nnz / stored, rel=1e-15)
    if case == "llt_split":
        # the split actually happened
        assert max(sym["cblk_fcol"][1:] - sym["cblk_fcol"][:-1]) <= 24
    pastix.finalize(h)
