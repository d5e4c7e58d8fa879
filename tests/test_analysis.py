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
    # (the smallest-split rule moves a cut by up to 10% of the step,
    # reference blend/splitsymbol.c:195-241)
    widths = np.diff(sym["cblk_fcol"])
    assert widths.max() <= bs[1] + -(-bs[1] // 10)
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
    # (and, with the leaf region's persistent launch on, its in-launch
    # updates and the level kernels' share together)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=20, max_bs=40)
    fl = pastix.flops(h)
    pastix.set_amd_iparm(h, pastix.amd_iparm.leaf_tasks, 0)
    off = pastix.update_flops(h)
    assert off["leaf"] == 0
    assert off["direct"] + off["multi"] + off["single"] == pytest.approx(fl[3], rel=1e-12)
    # the accounting of the (off by default) leaf-region launch
    pastix.set_amd_iparm(h, pastix.amd_iparm.leaf_tasks, 1)
    uf = pastix.update_flops(h)
    assert uf["leaf"] > 0
    total = uf["direct"] + uf["multi"] + uf["single"] + uf["leaf"]
    assert total == pytest.approx(fl[3], rel=1e-12)
    # the level kernels' split (leaf launch off) for the U side of LU
    uf = off
    total = uf["direct"] + uf["multi"] + uf["single"]
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
        assert max(sym["cblk_fcol"][1:] - sym["cblk_fcol"][:-1]) <= 24 + 3
    pastix.finalize(h)


@pytest.mark.parametrize("base", [0, 1])
def test_schur_unknowns_are_ordered_last(base):
    """Reference order/pastix_subtask_order.c:198-230 with graphIsolate and
    pastixOrderAddIsolate: the listed unknowns (in the matrix's numbering
    base) take the last positions, in increasing index order, as the last
    supernode(s); the rest is ordered as a graph of its own."""
    grid = (7, 6, 5)
    A = spm.laplacian7(*grid, diag=6.0)
    n = A.shape[0]
    S = spm.Spm(A, symmetric=True)
    if base:
        S.colptr += 1
        S.rowptr += 1
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = LLT
    ip[pastix.iparm.min_blocksize] = 8
    ip[pastix.iparm.max_blocksize] = 16
    h = pastix.init(ip, dp)
    lst = np.array([100, 3, 57, 4, 200, 199, 198, 12], dtype=np.int64)
    pastix.setSchurUnknownList(h, lst + base)
    pastix.task_analyze(h, S)
    sym = pastix.symbol(h)
    p = sym["permtab"]
    ns = len(lst)
    assert list(p[np.sort(lst)]) == list(range(n - ns, n))
    others = np.setdiff1d(np.arange(n), lst)
    assert sorted(p[others]) == list(range(n - ns))
    # the Schur columns start a cblk: no cblk straddles n - ns
    assert n - ns in set(sym["cblk_fcol"].tolist())
    pastix.finalize(h)
    # a repeated or out-of-range unknown is refused
    h = pastix.init(ip, dp)
    pastix.setSchurUnknownList(h, np.array([1, 1]) + base)
    with pytest.raises(pastix.PastixError):
        pastix.task_analyze(h, S)
    pastix.finalize(h)


def _apply_level_order(fath, sizes, level, width, base=0):
    import ctypes as ct
    n = int(sum(sizes))
    cb = len(sizes)
    perm = np.arange(n, dtype=np.int64) + base
    peri = np.arange(n, dtype=np.int64) + base
    rang = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64) + base
    tree = np.array([f if f < 0 else f + base for f in fath], dtype=np.int64)
    o = pastix.pastix_order_t()
    o.baseval, o.vertnbr, o.cblknbr = base, n, cb
    o.permtab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.peritab = peri.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.rangtab = rang.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.treetab = tree.ctypes.data_as(ct.POINTER(ct.c_int64))
    pastix.lib.pastixOrderApplyLevelOrder.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int64]
    pastix.lib.pastixOrderApplyLevelOrder.restype = ct.c_int
    assert pastix.lib.pastixOrderApplyLevelOrder(ct.byref(o), level, width) == 0
    return perm - base, peri - base, rang - base, np.where(tree < 0, -1, tree - base)


@pytest.mark.parametrize("base", [0, 1])
def test_level_order_golden(base):
    """Reference order/order_apply_level_order.c on a 7-supernode tree,
    derived by hand from its rules: root R (6) with sons B (1) and A (5); B
    has the leaf B1 (0); A has A1 (3) and A2 (4); A1 has the 3-column leaf
    A11 (2).  Plain walk (level_tasks2d >= 0): breadth-first from the root,
    sons pushed in reverse, numbers given backwards: A11, B1, A1, A2, B, A, R.
    Default walk (level_tasks2d < 0, width 3): A11 is wide, so its family,
    A1's, A's and R's are "2D" and walk ahead of the "1D" B1: B1, A11, A1,
    A2, B, A, R."""
    fath = [1, 6, 3, 5, 5, 6, -1]
    sizes = [1, 1, 3, 1, 1, 1, 1]
    perm, peri, rang, tree = _apply_level_order(fath, sizes, 0, 3, base)
    assert list(peri) == [2, 3, 4, 0, 5, 6, 1, 7, 8]
    assert list(rang) == [0, 3, 4, 5, 6, 7, 8, 9]
    assert list(tree) == [2, 4, 5, 5, 6, 6, -1]
    assert list(perm[peri]) == list(range(9))
    perm, peri, rang, tree = _apply_level_order(fath, sizes, -1, 3, base)
    assert list(peri) == [0, 2, 3, 4, 5, 6, 1, 7, 8]
    assert list(rang) == [0, 1, 4, 5, 6, 7, 8, 9]
    assert list(tree) == [4, 2, 5, 5, 6, 6, -1]


@pytest.mark.parametrize("rule", ["default", "CONSTANT"])
def test_split_smallest_golden(rule, monkeypatch):
    """Reference blend/splitsymbol.c (computeNbSplit, splitSmart with
    computeSmallestSplit, 10% window), derived by hand: seven one-column
    leaves face a 40-column root supernode with the row ranges [0,39] x3,
    [0,15], [16,39], [16,29], [30,39]; the root's diagonal block spans every
    row.  Blocks 8/16: three pieces, the first cut searched around row 13
    (rows 11-15; row 15 ends a range, so fewer blocks span 15|16) and the
    second around row 16+11 (rows 25-29; 29|30 is spanned least): pieces
    [0,15], [16,29], [30,39].  The constant rule cuts [0,13], [14,26],
    [27,39]."""
    import ctypes as ct
    import scipy.sparse as sps
    if rule != "default":
        monkeypatch.setenv("PASTIX_BLEND_SPLIT", rule)
    m, w = 7, 40
    n = m + w
    ranges = [(0, 39), (0, 39), (0, 39), (0, 15), (16, 39), (16, 29), (30, 39)]
    rows, cols = [], []
    for i, (a, b) in enumerate(ranges):
        for r in range(a, b + 1):
            rows += [m + r, i]
            cols += [i, m + r]
    for r in range(w - 1):
        rows += [m + r + 1, m + r]
        cols += [m + r, m + r + 1]
    A = sps.csc_matrix((-np.ones(len(rows)), (rows, cols)), shape=(n, n)) + 100.0 * sps.eye(n)
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = LLT
    ip[pastix.iparm.ordering] = pastix.order.Personal
    ip[pastix.iparm.min_blocksize] = 8
    ip[pastix.iparm.max_blocksize] = 16
    h = pastix.init(ip, dp)
    pastix.set_amd_iparm(h, pastix.amd_iparm.level_order, 0)
    perm = np.arange(n, dtype=np.int64)
    rang = np.array(list(range(m + 1)) + [n], dtype=np.int64)
    tree = np.array([m] * m + [-1], dtype=np.int64)
    o = pastix.pastix_order_t()
    o.baseval, o.vertnbr, o.cblknbr = 0, n, m + 1
    o.permtab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.peritab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.rangtab = rang.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.treetab = tree.ctypes.data_as(ct.POINTER(ct.c_int64))
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    # (no symbol reordering: it would permute the root's rows first)
    pastix.subtask_blend(h)
    sym = pastix.symbol(h)
    cuts = [int(c) - m for c in sym["cblk_fcol"] if c >= m]
    expect = [0, 16, 30, 40] if rule == "default" else [0, 14, 27, 40]
    assert cuts == expect
    pastix.finalize(h)


def test_symbol_reordering_golden():
    """Reference symbol/symbol_reordering.c (hamming_distance,
    symbol_reorder_tsp), traced by hand: leaves L0, L1 face the 4-row root
    supernode on rows {0, 2} and {1, 3}; every row's set of contributors is
    {L0} or {L1}.  Inserting rows 1, 2, 3 into the tour (ghost, 0): 1 goes
    before 0 (first position, cost 2), 2 next to its twin 0 (distance 0,
    search stops), 3 at the first position (next to its twin 1); the tour
    from after the ghost is rows 3, 1, 2, 0.  The leaves then face the root
    through one block each instead of two."""
    import ctypes as ct
    import scipy.sparse as sps
    n = 6
    rows = [2, 4, 3, 5]
    cols = [0, 0, 1, 1]
    A = sps.csc_matrix((-np.ones(4), (rows, cols)), shape=(n, n))
    A = A + A.T + 10.0 * sps.eye(n)
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = LLT
    ip[pastix.iparm.ordering] = pastix.order.Personal
    h = pastix.init(ip, dp)
    pastix.set_amd_iparm(h, pastix.amd_iparm.level_order, 0)
    perm = np.arange(n, dtype=np.int64)
    rang = np.array([0, 1, 2, 6], dtype=np.int64)
    tree = np.array([2, 2, -1], dtype=np.int64)
    o = pastix.pastix_order_t()
    o.baseval, o.vertnbr, o.cblknbr = 0, n, 3
    o.permtab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.peritab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.rangtab = rang.ctypes.data_as(ct.POINTER(ct.c_int64))
    o.treetab = tree.ctypes.data_as(ct.POINTER(ct.c_int64))
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    sym = pastix.symbol(h)
    assert list(sym["peritab"]) == [0, 1, 5, 3, 4, 2]
    assert sym["bloknbr"] == 3 + 2
    pastix.finalize(h)
    # without the reordering: two blocks per leaf
    h = pastix.init(ip, dp)
    pastix.set_amd_iparm(h, pastix.amd_iparm.level_order, 0)
    o.permtab = perm.ctypes.data_as(ct.POINTER(ct.c_int64))
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_blend(h)
    assert pastix.symbol(h)["bloknbr"] == 3 + 4
    pastix.finalize(h)


def test_ordergen_symbgen_save_and_load(tmp_path, monkeypatch):
    """IPARM_IO_STRATEGY (reference order/order_io.c, symbol/symbol_io.c):
    PastixIOSave writes ordergen and symbgen (the reference's text layouts)
    into a pastix-XXXXXX directory; a second analysis with PastixIOLoad and
    the Personal ordering reads them back (PASTIX_FILE_ORDER for the ordering,
    symbname for the symbol) and gets the same permutation and blocks."""
    monkeypatch.chdir(tmp_path)
    grid = (7, 6, 5)
    A = spm.laplacian7(*grid, diag=6.0)
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = LLT
    ip[pastix.iparm.min_blocksize] = 8
    ip[pastix.iparm.max_blocksize] = 16
    ip[pastix.iparm.io_strategy] = 2  # PastixIOSave
    h = pastix.init(ip, dp)
    pastix.task_analyze(h, S)
    sym1 = pastix.symbol(h)
    pastix.finalize(h)
    dirs = [p for p in tmp_path.iterdir() if p.name.startswith("pastix-")]
    assert len(dirs) == 1
    og = (dirs[0] / "ordergen").read_text().split()
    sg = (dirs[0] / "symbgen").read_text().split()
    n = A.shape[0]
    vals = [int(v) for v in og]
    assert vals[0] == 1 and vals[2] == n
    cb = vals[1]
    perm = vals[3 + cb + 1:3 + cb + 1 + n]
    assert perm == list(sym1["permtab"])
    hdr = [int(v) for v in sg[:5]]
    assert hdr[0] == 1 and hdr[3] == n and hdr[4] == 0
    # load both
    (tmp_path / "symbname").write_text((dirs[0] / "symbgen").read_text())
    monkeypatch.setenv("PASTIX_FILE_ORDER", str(dirs[0] / "ordergen"))
    ip[pastix.iparm.io_strategy] = 1  # PastixIOLoad
    ip[pastix.iparm.ordering] = pastix.order.Personal
    h = pastix.init(ip, dp)
    pastix.task_analyze(h, S)
    sym2 = pastix.symbol(h)
    assert np.array_equal(sym2["permtab"], sym1["permtab"])
    assert sym2["cblknbr"] == sym1["cblknbr"] and sym2["bloknbr"] == sym1["bloknbr"]
    for k in ("cblk_fcol", "blok_frow", "blok_lrow", "blok_fcblk"):
        assert np.array_equal(sym2[k], sym1[k])
    pastix.finalize(h)


def test_ordergen_golden_file(tmp_path, monkeypatch):
    """A hand-written ordergen file in base 1 (version 1: rangtab, permtab,
    treetab), as the reference's ordering_load reads it."""
    monkeypatch.chdir(tmp_path)
    (tmp_path / "ordername").write_text("1\n2\t4\n1\t3\t5\n2\t4\t1\t3\n2\t-1\n")
    import ctypes as ct
    o = pastix.pastix_order_t()
    pastix.lib.pastixOrderLoad.argtypes = [ct.c_void_p, ct.c_void_p]
    pastix.lib.pastixOrderLoad.restype = ct.c_int
    assert pastix.lib.pastixOrderLoad(None, ct.byref(o)) == 0
    assert (o.baseval, o.vertnbr, o.cblknbr) == (1, 4, 2)
    assert [o.rangtab[i] for i in range(3)] == [1, 3, 5]
    assert [o.permtab[i] for i in range(4)] == [2, 4, 1, 3]
    assert [o.peritab[i] for i in range(4)] == [3, 1, 4, 2]
    assert [o.treetab[i] for i in range(2)] == [2, -1]
    pastix.orderExit(o)


def _solve_nodes_from_symbol(sym, nmax):
    """Solve nodes restated from the symbol: runs of consecutive cblks, each
    the parent of the one before (its first off-diagonal block faces the
    next), every member's off-diagonal rows starting with exactly the columns
    of the members after it, at most nmax columns; then the solve levels over
    the nodes and the merge rounds of their inverses (64-column leaves,
    adjacent pairs merged per round), summed over the batches they are
    issued in."""
    fc, bn = sym["cblk_fcol"], sym["cblk_bloknum"]
    fr, lr, fb = sym["blok_frow"], sym["blok_lrow"], sym["blok_fcblk"]
    cb = sym["cblknbr"]
    w = np.diff(fc)

    def offrows(c):  # off-diagonal rows of c in block order
        return np.concatenate([np.arange(fr[b], lr[b] + 1) for b in range(bn[c] + 1, bn[c + 1])] or [[]])
    rows = [offrows(c) for c in range(cb)]
    nodes, c = [], 0
    while c < cb:
        e, K = c + 1, w[c]
        while nmax > 1 and e < cb:
            if bn[e] == bn[e - 1] + 1 or fb[bn[e - 1] + 1] != e or K + w[e] > nmax:
                break
            ok = all(len(rows[k]) >= fc[e + 1] - fc[k + 1] and
                     np.array_equal(rows[k][:fc[e + 1] - fc[k + 1]], np.arange(fc[k + 1], fc[e + 1]))
                     for k in range(c, e))
            if not ok:
                break
            K += w[e]
            e += 1
        nodes.append((c, e))
        c = e
    node_of = np.zeros(cb, int)
    for k, (a, b) in enumerate(nodes):
        node_of[a:b] = k
    lev = np.zeros(len(nodes), int)
    for c in range(cb):
        for b in range(bn[c] + 1, bn[c + 1]):
            t = node_of[fb[b]]
            if t != node_of[c]:
                lev[t] = max(lev[t], lev[node_of[c]] + 1)
    # factorisation level of every cblk; a node is complete with its last cblk
    clev = np.zeros(cb, int)
    for c in range(cb):
        for b in range(bn[c] + 1, bn[c + 1]):
            clev[fb[b]] = max(clev[fb[b]], clev[c] + 1)
    nrounds, tiles, flops, extra = {}, 0, 0.0, 0
    ntiles = {}
    for a, b in nodes:
        runs = []
        for c in range(a, b):
            o = fc[c] - fc[a]
            runs += [(o + s, o + min(w[c], s + 64)) for s in range(0, w[c], 64)]
        r, nt = 0, 0
        while len(runs) > 1:
            nxt = []
            for i in range(0, len(runs), 2):
                if i + 1 == len(runs):
                    nxt.append(runs[i])
                    break
                x0, m, x1 = runs[i][0], runs[i + 1][0], runs[i + 1][1]
                nt += -(-(x1 - m) // 64) * -(-(m - x0) // 64)
                flops += (x1 - m) * (m - x0) * ((m - x0) + (x1 - m))
                nxt.append((x0, x1))
            runs, r = nxt, r + 1
        nrounds[(a, b)], ntiles[(a, b)] = r, nt
        tiles += nt
        if b - a > 1:
            extra += (fc[b] - fc[a]) ** 2
    # batches: nodes by the level of their last cblk, consecutive levels
    # gathered until 1024 tiles, a node of several cblks or the last level;
    # the rounds of a batch are those of its deepest node
    rounds, brounds, btiles, bleaves, bgroup = 0, 0, 0, 0, False
    nlev = int(clev.max()) + 1
    for lv in range(nlev):
        bgroup = False
        for a, b in nodes:
            if clev[b - 1] != lv:
                continue
            brounds, btiles = max(brounds, nrounds[(a, b)]), btiles + ntiles[(a, b)]
            if b - a > 1:
                bgroup = True
                bleaves += sum(-(-w[c] // 64) for c in range(a, b))
        if brounds == 0 and bleaves == 0:
            continue
        if btiles < 1024 and not bgroup and lv + 1 < nlev:
            continue
        rounds += brounds
        brounds, btiles, bleaves = 0, 0, 0
    return nodes, int(lev.max()) + 1, rounds, tiles, flops, extra


@pytest.mark.parametrize("grid,nmax", [((20, 20, 20), -1), ((30, 30, 30), -1), ((30, 30, 30), 400),
                                       ((30, 30, 30), 0), ((14, 9, 11), -1)])
def test_solve_nodes_follow_their_rule(grid, nmax):
    A = spm.laplacian7(*grid)
    h, S, sym = analyse(A, True, pastix.factotype.LLT, grid=grid)
    pastix.set_amd_iparm(h, pastix.amd_iparm.solve_node_max, nmax)
    info = pastix.solve_info(h)
    sym = pastix.symbol(h)
    nodes, nsl, rounds, tiles, flops, extra = _solve_nodes_from_symbol(sym, 1024 if nmax < 0 else nmax)
    print(grid, nmax, info)
    assert info["nodes"] == len(nodes)
    assert info["merged_nodes"] == sum(1 for a, b in nodes if b - a > 1)
    assert info["solve_levels"] == nsl
    assert info["merge_rounds"] == rounds
    assert info["merge_tiles"] == tiles
    assert info["merge_flops"] == flops
    assert info["merged_entries"] == extra
    assert info["levels"] == sym["nlevels"]
    if nmax == 0:
        assert info["nodes"] == sym["cblknbr"] and info["solve_levels"] == info["levels"]
    elif nmax < 0 and grid[0] >= 20:
        assert info["merged_nodes"] >= 1 and info["solve_levels"] < info["levels"]
    pastix.finalize(h)
