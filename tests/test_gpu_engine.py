"""GPU parity tests: the HIP engine against the CPU oracle through the C-ABI.

Bars: the factor within 1e-12 relative Frobenius of the oracle's (FP64),
solves within the solution's conditioning, normwise backward error
max|b-Ax| / (||A||_inf max|x| + max|b|) <= 1e-12 after refinement.
"""
import numpy as np
import pytest

from pastix_amd import pastix, spm
from tests.helpers import analyse, oracle_factor, rel_frob, lower_mask_panel
import oracle

pytestmark = pytest.mark.gpu

LLT = pastix.factotype.LLT
LU = pastix.factotype.LU


def _berr(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


@pytest.mark.parametrize("grid,bs", [((5, 5, 5), (160, 320)), ((8, 8, 8), (4, 8)), ((10, 10, 10), (8, 16)),
                                     ((20, 20, 20), (160, 320)), ((12, 12, 12), (40, 70))])
def test_potrf_factor_parity(grid, bs):
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=bs[0], max_bs=bs[1])
    pastix.task_numfact(h, S)
    Lg, _ = pastix.coeftab(h)
    O, bc, Lo, _ = oracle_factor(A, sym)
    err = rel_frob(lower_mask_panel(sym, Lg), lower_mask_panel(sym, Lo))
    print(grid, bs, "cblk", sym["cblknbr"], "levels", sym["nlevels"], "rel frob", err)
    assert err <= 1e-12
    assert h.iparm[pastix.iparm.static_pivoting] == 0
    pastix.finalize(h)


@pytest.mark.parametrize("nrhs", [1, 4, 16, 64])
def test_potrf_solve(nrhs):
    grid = (12, 11, 10)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=16, max_bs=32)
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, nrhs, seed=7)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    assert np.abs(x - x0).max() <= 1e-12 * np.abs(x0).max() * 10
    for r in range(nrhs):
        assert _berr(A, x[:, r], b[:, r]) <= 1e-15
    # oracle solve on the oracle factor agrees
    O, bc, Lo, _ = oracle_factor(A, sym)
    p = sym["permtab"]
    xp = np.zeros_like(b)
    xp[p, :] = b
    xo = O.solve(Lo, None, xp)[p, :]
    assert np.abs(x - xo).max() <= 1e-13 * np.abs(xo).max()
    pastix.finalize(h)


def test_potrf_generic_ordering():
    # Scotch requested: the engine orders with its own nested dissection
    A = spm.laplacian7(9, 7, 6, coef=(1.0, 1.0, 0.3))
    h, S, sym = analyse(A, True, LLT)
    pastix.task_numfact(h, S)
    Lg, _ = pastix.coeftab(h)
    O, bc, Lo, _ = oracle_factor(A, sym)
    assert rel_frob(lower_mask_panel(sym, Lg), lower_mask_panel(sym, Lo)) <= 1e-12
    x0, b = spm.rhs_from_x(A, 1, seed=3)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    assert _berr(A, x[:, 0], b[:, 0]) <= 1e-15
    pastix.finalize(h)


@pytest.mark.parametrize("grid,bs", [((6, 6, 6), (160, 320)), ((8, 7, 6), (8, 16)), ((10, 10, 10), (30, 60))])
def test_getrf_factor_parity(grid, bs):
    A = spm.convection_diffusion27(*grid, seed=11)
    h, S, sym = analyse(A, False, LU, grid=grid, min_bs=bs[0], max_bs=bs[1])
    h.dparm[pastix.dparm.epsilon_magn_ctrl] = np.ldexp(1.0, -52)
    pastix.task_numfact(h, S)
    Lg, Ug = pastix.coeftab(h, lu=True)
    O, bc, Lo, Uo = oracle_factor(A, sym, lu=True, crit=np.ldexp(1.0, -52) * np.linalg.norm(A.data))
    eL, eU = rel_frob(Lg, Lo), rel_frob(Ug, Uo)
    print(grid, bs, "L", eL, "U", eU)
    assert eL <= 1e-12 and eU <= 1e-12
    x0, b = spm.rhs_from_x(A, 2, seed=5)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    for r in range(2):
        assert _berr(A, x[:, r], b[:, r]) <= 1e-14
    pastix.finalize(h)


def test_orsirr_lu_gmres_reference_check():
    """The reference's refinement test matrix (test/matrix/orsirr.rua, LU +
    GMRES) with its pass criterion ||b-Ax||_1 / (||A||_1 ||x||_1 eps) <= 1e2
    (spm/src/z_spm_genrhs.c:415-425)."""
    import os
    from tests.helpers import read_rua
    A = read_rua(os.path.join(os.path.dirname(__file__), "golden", "orsirr.rua"))
    h, S, sym = analyse(A, False, LU)
    h.iparm[pastix.iparm.refinement] = pastix.refine.GMRES
    pastix.task_numfact(h, S)
    Lg, Ug = pastix.coeftab(h, lu=True)
    O, bc, Lo, Uo = oracle_factor(A, sym, lu=True)
    assert rel_frob(Lg, Lo) <= 1e-12 and rel_frob(Ug, Uo) <= 1e-12
    x0, b = spm.rhs_from_x(A, 1, seed=24356)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    pastix.task_refine(h, S, b, x)
    r = b[:, 0] - A @ x[:, 0]
    crit = np.abs(r).sum() / (abs(A).sum(axis=0).max() * np.abs(x).sum() * np.finfo(float).eps)
    assert crit <= 1e2
    assert _berr(A, x[:, 0], b[:, 0]) <= 1e-15
    pastix.finalize(h)


def test_refine_cg_fp64():
    grid = (10, 10, 10)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid)
    h.iparm[pastix.iparm.refinement] = pastix.refine.CG
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, 1, seed=1)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    pastix.task_refine(h, S, b, x)
    assert _berr(A, x[:, 0], b[:, 0]) <= 1e-15
    assert pastix.backward_error(h, b[:, 0], x[:, 0]) <= 1e-15
    pastix.finalize(h)


@pytest.mark.parametrize("facto", [LLT, LU])
def test_fp32_factor_fp64_refinement(facto):
    grid = (12, 12, 12)
    if facto == LLT:
        A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
        sym_ = True
    else:
        A = spm.convection_diffusion27(*grid, seed=2)
        sym_ = False
    S = spm.Spm(A, symmetric=sym_)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = facto
    ip[pastix.iparm.ordering] = pastix.order.Personal
    ip[pastix.iparm.refinement] = pastix.refine.CG if facto == LLT else pastix.refine.GMRES
    h = pastix.init(ip, dp)
    pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, spm.SpmFloat)
    o = pastix.orderGrid(*grid)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_blend(h)
    if facto == LU:
        dp[pastix.dparm.epsilon_magn_ctrl] = np.ldexp(1.0, -52)
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, 1, seed=9)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    e0 = _berr(A, x[:, 0], b[:, 0])
    pastix.task_refine(h, S, b, x)
    e1 = _berr(A, x[:, 0], b[:, 0])
    it_host = int(ip[pastix.iparm.nbiter])
    print("facto", facto, "berr fp32 solve", e0, "refined", e1, "iters", it_host)
    assert e0 < 1e-4
    assert e1 <= 1e-12
    # the same solve + refinement on device vectors: same recurrence, same
    # values in the same order, so the same iterate
    db = pastix.DeviceArray(b)
    dx = pastix.DeviceArray(np.zeros_like(b))
    pastix.solve_refine_device(h, db, dx, A.shape[0], 1)
    xd = dx.download()
    assert int(ip[pastix.iparm.nbiter]) == it_host
    assert np.allclose(xd, x, rtol=1e-13, atol=0)
    # refinement from the device copy of the host solve's guess
    x2 = b.copy(order="F")
    pastix.task_solve(h, S, x2)
    dx.upload(x2)
    pastix.refine_device(h, db, dx, A.shape[0], 1)
    assert np.allclose(dx.download(), x, rtol=1e-13, atol=0)
    db.free()
    dx.free()
    pastix.finalize(h)


def test_batched_cg_columns_independent():
    """Several right-hand sides refined together (one CG recurrence per
    column, FP32 factor) give each column's result alone, within rounding."""
    grid = (11, 10, 9)
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
    x0, b = spm.rhs_from_x(A, 5, seed=21)
    b[:, 3] *= 1e-3  # columns of different scales converge at different steps
    db = pastix.DeviceArray(b)
    dx = pastix.DeviceArray(np.zeros_like(b))
    pastix.solve_refine_device(h, db, dx, A.shape[0], 5)
    X = dx.download()
    db.free()
    dx.free()
    for c in range(5):
        bc = np.asfortranarray(b[:, c:c + 1])
        d1 = pastix.DeviceArray(bc)
        d2 = pastix.DeviceArray(np.zeros_like(bc))
        pastix.solve_refine_device(h, d1, d2, A.shape[0], 1)
        xc = d2.download()[:, 0]
        d1.free()
        d2.free()
        assert _berr(A, X[:, c], b[:, c]) <= 1e-12
        assert np.linalg.norm(X[:, c] - xc) <= 1e-12 * np.linalg.norm(xc), c
    pastix.finalize(h)


def test_refactorization_same_pattern():
    grid = (9, 9, 9)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid)
    pastix.task_numfact(h, S)
    for shift in (0.05, 0.09):
        A2 = spm.laplacian7(*grid, diag=6.0 + shift)
        S.set_values(A2)
        pastix.task_numfact(h, S)
        Lg, _ = pastix.coeftab(h)
        O, bc, Lo, _ = oracle_factor(A2, sym)
        assert rel_frob(lower_mask_panel(sym, Lg), lower_mask_panel(sym, Lo)) <= 1e-12
    pastix.finalize(h)


def test_factor_bitwise_reproducible():
    grid = (12, 12, 12)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=20, max_bs=40)
    pastix.task_numfact(h, S)
    L1, _ = pastix.coeftab(h)
    pastix.task_numfact(h, S)
    L2, _ = pastix.coeftab(h)
    assert np.array_equal(L1, L2)
    pastix.finalize(h)


@pytest.mark.parametrize("facto", [LLT, LU])
def test_graph_replays_match_eager(facto):
    """Numfact and device solves replayed from captured HIP graphs give the
    eager results bit for bit, on every replay, with eager work and host
    synchronisation between the replays."""
    grid = (10, 9, 8)
    if facto == LLT:
        A = spm.laplacian7(*grid, diag=6.0)
    else:
        A = spm.convection_diffusion27(*grid, seed=3)
    h, S, sym = analyse(A, facto == LLT, facto, grid=grid, min_bs=16, max_bs=40)
    pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
    pastix.task_numfact(h, S)
    L0, U0 = pastix.coeftab(h, lu=facto == LU)
    x0, b = spm.rhs_from_x(A, 3, seed=11)
    db = pastix.DeviceArray(b)
    pastix.solve_device(h, db, A.shape[0], 3)
    xe = db.download()
    pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 1)
    for rep in range(3):
        pastix.refactorize(h)
        L1, U1 = pastix.coeftab(h, lu=facto == LU)
        assert np.array_equal(L0, L1), f"numfact replay {rep} differs"
        if facto == LU:
            assert np.array_equal(U0, U1), f"numfact replay {rep} differs (U)"
        db.upload(b)
        pastix.solve_device(h, db, A.shape[0], 3)
        assert np.array_equal(db.download(), xe), f"solve replay {rep} differs"
        # eager work in between
        pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 0)
        db.upload(b)
        pastix.solve_device(h, db, A.shape[0], 3)
        assert np.array_equal(db.download(), xe)
        pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 1)
    db.free()
    pastix.finalize(h)


@pytest.mark.parametrize("facto", [LLT, LU])
def test_overlap_matches_single_stream(facto):
    """Updates issued on a second stream beside the next level's panels give
    the single-stream factor bit for bit, eagerly and from a captured graph."""
    grid = (14, 13, 12)
    if facto == LLT:
        A = spm.laplacian7(*grid, diag=6.0)
    else:
        A = spm.convection_diffusion27(*grid, seed=5)
    h, S, sym = analyse(A, facto == LLT, facto, grid=grid, min_bs=16, max_bs=48)
    pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
    pastix.set_amd_iparm(h, pastix.amd_iparm.overlap, 0)
    pastix.task_numfact(h, S)
    L0, U0 = pastix.coeftab(h, lu=facto == LU)
    for graph in (0, 1, 1):
        pastix.set_amd_iparm(h, pastix.amd_iparm.graph, graph)
        pastix.set_amd_iparm(h, pastix.amd_iparm.overlap, 1)
        pastix.refactorize(h)
        L1, U1 = pastix.coeftab(h, lu=facto == LU)
        assert np.array_equal(L0, L1), f"graph={graph}"
        if facto == LU:
            assert np.array_equal(U0, U1), f"graph={graph} (U)"
    pastix.finalize(h)


def test_llt_threshold_above_one_counts_no_padding():
    """With an absolute threshold above 1 no real pivot of a well-scaled SPD
    matrix is replaced; the identity padding of diagonal blocks narrower than
    64 columns (pivots of 1) must not be counted either, and the factor
    matches the oracle's with the same threshold."""
    grid = (9, 8, 7)
    A = (spm.laplacian7(*grid, diag=6.0) * 1e8).tocsr()
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=12, max_bs=40)
    thr = 10.0
    h.dparm[pastix.dparm.epsilon_magn_ctrl] = -thr  # negative: absolute threshold
    pastix.task_numfact(h, S)
    assert h.iparm[pastix.iparm.static_pivoting] == 0
    Lg, _ = pastix.coeftab(h)
    O, bc, Lo, _ = oracle_factor(A, sym, crit=thr)
    assert rel_frob(lower_mask_panel(sym, Lg), lower_mask_panel(sym, Lo)) <= 1e-12
    pastix.finalize(h)


def test_static_pivot_on_device():
    """A zero LU pivot (the first eliminated unknown's diagonal) is replaced by
    the threshold and counted on the device exactly as the oracle does
    (kernels/core_zgetf2sp.c:80-88); the rest of the factor follows from the
    replaced value."""
    facto = LU
    grid = (6, 6, 5)
    A = spm.convection_diffusion27(*grid, seed=1)
    h, S, sym = analyse(A, False, facto, grid=grid, min_bs=8, max_bs=16)
    v = sym["peritab"][0]
    A2 = A.copy().tocsc()
    A2.sort_indices()
    k = A2.indptr[v] + np.searchsorted(A2.indices[A2.indptr[v]:A2.indptr[v + 1]], v)
    A2.data[k] = 0.0
    S2 = spm.Spm(A2, symmetric=False)
    thr = 1e-3
    h.dparm[pastix.dparm.epsilon_magn_ctrl] = -thr  # negative: absolute threshold
    pastix.task_numfact(h, S2)
    assert h.iparm[pastix.iparm.static_pivoting] == 1
    Lg, Ug = pastix.coeftab(h, lu=True)
    O, bc, Lo, Uo = oracle_factor(A2, sym, lu=True, crit=thr)
    assert rel_frob(Lg, Lo) <= 1e-12 and rel_frob(Ug, Uo) <= 1e-12
    assert Lg[0] == pytest.approx(thr, rel=1e-15)
    pastix.finalize(h)


@pytest.mark.parametrize("method", ["SR", "BiCGSTAB", "GMRES"])
def test_refine_sr_bicgstab_fp64_factor_matches_oracle(method):
    """Simple refinement, BiCGSTAB and GMRES (reference
    refinement/z_refine_pivot.c, z_refine_bicgstab.c, z_refine_gmres.c) on
    the device with the FP64 factor: the oracle's drivers on the oracle's
    factor take the same number of steps (the exact preconditioner: 2 passes
    of simple refinement, 1 BiCGSTAB iteration, 1 GMRES step), and the
    iterates agree."""
    grid = (10, 9, 8)
    A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
    h, S, sym = analyse(A, True, LLT, grid=grid)
    h.iparm[pastix.iparm.refinement] = getattr(pastix.refine, method)
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, 1, seed=31)
    x = np.zeros_like(b, order="F")
    pastix.task_refine(h, S, b, x)
    it = int(h.iparm[pastix.iparm.nbiter])
    O, bc, Lo, _ = oracle_factor(A, sym)
    p = sym["permtab"]
    bp = np.zeros(A.shape[0])
    bp[p] = b[:, 0]
    fn = {"SR": O.sr, "BiCGSTAB": O.bicgstab, "GMRES": O.gmres}[method]
    xo, ito, reso = fn(Lo, None, bc, bp, np.zeros(A.shape[0]), itermax=250, eps=1e-12)
    assert it == ito == (2 if method == "SR" else 1)
    assert np.abs(x[:, 0] - xo[p]).max() <= 1e-13 * np.abs(xo).max()
    assert _berr(A, x[:, 0], b[:, 0]) <= 1e-15
    pastix.finalize(h)


@pytest.mark.parametrize("method", ["SR", "BiCGSTAB", "GMRES"])
@pytest.mark.parametrize("facto", [LLT, LU])
def test_refine_sr_bicgstab_fp32_factor(method, facto):
    """FP32 factor + FP64 simple refinement / BiCGSTAB / GMRES reach berr <= 1e-12;
    the device-vector entry point runs the same recurrence (same iterate,
    same count), and the oracle's driver preconditioned with the same FP32
    factor values (applied in FP64) takes about as many steps."""
    grid = (12, 11, 10)
    if facto == LLT:
        A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
    else:
        A = spm.convection_diffusion27(*grid, seed=5)
    h, S, sym = analyse(A, facto == LLT, facto, grid=grid)
    h.iparm[pastix.iparm.refinement] = getattr(pastix.refine, method)
    pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, spm.SpmFloat)
    if facto == LU:
        h.dparm[pastix.dparm.epsilon_magn_ctrl] = np.ldexp(1.0, -52)
    pastix.task_numfact(h, S)
    x0, b = spm.rhs_from_x(A, 1, seed=17)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    pastix.task_refine(h, S, b, x)
    it = int(h.iparm[pastix.iparm.nbiter])
    assert _berr(A, x[:, 0], b[:, 0]) <= 1e-12
    db = pastix.DeviceArray(b)
    dx = pastix.DeviceArray(np.zeros_like(b))
    pastix.solve_refine_device(h, db, dx, A.shape[0], 1)
    assert int(h.iparm[pastix.iparm.nbiter]) == it
    assert np.allclose(dx.download(), x, rtol=1e-13, atol=0)
    Lg, Ug = pastix.coeftab(h, dtype=np.float32, lu=facto == LU)
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A, sym["permtab"])
    p = sym["permtab"]
    bp = np.zeros(A.shape[0])
    bp[p] = b[:, 0]
    xs = bp.copy()
    O.solve(Lg.astype(np.float64), None if Ug is None else Ug.astype(np.float64), xs)
    fn = {"SR": O.sr, "BiCGSTAB": O.bicgstab, "GMRES": O.gmres}[method]
    xo, ito, reso = fn(Lg.astype(np.float64), None if Ug is None else Ug.astype(np.float64), bc, bp, xs, eps=1e-12)
    print(method, facto, "device steps", it, "oracle steps", ito, "oracle res", reso)
    assert reso <= 1e-12 and abs(it - ito) <= 1
    assert np.abs(x[:, 0] - xo[p]).max() <= 1e-9 * np.abs(xo).max()
    pastix.finalize(h)
