"""Pins the CPU oracle against what the mathematics fixes.

- LL^t: the Cholesky factor of P A P^T is unique, so a dense
  numpy.linalg.cholesky of the permuted matrix fixes every factor entry.
- LU without pivoting is unique when no static pivot fires: a plain dense
  Doolittle elimination (written here, scipy's LU pivots) fixes the factor.
- Solves: dense numpy.linalg.solve.  SpMV: scipy.sparse products.
- Flop counts: closed forms checked against brute-force operation counts.
- CG refinement: convergence to the stop criterion.
- Simple refinement: on A = I with M = c I every pass contracts the error by
  rho = 1 - 1/c exactly, so the pass count and the stagnation stop
  (berr must halve) have closed forms.
- BiCGSTAB: with the exact factor as preconditioner it finishes in one
  iteration; with a perturbed one it converges and its recursive residual
  stays the true one.
A planted mistake (dropping an update, a transposed operand, a wrong pivot
rule) fails at least one of these.
"""
import numpy as np
import pytest

from pastix_amd import pastix, spm
from tests.helpers import analyse, oracle_factor, rel_frob
import oracle

LLT, LU = pastix.factotype.LLT, pastix.factotype.LU


def _dense_lu_nopiv(B):
    A = B.copy()
    n = A.shape[0]
    for k in range(n):
        A[k + 1:, k] /= A[k, k]
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    return np.tril(A, -1) + np.eye(n), np.triu(A)


@pytest.mark.parametrize("grid,bs", [((4, 4, 4), (160, 320)), ((6, 5, 4), (3, 6)), ((7, 7, 7), (8, 16)),
                                     ((10, 10, 10), (160, 320))])
def test_potrf_matches_dense_cholesky(grid, bs):
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=bs[0], max_bs=bs[1])
    O, bc, L, _ = oracle_factor(A, sym)
    Lref = np.linalg.cholesky(bc["B"].toarray())
    assert rel_frob(O.to_dense(L), Lref) <= 1e-14
    pastix.finalize(h)


def test_potrf_general_ordering_matches_dense():
    A = spm.laplacian7(8, 6, 5, coef=(1.0, 0.5, 0.25))
    h, S, sym = analyse(A, True, LLT)
    O, bc, L, _ = oracle_factor(A, sym)
    Lref = np.linalg.cholesky(bc["B"].toarray())
    assert rel_frob(O.to_dense(L), Lref) <= 1e-14
    pastix.finalize(h)


@pytest.mark.parametrize("grid,bs", [((5, 5, 5), (160, 320)), ((6, 6, 4), (4, 8))])
def test_getrf_matches_dense_doolittle(grid, bs):
    A = spm.convection_diffusion27(*grid, seed=4)
    h, S, sym = analyse(A, False, LU, grid=grid, min_bs=bs[0], max_bs=bs[1])
    O, bc, L, U = oracle_factor(A, sym, lu=True, crit=0.0)
    Lref, Uref = _dense_lu_nopiv(bc["B"].toarray())
    Ld = O.to_dense(L, lower_only=False)
    Ud = O.to_dense(U, lower_only=False)
    n = A.shape[0]
    # L panel holds unit-lower L below the diagonal and U on/above it inside
    # diagonal blocks; U panel holds U^T below the diagonal blocks.
    strict = np.tril(np.ones((n, n)), -1).astype(bool)
    assert np.abs(Ld[strict] - Lref[strict]).max() <= 1e-13 * np.abs(Lref).max()
    diagd = np.diag(Ld)
    assert np.abs(diagd - np.diag(Uref)).max() <= 1e-13 * np.abs(Uref).max()
    assert np.abs(Ud[strict] - Uref.T[strict]).max() <= 1e-13 * np.abs(Uref).max()
    pastix.finalize(h)


def test_static_pivot_counts_and_replaces():
    # a zero pivot is replaced by the threshold and counted (core_zgetf2sp rule)
    A = spm.convection_diffusion27(4, 4, 4, seed=1)
    h, S, sym = analyse(A, False, LU, grid=(4, 4, 4))
    v = sym["peritab"][0]  # the first eliminated unknown: its pivot is its diagonal entry
    A2 = A.copy()
    k = A2.indptr[v] + np.searchsorted(A2.indices[A2.indptr[v]:A2.indptr[v + 1]], v)
    A2.data[k] = 0.0
    O = oracle.Oracle(sym)
    L, U, _ = O.fill(oracle.bcsc(A2, sym["permtab"]), lu=True)
    rc, npiv = O.getrf(L, U, 1e-3)
    assert rc == 0 and npiv == 1
    assert L[0] == pytest.approx(1e-3)  # replaced by +threshold (non-negative zero)
    L2, U2, _ = O.fill(oracle.bcsc(A, sym["permtab"]), lu=True)
    assert O.getrf(L2, U2, 1e-3) == (0, 0)
    pastix.finalize(h)


@pytest.mark.parametrize("lu", [False, True])
def test_oracle_solve_matches_dense(lu):
    grid = (6, 5, 5)
    A = spm.convection_diffusion27(*grid, seed=3) if lu else spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, not lu, LU if lu else LLT, grid=grid, min_bs=4, max_bs=8)
    O, bc, L, U = oracle_factor(A, sym, lu=lu, crit=0.0 if lu else None)
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, size=(A.shape[0], 3))
    p = sym["permtab"]
    bp = np.zeros_like(b)
    bp[p, :] = b
    x = O.solve(L, U, bp)[p, :]
    xref = np.linalg.solve(A.toarray(), b)
    assert np.abs(x - xref).max() <= 1e-12 * np.abs(xref).max()
    pastix.finalize(h)


def test_oracle_spmv_matches_scipy():
    A = spm.convection_diffusion27(5, 4, 3, seed=8)
    perm = np.random.default_rng(1).permutation(A.shape[0])
    bc = oracle.bcsc(A, perm)
    sym = {"n": A.shape[0], "cblknbr": 1, "bloknbr": 1, "coefnbr": 1,
           "cblk_fcol": np.array([0, A.shape[0]]), "cblk_bloknum": np.array([0, 1]),
           "cblk_stride": np.array([A.shape[0]]), "cblk_coef": np.array([0, 1]),
           "blok_frow": np.array([0]), "blok_lrow": np.array([A.shape[0] - 1]), "blok_fcblk": np.array([0]),
           "blok_roff": np.array([0])}
    O = oracle.Oracle(sym)
    x = np.random.default_rng(2).uniform(-1, 1, A.shape[0])
    y0 = np.random.default_rng(3).uniform(-1, 1, A.shape[0])
    y = O.spmv(bc, x, alpha=-0.5, beta=2.0, y=y0)
    yref = -0.5 * (bc["B"] @ x) + 2.0 * y0
    # reference bound of test/z_bcsc_tests.c:104-112
    n = A.shape[0]
    eps = np.finfo(float).eps
    bound = (abs(bc["B"]).sum(axis=1).max() + np.abs(x).max() + np.abs(y0).max()) * n * eps * 10
    assert np.abs(y - yref).max() <= bound


def _count_cholesky_ops(n):
    """Operation count of an unblocked Cholesky: (muls incl. sqrt/div, adds)."""
    mul = add = 0
    for k in range(n):
        mul += 1  # sqrt
        mul += n - k - 1  # scale column
        m = n - k - 1
        mul += m * (m + 1) // 2
        add += m * (m + 1) // 2
    return mul, add


def _count_lu_ops(n):
    mul = add = 0
    for k in range(n):
        mul += n - k - 1  # divisions
        mul += (n - k - 1) ** 2
        add += (n - k - 1) ** 2
    return mul, add


def _one_cblk_symbol(n, extra_rows=0):
    st = n + extra_rows
    nb = 1 if extra_rows == 0 else 2
    return {"n": st, "cblknbr": 1 if extra_rows == 0 else 2, "bloknbr": nb, "coefnbr": st * n,
            "cblk_fcol": np.array([0, n] + ([st] if extra_rows else [])),
            "cblk_bloknum": np.array([0, nb] + ([nb] if extra_rows else [])),
            "cblk_stride": np.array([st] + ([extra_rows] if extra_rows else [])),
            "cblk_coef": np.array([0, st * n] + ([st * n] if extra_rows else [])),
            "blok_frow": np.array([0] + ([n] if extra_rows else [])),
            "blok_lrow": np.array([n - 1] + ([st - 1] if extra_rows else [])),
            "blok_fcblk": np.array([0] + ([1] if extra_rows else [])),
            "blok_roff": np.array([0] + ([n] if extra_rows else []))}


@pytest.mark.parametrize("n", [1, 2, 3, 7, 40])
def test_thflops_closed_forms_vs_brute_force(n):
    O = oracle.Oracle(_one_cblk_symbol(n))
    mul, add = _count_cholesky_ops(n)
    assert O.thflops(lu=False) == pytest.approx(mul + add, rel=1e-12)
    mul, add = _count_lu_ops(n)
    # LAPACK's FLOPS_DGETRF(n, n) counts n operations more than the
    # elimination performs (its 2n/3 + n/6 lower-order terms)
    assert O.thflops(lu=True) == pytest.approx(mul + add + n, rel=1e-12)


def test_thflops_offdiag_terms():
    # one cblk of width N with M rows below: POTRF(N) + TRSM(Right, M, N) + SYRK(N, M)
    N, M = 5, 9
    O = oracle.Oracle(_one_cblk_symbol(N, extra_rows=M))
    rest = O.thflops() - oracle.Oracle(_one_cblk_symbol(N)).thflops() - oracle.Oracle(_one_cblk_symbol(M)).thflops()
    trsm_ops = M * N * N        # M triangular solves of order N: N divisions + N(N-1)/2 (mul + add)
    syrk_ops = 2 * N * (M * (M + 1) // 2)  # lower triangle of the M x M product, N mul + N add each
    # the reference defines FADDS_TRSM as FMULS_TRMM, which adds one count per (row, column)
    assert rest == pytest.approx(trsm_ops + M * N + syrk_ops, rel=1e-12)


def test_cg_refinement_converges():
    grid = (7, 6, 5)
    A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
    h, S, sym = analyse(A, True, LLT, grid=grid)
    O, bc, L, _ = oracle_factor(A, sym)
    # perturb the factor so that refinement has work to do
    Lp = L * (1.0 + 1e-6 * np.random.default_rng(0).standard_normal(L.shape))
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, A.shape[0])
    x, it, res = O.cg(Lp, bc, b, np.zeros(A.shape[0]), itermax=50, eps=1e-12)
    assert res <= 1e-12 and 1 <= it <= 10
    assert np.abs(bc["B"] @ x - b).max() <= 1e-10
    pastix.finalize(h)


def _identity_case(c):
    """A = I on a 3D grid's symbol with the factor sqrt(c) I: M = c I."""
    grid = (6, 5, 4)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid)
    import scipy.sparse as sps
    I = sps.identity(A.shape[0], format="csc")
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(I, sym["permtab"])
    L, _, rc = O.fill(bc)
    assert rc == 0
    pastix.finalize(h)
    return O, bc, L * np.sqrt(c)


@pytest.mark.parametrize("rho,passes,stop_berr", [(0.25, 21, None), (0.6, 2, 0.6), (0.0, 2, None)])
def test_sr_pass_counts_closed_form(rho, passes, stop_berr):
    # x_{k+1} = x_k + (b - x_k) / c: berr_k = rho^k with rho = 1 - 1/c.  The
    # reference loop (z_refine_pivot.c) corrects while berr > eps and berr <=
    # last_berr / 2: rho = 0.25 needs 20 corrections to reach 1e-12 (21 passes
    # with the stopping one); rho = 0.6 stops after one correction, berr 0.6
    # not having halved; the exact preconditioner takes one correction.
    O, bc, L = _identity_case(1.0 / (1.0 - rho))
    b = np.random.default_rng(3).uniform(-1, 1, O.n)
    x, it, res = O.sr(L, None, bc, b, np.zeros(O.n), itermax=100, eps=1e-12)
    assert it == passes
    if stop_berr is None:
        assert res <= 1e-12
        assert np.abs(x - b).max() <= 1e-11
    else:
        assert res == pytest.approx(stop_berr, rel=1e-12)
        assert np.abs(x - b / (1.0 / (1.0 - rho))).max() <= 1e-14


def test_sr_itermax_bound():
    O, bc, L = _identity_case(1.0 / (1.0 - 0.25))
    b = np.random.default_rng(4).uniform(-1, 1, O.n)
    x, it, res = O.sr(L, None, bc, b, np.zeros(O.n), itermax=3, eps=1e-12)
    # 3 corrections, then the pass that stops on itermax
    assert it == 4 and res == pytest.approx(0.25 ** 3, rel=1e-10)


@pytest.mark.parametrize("lu", [False, True])
def test_bicgstab_refinement(lu):
    grid = (7, 6, 5)
    if lu:
        A = spm.convection_diffusion27(*grid, seed=4)
    else:
        A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
    h, S, sym = analyse(A, not lu, LU if lu else LLT, grid=grid)
    O, bc, L, U = oracle_factor(A, sym, lu=lu)
    b = np.random.default_rng(6).uniform(-1, 1, A.shape[0])
    # exact preconditioner: alpha = 1 and s is rounding noise, one iteration
    x, it, res = O.bicgstab(L, U, bc, b, np.zeros(A.shape[0]), itermax=50, eps=1e-12)
    assert it == 1 and res <= 1e-12
    # perturbed preconditioner: converges, and the final iterate's true
    # residual agrees with the recursive one
    rng = np.random.default_rng(0)
    Lp = L * (1.0 + 1e-5 * rng.standard_normal(L.shape))
    Up = None if U is None else U * (1.0 + 1e-5 * rng.standard_normal(U.shape))
    x, it, res = O.bicgstab(Lp, Up, bc, b, np.zeros(A.shape[0]), itermax=50, eps=1e-12)
    assert 2 <= it <= 20 and res <= 1e-12
    true = np.linalg.norm(bc["B"] @ x - b) / np.linalg.norm(b)
    assert true <= 1e-11
    # simple refinement on the same perturbed factor converges too
    x2, it2, res2 = O.sr(Lp, Up, bc, b, np.zeros(A.shape[0]), itermax=50, eps=1e-12)
    assert res2 <= 1e-12 and np.abs(x2 - x).max() <= 1e-10 * np.abs(x).max()
    pastix.finalize(h)


# ---- GMRES (reference refinement/z_refine_gmres.c) -------------------------
# Any correct right-preconditioned GMRES cycle of k steps returns the point of
# x0 + M^{-1} K_k(A M^{-1}, r0) with the smallest ||b - A x||_2; on A = I with
# M = diag(c) and c taking m distinct values, A M^{-1} has a minimal
# polynomial of degree m, so the exact solution is reached at step m.

def _dense_precond(O, L, U):
    """M^{-1} as a dense matrix, one oracle solve per unit vector."""
    E = np.eye(O.n)
    return O.solve(L, U, E.copy(order="F"))


def _minres_cycle(A, Minv, b, x0, k):
    r0 = b - A @ x0
    AM = A @ Minv
    K = np.empty((len(b), k))
    v = r0.copy()
    for j in range(k):
        K[:, j] = v
        v = AM @ v
    Q, _ = np.linalg.qr(K)
    c, *_ = np.linalg.lstsq(AM @ Q, r0, rcond=None)
    return x0 + Minv @ (Q @ c)


def test_gmres_exact_preconditioner_one_step():
    for lu in (False, True):
        grid = (7, 6, 5)
        A = spm.convection_diffusion27(*grid, seed=4) if lu else spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
        h, S, sym = analyse(A, not lu, LU if lu else LLT, grid=grid)
        O, bc, L, U = oracle_factor(A, sym, lu=lu)
        b = np.random.default_rng(8).uniform(-1, 1, A.shape[0])
        x, it, res = O.gmres(L, U, bc, b, np.zeros(A.shape[0]), itermax=50, im=25, eps=1e-12)
        assert it == 1 and res <= 1e-12
        assert np.abs(bc["B"] @ x - b).max() <= 1e-12
        pastix.finalize(h)


@pytest.mark.parametrize("values,steps", [((1.0, 2.0, 3.0), 3), ((1.5, 4.0), 2), ((2.0, 2.5, 3.0, 7.0, 9.0), 5)])
def test_gmres_steps_equal_distinct_eigenvalues(values, steps):
    O, bc, L1 = _identity_case(1.0)
    c = np.array(values)[np.arange(O.n) % len(values)]
    # permuted order: the factor's diagonal holds sqrt(c) (M = diag(c))
    sym = O.sym
    L = L1.copy()
    for k in range(len(sym["cblk_stride"])):
        fc, lc, st = sym["cblk_fcol"][k], sym["cblk_fcol"][k + 1], sym["cblk_stride"][k]
        for j in range(lc - fc):
            L[sym["cblk_coef"][k] + j + j * st] = np.sqrt(c[fc + j])
    b = np.random.default_rng(9).uniform(-1, 1, O.n)
    x, it, res = O.gmres(L, None, bc, b, np.zeros(O.n), itermax=50, im=25, eps=1e-12)
    assert it == steps and res <= 1e-12
    assert np.abs(x - b).max() <= 1e-11


def test_gmres_iterates_are_minimal_residual_and_restart():
    grid = (6, 5, 4)
    A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
    h, S, sym = analyse(A, True, LLT, grid=grid)
    O, bc, L, _ = oracle_factor(A, sym)
    Lp = L * (1.0 + 1e-2 * np.random.default_rng(1).standard_normal(L.shape))
    Minv = _dense_precond(O, Lp, None)
    B = bc["B"].toarray()
    b = np.random.default_rng(2).uniform(-1, 1, O.n)
    x0 = np.zeros(O.n)
    # one cycle cut at k steps by the step limit: the k-step minimal residual point
    for k in (1, 2, 3):
        x, it, _ = O.gmres(Lp, None, bc, b, x0, itermax=k, im=25, eps=1e-30)
        xr = _minres_cycle(B, Minv, b, x0, k)
        assert it == k
        assert np.abs(x - xr).max() <= 1e-9 * np.abs(xr).max()
    # GMRES(2), two cycles: the second restarts from the true residual of the first
    x, it, _ = O.gmres(Lp, None, bc, b, x0, itermax=4, im=2, eps=1e-30)
    xr = _minres_cycle(B, Minv, b, _minres_cycle(B, Minv, b, x0, 2), 2)
    assert it == 4
    assert np.abs(x - xr).max() <= 1e-9 * np.abs(xr).max()
    # a nonzero initial guess enters through r0 = b - A x0
    x1 = np.random.default_rng(3).uniform(-1, 1, O.n)
    x, it, _ = O.gmres(Lp, None, bc, b, x1, itermax=2, im=25, eps=1e-30)
    xr = _minres_cycle(B, Minv, b, x1, 2)
    assert np.abs(x - xr).max() <= 1e-9 * np.abs(xr).max()
    pastix.finalize(h)


# ---- static scheduler (sopalin/sequential_zpotrf.c:71-120) -----------------
@pytest.mark.parametrize("lu", [False, True])
def test_static_scheduler_matches_sequential(lu):
    """One thread walks the cblks in order: the sequential factor bit for bit.
    Several threads apply a target's updates in another order: the same factor
    to rounding, and (LL^t) the dense Cholesky of P A P^T."""
    grid = (14, 13, 12)
    A = spm.convection_diffusion27(*grid, seed=2) if lu else spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, not lu, LU if lu else LLT, grid=grid, min_bs=8, max_bs=24)
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A, sym["permtab"])
    crit = np.sqrt(np.ldexp(1.0, -53)) * np.linalg.norm(bc["vals"])
    L0, U0, _ = O.fill(bc, lu=lu)
    L1, U1 = L0.copy(), None if U0 is None else U0.copy()
    if lu:
        O.getrf(L1, U1, crit)
    else:
        O.potrf(L1, crit)
    L2, U2 = L0.copy(), None if U0 is None else U0.copy()
    rc, npiv = O.factor_static(L2, U2, crit, 1)
    assert rc == 0 and npiv == 0
    assert np.array_equal(L1, L2) and (U1 is None or np.array_equal(U1, U2))
    for nt in (3, 8):
        L3, U3 = L0.copy(), None if U0 is None else U0.copy()
        rc, npiv = O.factor_static(L3, U3, crit, nt)
        assert rc == 0
        assert rel_frob(L3, L1) <= 1e-13 and (U1 is None or rel_frob(U3, U1) <= 1e-13)
    if not lu:
        D = O.to_dense(L3)
        Bd = bc["B"].toarray()
        assert rel_frob(D, np.linalg.cholesky(Bd)) <= 1e-13
    pastix.finalize(h)


def test_static_scheduler_negative_pivot_ends_every_thread():
    grid = (10, 9, 8)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, LLT, grid=grid, min_bs=8, max_bs=24)
    import scipy.sparse as sps
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A - 7.0 * sps.eye(A.shape[0]), sym["permtab"])
    L, _, _ = O.fill(bc)
    rc, _ = O.factor_static(L, None, 1e-8, 6)
    assert rc == -1
    pastix.finalize(h)
