"""Residual SpMV and refinement vector kernels on the device, through the
C-ABI (pastix_amd_bcsc_spmv, pastix_amd_bvec_*), checked as the reference's
own tests check them:

- SpMV: test/z_bcsc_tests.c:38-124.  y = alpha op(A) x + beta y0 on the
  permuted matrix, permuted back, against the product on the user's matrix:
  ||dy||_max / ((||A||_inf + ||x||_inf + ||y0||_inf) n eps) <= 10
  (:104-112); also against the oracle's SpMV (reference
  bcsc/bcsc_zspmv.c:55-202) at 1e-15 relative.
- gemv: test/z_bvec_tests.c:64-140 at its default sizes (m = 10000, n =
  IPARM_GMRES_IM = 25, alpha 3.5, beta -4.8): ||dy||_inf / ((||A||_inf +
  ||x||_1 + ||y||_inf) n eps) <= 10.
- dot, nrm2, axpy, scal, copy (timed only in the reference's bvec tests):
  numpy with the standard rounding bounds; nrm2 also where the plain sum of
  squares over- or underflows (the reference's scaled sum of squares,
  bcsc/bvec_zcompute.c:67-120, stays exact there).
"""
import numpy as np
import pytest

from pastix_amd import pastix, spm
from tests.helpers import analyse, read_rua
import oracle

pytestmark = pytest.mark.gpu
EPS = np.ldexp(1.0, -53)
LLT, LU = pastix.factotype.LLT, pastix.factotype.LU


def _instance(case):
    if case == "laplacian_llt":
        grid = (11, 10, 9)
        A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
        h, S, sym = analyse(A, True, LLT, grid=grid)
    elif case == "convdiff_lu":
        grid = (10, 9, 8)
        A = spm.convection_diffusion27(*grid, seed=2)
        h, S, sym = analyse(A, False, LU, grid=grid)
    else:  # orsirr: the reference's bundled general matrix (test/matrix/orsirr.rua)
        import os
        A = read_rua(os.path.join(os.path.dirname(__file__), "golden", "orsirr.rua"))
        h, S, sym = analyse(A, False, LU)
    pastix.task_numfact(h, S)
    return A.tocsc(), h, sym


@pytest.mark.parametrize("case", ["laplacian_llt", "convdiff_lu", "orsirr_lu"])
@pytest.mark.parametrize("trans", ["NoTrans", "Trans"])
def test_bcsc_spmv_reference_bound(case, trans):
    A, h, sym = _instance(case)
    n = A.shape[0]
    rng = np.random.default_rng(35469)
    alpha, beta = rng.uniform(-1, 1, 2)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    op = A if trans == "NoTrans" else A.T.tocsc()
    ys = alpha * (op @ x) + beta * y0
    p = sym["permtab"]
    xp, yp = np.empty(n), np.empty(n)
    xp[p], yp[p] = x, y0
    dx, dy = pastix.DeviceArray(xp), pastix.DeviceArray(yp)
    pastix.bcsc_spmv(h, getattr(pastix.trans, trans), alpha, dx, beta, dy)
    yd = dy.download()[p]
    anorm = abs(A).sum(axis=1).max()
    result = np.abs(yd - ys).max() / ((anorm + np.abs(x).max() + np.abs(y0).max()) * n * EPS)
    assert np.isfinite(result) and result <= 10.0, result
    if trans == "NoTrans":
        # the oracle's SpMV on the same permuted matrix
        bc = oracle.bcsc(A, p)
        yo = oracle.Oracle(sym).spmv(bc, xp, alpha=alpha, beta=beta, y=yp)
        assert np.abs(dy.download() - yo).max() <= 1e-15 * np.abs(yo).max()
    # beta = 0 does not read y (NaN in y stays out of the result)
    dy.upload(np.full(n, np.nan))
    pastix.bcsc_spmv(h, getattr(pastix.trans, trans), 1.0, dx, 0.0, dy)
    assert np.isfinite(dy.download()).all()
    dx.free(), dy.free()
    pastix.finalize(h)


def test_bvec_gemv_reference_bound():
    A_, h, _ = _instance("laplacian_llt")
    m, n = 10000, 25
    rng = np.random.default_rng(7213)
    A = np.asfortranarray(rng.uniform(-0.5, 0.5, (m, n)))
    x = rng.uniform(-0.5, 0.5, n)
    y = rng.uniform(-0.5, 0.5, m)
    dA, dx, dy = pastix.DeviceArray(A), pastix.DeviceArray(x), pastix.DeviceArray(y)
    pastix.bvec_gemv(h, m, n, 3.5, dA, m, dx, -4.8, dy)
    y0 = 3.5 * (A @ x) - 4.8 * y
    normR = np.abs(dy.download() - y0).max()
    result = normR / ((np.abs(A).sum(axis=1).max() + np.abs(x).sum() + np.abs(y).max()) * n * EPS)
    assert np.isfinite(result) and result <= 10.0, result
    # leading dimension larger than m, beta = 0 ignores y
    A2 = np.asfortranarray(rng.uniform(-1, 1, (m + 7, n)))
    dA2 = pastix.DeviceArray(A2)
    dy.upload(np.full(m, np.nan))
    pastix.bvec_gemv(h, m, n, 1.0, dA2, m + 7, dx, 0.0, dy)
    assert np.abs(dy.download() - A2[:m] @ x).max() <= 10 * n * EPS * np.abs(A2[:m]).sum(axis=1).max()
    for d in (dA, dx, dy, dA2):
        d.free()
    pastix.finalize(h)


def test_bvec_blas1_against_numpy():
    A_, h, _ = _instance("laplacian_llt")
    rng = np.random.default_rng(11)
    for n in (1, 255, 256, 257, 100003):
        x = rng.uniform(-1, 1, n)
        y = rng.uniform(-1, 1, n)
        dx, dy = pastix.DeviceArray(x), pastix.DeviceArray(y)
        d = pastix.bvec_dot(h, n, dx, dy)
        assert abs(d - np.dot(x, y)) <= n * EPS * np.abs(x * y).sum()
        nr = pastix.bvec_nrm2(h, n, dx)
        assert nr == pytest.approx(np.linalg.norm(x), rel=4 * np.sqrt(n) * EPS)
        pastix.bvec_axpy(h, n, -0.75, dx, dy)
        # one rounding per entry (fused or not): within one ulp of numpy's
        assert np.abs(dy.download() - (y - 0.75 * x)).max() <= 2 * EPS * (np.abs(y) + 0.75 * np.abs(x)).max()
        pastix.bvec_scal(h, n, 3.0, dx)
        assert np.array_equal(dx.download(), 3.0 * x)
        pastix.bvec_copy(h, n, dx, dy)
        assert np.array_equal(dy.download(), 3.0 * x)
        # the reductions are reproducible bit for bit
        assert pastix.bvec_dot(h, n, dx, dy) == pastix.bvec_dot(h, n, dx, dy)
        dx.free(), dy.free()
    # empty vectors
    dz = pastix.DeviceArray(np.zeros(4))
    assert pastix.bvec_dot(h, 0, dz, dz) == 0.0
    assert pastix.bvec_nrm2(h, 0, dz) == 0.0
    pastix.bvec_axpy(h, 0, 1.0, dz, dz)
    dz.free()
    pastix.finalize(h)


@pytest.mark.parametrize("scale", [1e200, 1e-200, 1e-310])
def test_bvec_nrm2_scaled_sum_of_squares(scale):
    """Entries whose squares over- or underflow: sqrt(x.x) gives inf or 0,
    the scaled sum of squares gives the norm."""
    A_, h, _ = _instance("laplacian_llt")
    rng = np.random.default_rng(3)
    n = 50000
    u = rng.uniform(-1, 1, n)
    x = u * scale
    dx = pastix.DeviceArray(x)
    nr = pastix.bvec_nrm2(h, n, dx)
    assert nr == pytest.approx(np.linalg.norm(u) * scale, rel=1e-12 if scale > 1e-300 else 1e-6)
    dx.free()
    pastix.finalize(h)
