"""pastix_amd_backward_error_device_noalloc (no reference counterpart; for
callers that cannot allocate) against pastix_amd_backward_error on the same
vectors, through the C-ABI: two calls in a row on two instances, the second
with a smaller input (fewer rows and columns), so that it reuses the scratch
the first one filled.  The partial maxima are the same products in the same
order as the existing entry point's kernel and the final quotient is rounded
step by step as the host rounds it, so the results agree to the last bit or
within 2 ulps where the host compiler contracts."""
import numpy as np
import pytest

from pastix_amd import pastix, spm

pytestmark = pytest.mark.gpu


def _factored(grid):
    A = spm.laplacian7(*grid, diag=6.0)
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    ip[pastix.iparm.ordering] = pastix.order.Personal
    h = pastix.init(ip, dp)
    o = pastix.orderGrid(*grid)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    pastix.task_numfact(h, S)
    return A, S, h


def _check(A, S, h, nrhs, seed):
    n = A.shape[0]
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-1, 1, size=(n, nrhs))
    b = np.asfortranarray(A @ x0)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    # perturb the solutions so that the backward errors are far from rounding
    x[: n // 3] += 1e-9 * rng.uniform(-1, 1, size=(n // 3, nrhs))
    x = np.asfortranarray(x)
    db, dx = pastix.DeviceArray(b), pastix.DeviceArray(x)
    dberr = pastix.DeviceArray(np.full(nrhs, -1.0))
    pastix.backward_error_device_noalloc(h, db, dx, nrhs, dberr)
    got = dberr.download()
    for d in (db, dx, dberr):
        d.free()
    want = np.array([pastix.backward_error(h, b[:, c], x[:, c]) for c in range(nrhs)])
    print(f"n={n} nrhs={nrhs} berr {got.min():.3e}..{got.max():.3e}")
    assert np.all(want > 0)
    assert np.all(np.abs(got - want) <= 2 * np.spacing(want)), (got, want)


def test_noalloc_backward_error_matches_existing_entry_point():
    A1, S1, h1 = _factored((12, 11, 10))
    A2, S2, h2 = _factored((7, 6, 5))
    _check(A1, S1, h1, 9, seed=1)
    _check(A2, S2, h2, 3, seed=2)
    # beyond the scratch: a clear error, nothing launched
    n = A2.shape[0]
    db = pastix.DeviceArray(np.zeros((n, 65), order="F"))
    dberr = pastix.DeviceArray(np.zeros(65))
    with pytest.raises(pastix.PastixError):
        pastix.backward_error_device_noalloc(h2, db, db, 65, dberr)
    db.free()
    dberr.free()
    pastix.finalize(h1)
    pastix.finalize(h2)
