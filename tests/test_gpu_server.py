"""Serving loop on the GPU: continuous batching of solves and refactorised
versions A + shift*I overlapped on two slots.

Bars: every solve reaches berr = max|b-Ax| / (||A||_inf max|x| + max|b|)
<= 1e-12 against the matrix version it was bound to at submission, and
equals the same request solved alone (one request, one version, plain
instance) within cond(A) * 1e-12, taken as 1e-8 relative in max norm for
these anisotropic Laplacians (cond(A) < 1e4).
"""
import numpy as np
import pytest
import scipy.sparse as sps

from pastix_amd import pastix, spm

pytestmark = pytest.mark.gpu

GRID = (12, 11, 10)


def _berr(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


def _slot(A, prec):
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    ip[pastix.iparm.ordering] = pastix.order.Personal
    ip[pastix.iparm.refinement] = pastix.refine.CG
    h = pastix.init(ip, dp)
    pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
    o = pastix.orderGrid(*GRID)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    pastix.task_numfact(h, S)
    return h, S


def test_server_mix_equals_alone():
    A = spm.laplacian7(*GRID, coef=(1.0, 1.0, 0.01))
    n = A.shape[0]
    slots = [_slot(A, spm.SpmFloat) for _ in range(2)]
    srv = pastix.Server([h for h, _ in slots], max_rhs=16)
    rng = np.random.default_rng(21)
    shifts = {0: 0.0}
    reqs = []
    for k in range(14):
        if k in (4, 9):
            shift = float(rng.uniform(0.0, 0.1))
            t = srv.submit_refactor(shift)
            reqs.append(("refactor", t, shift, None))
        else:
            nrhs = int(rng.choice([1, 4, 16]))
            b = rng.uniform(-1.0, 1.0, size=(n, nrhs))
            t = srv.submit_solve(b.copy(order="F"))
            reqs.append(("solve", t, None, b))
    results = []
    for kind, t, shift, b in reqs:
        x, lat, ver = srv.wait(t)
        if kind == "refactor":
            shifts[ver] = shift
        else:
            results.append((b, x, ver, lat))
    st = srv.stats()
    srv.close()
    assert st["pending"] == 0 and st["version"] == 2
    # alone: one request, its own version, a plain instance
    h1, _ = slots[0]
    for b, x, ver, lat in results:
        Av = A + shifts[ver] * sps.eye(n)
        for c in range(b.shape[1]):
            assert _berr(Av, x[:, c], b[:, c]) <= 1e-12
        pastix.refactorize_shift(h1, shifts[ver])
        xa = b.copy(order="F")
        pastix.task_solve(h1, slots[0][1], xa)
        pastix.task_refine(h1, slots[0][1], b.copy(order="F"), xa)
        assert np.abs(x - xa).max() <= 1e-8 * np.abs(xa).max()
    for h, _ in slots:
        pastix.finalize(h)


def test_server_failed_refactorisation_ends_its_solves():
    """A refactorisation that fails (A - 7 I is not positive definite: LL^t
    stops on a negative diagonal term) ends the solves bound to its version
    with its error status instead of leaving them queued; solves of the
    earlier version still succeed, and a later refactorisation serves its
    solves again.  One slot, so a stuck version would stall everything."""
    A = spm.laplacian7(*GRID, coef=(1.0, 1.0, 0.01))
    n = A.shape[0]
    h, S = _slot(A, spm.SpmFloat)
    srv = pastix.Server([h], max_rhs=8)
    rng = np.random.default_rng(5)
    b0 = rng.uniform(-1.0, 1.0, size=(n, 2))
    t0 = srv.submit_solve(b0.copy(order="F"))
    tbad = srv.submit_refactor(-7.0)
    b1 = rng.uniform(-1.0, 1.0, size=(n, 1))
    t1 = srv.submit_solve(b1.copy(order="F"))
    x0, _, v0 = srv.wait(t0)
    assert v0 == 0 and all(_berr(A, x0[:, c], b0[:, c]) <= 1e-12 for c in range(2))
    with pytest.raises(pastix.PastixError):
        srv.wait(tbad)
    with pytest.raises(pastix.PastixError):
        srv.wait(t1)
    # bound to the failed version after it failed: ends at once with its status
    t2 = srv.submit_solve(b1.copy(order="F"))
    with pytest.raises(pastix.PastixError):
        srv.wait(t2)
    tok = srv.submit_refactor(0.05)
    b3 = rng.uniform(-1.0, 1.0, size=(n, 3))
    t3 = srv.submit_solve(b3.copy(order="F"))
    srv.wait(tok)
    x3, _, v3 = srv.wait(t3)
    assert v3 == 2
    A2 = A + 0.05 * sps.eye(n)
    assert all(_berr(A2, x3[:, c], b3[:, c]) <= 1e-12 for c in range(3))
    srv.close()
    pastix.finalize(h)


def test_server_hbm_budget_admission():
    """HBM-budget admission: with a work budget of about 20 right-hand sides,
    two slots cannot both hold 16-column batches; merged solves are admitted
    as memory frees up (the most the slots held at once stays within the
    budget; whether a request had to wait for memory depends on how the two
    workers interleave, so the count is printed, not asserted), every result
    is right for its matrix version, and a request larger than the whole
    budget ends with PASTIX_ERR_OUTOFMEMORY instead of stalling the queue."""
    A = spm.laplacian7(*GRID, coef=(1.0, 1.0, 0.01))
    n = A.shape[0]
    slots = [_slot(A, spm.SpmFloat) for _ in range(2)]
    srv = pastix.Server([h for h, _ in slots], max_rhs=64)
    assert srv.memory()["budget"] > 0
    # one column's blocks, measured
    rng = np.random.default_rng(41)
    srv.wait(srv.submit_solve(rng.uniform(-1, 1, (n, 1))))
    one = srv.memory()["held"]
    assert one > 0
    srv.set_budget(20 * one)
    shifts = {0: 0.0}
    reqs = []
    for k in range(16):
        if k == 3:
            sh = 0.05
            reqs.append(("refactor", srv.submit_refactor(sh), sh, None))
        else:
            nrhs = int(rng.choice([4, 16]))
            b = rng.uniform(-1.0, 1.0, size=(n, nrhs))
            reqs.append(("solve", srv.submit_solve(b.copy(order="F")), None, b))
    big = rng.uniform(-1.0, 1.0, size=(n, 64))
    tbig = srv.submit_solve(big.copy(order="F"))
    for kind, t, sh, b in reqs:
        x, lat, ver = srv.wait(t)
        if kind == "refactor":
            shifts[ver] = sh
        else:
            Av = A + shifts[ver] * sps.eye(n)
            for c in range(b.shape[1]):
                assert _berr(Av, x[:, c], b[:, c]) <= 1e-12
    with pytest.raises(pastix.PastixError) as e:
        srv.wait(tbig)
    assert e.value.code == pastix.error.OUTOFMEMORY
    mem = srv.memory()
    print("server memory", mem, "one column", one)
    assert mem["peak"] <= 20 * one or mem["peak"] == one
    srv.close()
    for h, _ in slots:
        pastix.finalize(h)
