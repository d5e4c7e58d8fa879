"""Shared helpers for the tests: run the engine's analysis through the C-ABI
and build the oracle on the same structure."""
import numpy as np

from pastix_amd import pastix, spm
import oracle


def analyse(A, symmetric, facto, grid=None, min_bs=160, max_bs=320, level_order=False):
    """Init + analysis; returns (handle, Spm, symbol dict, grid order or None)."""
    S = spm.Spm(A, symmetric=symmetric)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = facto
    ip[pastix.iparm.min_blocksize] = min_bs
    ip[pastix.iparm.max_blocksize] = max_bs
    h = pastix.init(ip, dp)
    if level_order:
        pastix.set_amd_iparm(h, pastix.amd_iparm.level_order, 1)
    o = None
    if grid is not None:
        o = pastix.orderGrid(*grid)
        h.order = o
        ip[pastix.iparm.ordering] = pastix.order.Personal
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    return h, S, pastix.symbol(h)


def oracle_factor(A, sym, lu=False, crit=None):
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A, sym["permtab"])
    L, U, rc = O.fill(bc, lu=lu)
    assert rc == 0, "fill dropped entries outside the symbol"
    if crit is None:
        crit = np.sqrt(np.ldexp(1.0, -53)) * np.linalg.norm(bc["vals"])
    if lu:
        O.getrf(L, U, crit)
    else:
        O.potrf(L, crit)
    return O, bc, L, U


def rel_frob(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def lower_mask_panel(sym, L):
    """Zero the strictly-upper part of every diagonal block (never read for LL^t)."""
    L = L.copy()
    for c in range(sym["cblknbr"]):
        fc, lc, st = sym["cblk_fcol"][c], sym["cblk_fcol"][c + 1], sym["cblk_stride"][c]
        P = L[sym["cblk_coef"][c]:sym["cblk_coef"][c + 1]].reshape((lc - fc, st)).T
        K = lc - fc
        P[:K, :K] = np.tril(P[:K, :K])
        L[sym["cblk_coef"][c]:sym["cblk_coef"][c + 1]] = P.T.ravel()
    return L
