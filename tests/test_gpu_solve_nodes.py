"""Solve nodes (chains of supernodes solved as one diagonal block with its
whole inverse, PXA_IPARM_SOLVE_NODE_MAX): device solves with merged nodes
against the same factor solved supernode by supernode, and against the
oracle's supernodal solve (reference kernels/solve_ztrsmsp.c), LL^t and LU,
one and several right-hand sides, FP64 and FP32 factors."""
import numpy as np
import pytest

from pastix_amd import pastix, spm
from tests.helpers import analyse, oracle_factor

pytestmark = pytest.mark.gpu

LLT = pastix.factotype.LLT
LU = pastix.factotype.LU


def _berr(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


def _solve(A, sym_flag, facto, grid, node_max, b, prec=None):
    h, S, sym = analyse(A, sym_flag, facto, grid=grid)
    if prec is not None:
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    pastix.set_amd_iparm(h, pastix.amd_iparm.solve_node_max, node_max)
    info = pastix.solve_info(h)
    sym = pastix.symbol(h)
    pastix.task_numfact(h, S)
    x = b.copy(order="F")
    pastix.task_solve(h, S, x)
    pastix.finalize(h)
    return x, info, sym


@pytest.mark.parametrize("kind", ["llt", "lu"])
@pytest.mark.parametrize("nrhs", [1, 5])
def test_merged_nodes_match_supernode_solve_and_oracle(kind, nrhs):
    if kind == "llt":
        grid = (30, 30, 30)
        A = spm.laplacian7(*grid, diag=6.0)
        sym_flag, facto = True, LLT
    else:
        grid = (16, 16, 16)
        A = spm.convection_diffusion27(*grid, seed=5)
        sym_flag, facto = False, LU
    x0, b = spm.rhs_from_x(A, nrhs, seed=17)
    xm, im, sym = _solve(A, sym_flag, facto, grid, -1, b)
    xs, i0, _ = _solve(A, sym_flag, facto, grid, 0, b)
    print(kind, nrhs, "merged", im, "single", i0)
    assert im["merged_nodes"] >= 1 and im["solve_levels"] < i0["solve_levels"]
    assert i0["merged_nodes"] == 0 and i0["nodes"] == sym["cblknbr"] and i0["solve_levels"] == i0["levels"]
    scale = np.abs(xs).max()
    d = np.abs(xm - xs).max() / scale
    print("merged vs supernode solve", d)
    assert d <= 1e-12
    for r in range(nrhs):
        assert _berr(A, xm[:, r], b[:, r]) <= 1e-15
    # the oracle's supernodal solve on the oracle's factor
    O, bc, Lo, Uo = oracle_factor(A, sym, lu=(kind == "lu"))
    p = sym["permtab"]
    xp = np.zeros_like(b)
    xp[p, :] = b
    xo = O.solve(Lo, Uo, xp)[p, :]
    e = np.abs(xm - xo).max() / np.abs(xo).max()
    print("merged vs oracle", e)
    assert e <= 1e-11


def test_merged_nodes_fp32_factor():
    grid = (30, 30, 30)
    A = spm.laplacian7(*grid, diag=6.0)
    x0, b = spm.rhs_from_x(A, 2, seed=23)
    xm, im, _ = _solve(A, True, LLT, grid, -1, b, prec=spm.SpmFloat)
    xs, i0, _ = _solve(A, True, LLT, grid, 0, b, prec=spm.SpmFloat)
    d = np.abs(xm - xs).max() / np.abs(xs).max()
    print("fp32 merged vs supernode solve", d, im)
    assert im["merged_nodes"] >= 1
    assert d <= 1e-5
