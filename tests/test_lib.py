"""The C-ABI library: loads, exports every declared entry point, keeps the
reference parameter defaults and the task state machine.  No GPU needed."""
import ctypes as ct
import os
import re

import numpy as np

from pastix_amd import pastix, spm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in ("pastix.h", "pastix_amd.h", os.path.join("pastix", "order.h")):
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(?:int|void)\s+(\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert "pastix_task_numfact" in names and "pastixOrderGrid" in names and len(names) > 30
    lib = ct.CDLL(pastix.LIBPATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_init_param_defaults_match_reference():
    # reference common/api.c:363-538
    ip, dp = pastix.initParam()
    assert ip[pastix.iparm.ordering] == pastix.order.Scotch
    assert ip[pastix.iparm.min_blocksize] == 160 and ip[pastix.iparm.max_blocksize] == 320
    assert ip[pastix.iparm.factorization] == pastix.factotype.LU
    assert ip[pastix.iparm.scheduler] == pastix.scheduler.Static
    assert ip[pastix.iparm.refinement] == pastix.refine.GMRES
    assert ip[pastix.iparm.itermax] == 250 and ip[pastix.iparm.gmres_im] == 25
    assert ip[pastix.iparm.gpu_memory_percentage] == 95
    assert ip[pastix.iparm.tasks2d_level] == -1 and ip[pastix.iparm.tasks2d_width] == 160
    assert ip[pastix.iparm.amalgamation_lvlcblk] == 5 and ip[pastix.iparm.dof_nbr] == 1
    assert dp[pastix.dparm.epsilon_refinement] == -1.0 and dp[pastix.dparm.epsilon_magn_ctrl] == 0.0
    assert dp[pastix.dparm.a_norm] == -1.0 and dp[pastix.dparm.compress_tolerance] == 0.01
    assert pastix.iparm.size == 71 and pastix.dparm.size == 18


def test_step_order_is_enforced():
    A = spm.laplacian7(4, 4, 4, diag=6.0)
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    h = pastix.init(ip, dp)
    for call in (lambda: pastix.task_numfact(h, S), lambda: pastix.subtask_symbfact(h),
                 lambda: pastix.subtask_blend(h)):
        try:
            call()
            raise AssertionError("accepted out of order")
        except pastix.PastixError as e:
            assert e.code == pastix.error.BADPARAMETER
    x = np.zeros((64, 1), order="F")
    try:
        pastix.task_solve(h, S, x)
        raise AssertionError("solve accepted before numfact")
    except pastix.PastixError as e:
        assert e.code == pastix.error.BADPARAMETER
    pastix.task_analyze(h, S)  # analysis alone runs on the host
    assert ip[pastix.iparm.nnzeros] > 0 and dp[pastix.dparm.fact_thflops] > 0
    pastix.finalize(h)


def test_applyorder_round_trip():
    A = spm.laplacian7(5, 4, 3, diag=6.0)
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    h = pastix.init(ip, dp)
    pastix.task_analyze(h, S)
    sym = pastix.symbol(h)
    b = np.asfortranarray(np.random.default_rng(0).standard_normal((60, 3)))
    b0 = b.copy()
    pastix.subtask_applyorder(h, pastix.dir.Forward, b)
    assert np.array_equal(b[sym["permtab"], :], b0)
    pastix.subtask_applyorder(h, pastix.dir.Backward, b)
    assert np.array_equal(b, b0)
    pastix.finalize(h)


def test_orsirr_fixture_loads_and_analyses():
    from tests.helpers import read_rua
    A = read_rua(os.path.join(ROOT, "tests", "golden", "orsirr.rua"))
    assert A.shape == (1030, 1030) and A.nnz == 6858
    S = spm.Spm(A, symmetric=False)
    ip, dp = pastix.initParam()
    h = pastix.init(ip, dp)
    pastix.task_analyze(h, S)
    sym = pastix.symbol(h)
    assert sym["n"] == 1030 and sym["cblknbr"] > 1
    pastix.finalize(h)
