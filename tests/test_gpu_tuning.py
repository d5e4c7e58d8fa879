"""Every schedule choice the tuner can make gives the same factor: with each
level's update form (one pass / product workspace) and panel direction
(left / right-looking) flipped from the plan's rules, the LL^t and LU factors
match the oracle at the plain method's 1e-12 relative Frobenius."""
import numpy as np
import pytest

from pastix_amd import pastix, spm
from tests.helpers import analyse, oracle_factor, rel_frob, lower_mask_panel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("facto", [pastix.factotype.LLT, pastix.factotype.LU])
def test_flipped_level_schedule_factor_parity(facto):
    grid = (18, 17, 16)
    if facto == pastix.factotype.LLT:
        A = spm.laplacian7(*grid, diag=6.0)
    else:
        A = spm.convection_diffusion27(*grid, seed=3)
    h, S, sym = analyse(A, facto == pastix.factotype.LLT, facto, grid=grid, min_bs=16, max_bs=96)
    info = pastix.level_info(h)
    multi = info[:, 1] > 0
    sp = np.where(multi, 1 - info[:, 2], -1).astype(np.int8)
    pr = (1 - info[:, 3]).astype(np.int8)
    pastix.set_level_knobs(h, sp, pr)
    pastix.task_numfact(h, S)
    lu = facto == pastix.factotype.LU
    Lg, Ug = pastix.coeftab(h, lu=lu)
    O, bc, Lo, Uo = oracle_factor(A, sym, lu=lu)
    if lu:
        err = max(rel_frob(Lg, Lo), rel_frob(Ug, Uo))
    else:
        err = rel_frob(lower_mask_panel(sym, Lg), lower_mask_panel(sym, Lo))
    print("flipped schedule", facto, "levels", len(info), "multi-source levels", int(multi.sum()), "rel frob", err)
    assert err <= 1e-12
    pastix.finalize(h)
