"""Level-schedule tuning tables (host side): per-level choices override the
plan's rules, survive a save/load round trip keyed by the analysed partition,
and an entry for another partition is not applied."""
import os

import numpy as np

from pastix_amd import pastix, spm
from tests.helpers import analyse


def test_knobs_table_round_trip(tmp_path):
    grid = (16, 15, 14)
    A = spm.laplacian7(*grid, diag=6.0)
    h, S, sym = analyse(A, True, pastix.factotype.LLT, grid=grid, min_bs=8, max_bs=32)
    info = pastix.level_info(h)
    nl = len(info)
    assert nl == sym["nlevels"]
    multi = info[:, 1] > 0
    assert multi.any()
    # flip every choice the rules made
    sp = np.where(multi, 1 - info[:, 2], -1).astype(np.int8)
    pr = (1 - info[:, 3]).astype(np.int8)
    pastix.set_level_knobs(h, sp, pr)
    got = pastix.level_info(h)
    assert np.array_equal(got[multi, 2], sp[multi])
    assert np.array_equal(got[:, 3], pr)
    # the plan's update-flop partition still covers the reference count
    uf = pastix.update_flops(h)
    fl = pastix.flops(h)
    assert uf["direct"] + uf["multi"] + uf["single"] == __import__("pytest").approx(fl[3], rel=1e-12)
    table = os.path.join(tmp_path, "table.txt")
    pastix.save_tuning(h, table)
    pastix.save_tuning(h, table)  # replaces its own entry
    with open(table) as f:
        assert sum(1 for l in f if l.startswith("entry ")) == 1
    # a fresh analysis of the same problem takes the table's choices
    h2, S2, _ = analyse(A, True, pastix.factotype.LLT, grid=grid, min_bs=8, max_bs=32)
    assert not np.array_equal(pastix.level_info(h2)[:, 3], pr)
    assert pastix.load_tuning(h2, table)
    assert np.array_equal(pastix.level_info(h2), got)
    # another partition does not match the entry
    h3, S3, _ = analyse(A, True, pastix.factotype.LLT, grid=grid, min_bs=16, max_bs=48)
    before = pastix.level_info(h3)
    assert not pastix.load_tuning(h3, table)
    assert np.array_equal(pastix.level_info(h3), before)
    for x in (h, h2, h3):
        pastix.finalize(x)
