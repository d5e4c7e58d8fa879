"""Level-schedule autotuner.

The level schedule makes two choices per dependency level: its multi-source
update tiles are summed in one pass (k_update_seq in FP64, k_update_acc in
FP32) or through the product workspace (k_update_wgemm + k_update_gather),
and its panels are updated
left-looking or right-looking.  The plan's rules pick them from counts
(average entries per tile, row tiles per level); this tuner times both
choices of every level on the device and keeps the faster one per level:

1. one-stream profiled factorisations with the rules' choices and with every
   choice flipped give each level's panel and update times for both;
2. per level, the faster panel form and the faster update form are kept;
3. the result is checked against the rules' schedule on the real (two-stream)
   factorisation, median of several runs, and written to the tuning table
   only if it is faster.

The table (pastix_amd/tuning/table.txt) holds one entry per analysed
partition and factorisation; the library applies a matching entry through
pastix_amd_load_tuning or the PASTIX_AMD_TUNING environment variable.

    python -m pastix_amd.autotune --workload laplacian7 --grid 80 80 80
    python -m pastix_amd.autotune --workload laplacian7 --grid 80 80 80 --precision float
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from pastix_amd import pastix, spm  # noqa: E402

TABLE = os.path.join(HERE, "tuning", "table.txt")


def instance(workload, grid, precision, seed=5):
    if workload == "laplacian7":
        A = spm.laplacian7(*grid, diag=6.0)
        facto = pastix.factotype.LLT
    elif workload == "laplacian7_aniso":
        A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
        facto = pastix.factotype.LLT
    elif workload == "convdiff27":
        A = spm.convection_diffusion27(*grid, seed=seed)
        facto = pastix.factotype.LU
    else:
        raise ValueError(f"unknown workload {workload}")
    S = spm.Spm(A, symmetric=facto == pastix.factotype.LLT)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = facto
    ip[pastix.iparm.ordering] = pastix.order.Personal
    if facto == pastix.factotype.LU:
        dp[pastix.dparm.epsilon_magn_ctrl] = np.ldexp(1.0, -52)
    h = pastix.init(ip, dp)
    if precision == "float":
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, spm.SpmFloat)
    pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
    o = pastix.orderGrid(*grid)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    return h, S


def profiled_levels(h, S, reps=2):
    """Per-level (panel, update) seconds, one stream, the minimum of reps."""
    pastix.task_numfact(h, S)
    pastix.set_amd_iparm(h, pastix.amd_iparm.profile, 1)
    pastix.set_amd_iparm(h, pastix.amd_iparm.overlap, 0)
    best = None
    for _ in range(reps):
        pastix.refactorize(h)
        t = pastix.level_times(h)
        best = t if best is None else np.minimum(best, t)
    pastix.set_amd_iparm(h, pastix.amd_iparm.profile, 0)
    pastix.set_amd_iparm(h, pastix.amd_iparm.overlap, 1)
    return best


def timed(h, S, reps):
    """Median device time of the real (two-stream) factorisation."""
    pastix.task_numfact(h, S)
    pastix.refactorize(h)
    ts = []
    for _ in range(reps):
        pastix.refactorize(h)
        ts.append(float(h.dparm[pastix.dparm.fact_time]))
    return float(np.median(ts))


def tune(workload, grid, precision, reps, table, summary):
    h, S = instance(workload, grid, precision)
    info = pastix.level_info(h)
    nl = len(info)
    auto_sp, auto_pr = info[:, 2].astype(np.int8), info[:, 3].astype(np.int8)
    multi = info[:, 1] > 0
    print(f"[autotune] {workload} {grid} {precision}: {nl} levels, {int(multi.sum())} with multi-source tiles",
          flush=True)
    ta = profiled_levels(h, S)
    flip_sp = np.where(multi, 1 - auto_sp, -1).astype(np.int8)
    flip_pr = (1 - auto_pr).astype(np.int8)
    pastix.set_level_knobs(h, flip_sp, flip_pr)
    tb = profiled_levels(h, S)
    sp = np.where(multi & (tb[:, 1] < ta[:, 1]), flip_sp, auto_sp).astype(np.int8)
    pr = np.where(tb[:, 0] < ta[:, 0], flip_pr, auto_pr).astype(np.int8)
    print(f"[autotune] one stream: rules {ta.sum()*1e3:.2f} ms, flipped {tb.sum()*1e3:.2f} ms, "
          f"per-level best {np.minimum(ta, tb).sum()*1e3:.2f} ms", flush=True)
    pastix.set_level_knobs(h, np.full(nl, -1, np.int8), np.full(nl, -1, np.int8))
    t_rules = timed(h, S, reps)
    pastix.set_level_knobs(h, sp, pr)
    t_tuned = timed(h, S, reps)
    keep = t_tuned < t_rules
    print(f"[autotune] two streams: rules {t_rules*1e3:.2f} ms, tuned {t_tuned*1e3:.2f} ms -> "
          f"{'tuned' if keep else 'rules'}", flush=True)
    if not keep:
        pastix.set_level_knobs(h, np.full(nl, -1, np.int8), np.full(nl, -1, np.int8))
    os.makedirs(os.path.dirname(os.path.abspath(table)), exist_ok=True)
    pastix.save_tuning(h, table)
    if summary:
        with open(summary, "w") as f:
            f.write(f"# Level-schedule autotune: {workload} {grid[0]}x{grid[1]}x{grid[2]}, {precision} factor\n\n")
            f.write("Command: `python -m pastix_amd.autotune --workload {} --grid {} {} {} --precision {} "
                    "--reps {}`\n\n".format(workload, *grid, precision, reps))
            f.write(f"Numfact device time, two streams, median of {reps}: rules {t_rules*1e3:.2f} ms, "
                    f"tuned {t_tuned*1e3:.2f} ms; table entry: {'tuned' if keep else 'rules'}.\n\n")
            f.write("One-stream per-level times (ms), rules' choice (sp = single-pass updates, "
                    "pr = right-looking panels) against the flipped one:\n\n")
            f.write("| level | cblks | multi tiles | sp rules | upd rules | upd flipped | pr rules | "
                    "panel rules | panel flipped | chosen sp | chosen pr |\n")
            f.write("|---|---|---|---|---|---|---|---|---|---|---|\n")
            for lv in range(nl):
                f.write(f"| {lv} | {info[lv, 0]} | {info[lv, 1]} | {auto_sp[lv]} | {ta[lv, 1]*1e3:.3f} | "
                        f"{tb[lv, 1]*1e3:.3f} | {auto_pr[lv]} | {ta[lv, 0]*1e3:.3f} | {tb[lv, 0]*1e3:.3f} | "
                        f"{sp[lv]} | {pr[lv]} |\n")
    pastix.finalize(h)
    return t_rules, t_tuned, keep


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", default="laplacian7", choices=["laplacian7", "laplacian7_aniso", "convdiff27"])
    ap.add_argument("--grid", type=int, nargs=3, default=[80, 80, 80])
    ap.add_argument("--precision", default="double", choices=["double", "float"])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--table", default=TABLE)
    ap.add_argument("--summary", default=None, help="markdown summary of the per-level times")
    a = ap.parse_args()
    tune(a.workload, tuple(a.grid), a.precision, a.reps, a.table, a.summary)


if __name__ == "__main__":
    main()
