#!/usr/bin/env python3
"""Headline benchmark: numfact GFLOP/s + solve p50/p99 on one MI355X.

Workload: 3D 7-point Laplacian on an 80^3 grid (n = 512,000), diagonal 6,
off-diagonals -1, grid nested dissection, full-rank LL^t in FP64 (the plain
method).  A step is one numerical factorisation with the filled coefficient
table already resident in HBM (pastix_amd_refactorize; the analysis and the
host-side fill are outside the timed region, as DPARM_FACT_TIME excludes
them).  value = DPARM_FACT_THFLOPS / median device factorisation time.

Also reported on the same matrix: solve p50/p99 over device-resident
right-hand sides, the FP32-factor + FP64-CG form to berr <= 1e-12, the
roofline of the dominant kernel (a Schur-update kernel, FP64 MFMA) and the oracle's CPU
factorisation timed on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "numfact GFLOP/s + solve p50/p99 ms; FP32+IR vs FP64 GPU at berr<=1e-12"
FP64_MFMA_PEAK_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def berr_inf(A, x, b):
    r = b - A @ x
    return float(np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()))


def setup(pastix, spm, A, grid, prec=None, resident=True):
    S = spm.Spm(A, symmetric=True)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    ip[pastix.iparm.ordering] = pastix.order.Personal
    ip[pastix.iparm.refinement] = pastix.refine.CG
    h = pastix.init(ip, dp)
    if prec is not None:
        pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
    if resident:
        pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
    o = pastix.orderGrid(*grid)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    return h, S


def cpu_baseline(pastix, spm, grid):
    """Oracle (plain-C sequential restatement of the reference) on a bounded grid."""
    import oracle
    A = spm.laplacian7(*grid, diag=6.0)
    h, S = setup(pastix, spm, A, grid, resident=False)
    sym = pastix.symbol(h)
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A, sym["permtab"])
    L, _, rc = O.fill(bc)
    t0 = time.perf_counter()
    O.potrf(L, np.sqrt(np.ldexp(1.0, -53)) * np.linalg.norm(bc["vals"]))
    t = time.perf_counter() - t0
    th = O.thflops()
    pastix.finalize(h)
    return th / t / 1e9, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, nargs=3, default=[80, 80, 80])
    ap.add_argument("--solves", type=int, default=100)
    ap.add_argument("--cpu-grid", type=int, nargs=3, default=[48, 48, 48])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--profile-only", action="store_true", help="only the timed factorisations (for rocprofv3)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only three device-to-device copies of 256 MiB (reference dispatch for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU)", file=sys.stderr)
        sys.exit(2)

    from pastix_amd import pastix, spm

    if args.calibrate_copy:
        # a plain copy of known size: each dispatch reads and writes 256 MiB
        n = 256 * 1024 * 1024 // 8
        src = pastix.DeviceArray(np.ones(n))
        dst = pastix.DeviceArray(np.zeros(n))
        for _ in range(3):
            dst.copy_from(src)
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "dispatches": 3}))
        return

    grid = tuple(args.grid)
    log(f"building the {grid} 7-point Laplacian")
    A = spm.laplacian7(*grid, diag=6.0)
    t0 = time.perf_counter()
    h, S = setup(pastix, spm, A, grid)
    t_an = time.perf_counter() - t0
    sym = pastix.symbol(h)
    fl = pastix.flops(h)
    thflops = float(fl[0])
    log(f"analysis {t_an:.2f}s cblk {sym['cblknbr']} levels {sym['nlevels']} nnz(L) {h.iparm[pastix.iparm.nnzeros]} "
        f"THFLOPS {thflops:.3e}")
    t0 = time.perf_counter()
    pastix.task_numfact(h, S)
    log(f"first numfact (fill + upload + factor) {time.perf_counter() - t0:.2f}s, device {h.dparm[pastix.dparm.fact_time]*1e3:.1f} ms")

    for _ in range(args.warmup):
        pastix.refactorize(h)
    times = []
    w0 = time.perf_counter()
    for k in range(args.steps):
        pastix.refactorize(h)
        times.append(float(h.dparm[pastix.dparm.fact_time]))
        log(f"step {k}: numfact {times[-1]*1e3:.2f} ms")
    wall = time.perf_counter() - w0
    tmed = float(np.median(times))
    gflops = thflops / tmed / 1e9
    if args.profile_only:
        print(json.dumps({"numfact_ms": tmed * 1e3, "gflops": gflops}))
        return

    # per-kernel split of one more factorisation (HIP events around every
    # update launch on the engine's stream)
    pastix.set_amd_iparm(h, pastix.amd_iparm.profile, 1)
    pastix.refactorize(h)
    ft = pastix.fact_times(h)
    kt = pastix.kernel_times(h)
    pastix.set_amd_iparm(h, pastix.amd_iparm.profile, 0)
    pastix.refactorize(h)
    uf = pastix.update_flops(h)
    kern = {}
    for cls, name, fl_ in (("update_direct", "k_update_direct<double,0>", uf["direct"]),
                           ("update_wgemm", "k_update_wgemm<double>", uf["multi"]),
                           ("update_acc", "k_update_acc<double,0>", uf["multi"])):
        t_, n_ = kt[cls]
        if n_ > 0:
            kern[name] = {"ms": t_ * 1e3, "launches": n_, "avg_launch_ms": t_ / n_ * 1e3,
                          "flops": fl_, "tflops": fl_ / t_ / 1e12}
    for cls, name in (("panel", "panel steps (k_potrf64 + k_panel_trsm + k_panel_upd)"),
                      ("update_gather", "k_update_gather<double,0>")):
        t_, n_ = kt[cls]
        if n_ > 0:
            kern[name] = {"ms": t_ * 1e3, "spans": n_}
    dom = max((k for k in kern if "flops" in kern[k]), key=lambda k: kern[k]["ms"])
    t_upd, n_upd = float(ft[3]), int(ft[6])
    upd_tf = float(fl[3]) / t_upd / 1e12
    log(f"split: panel {ft[1]*1e3:.2f} ms update {ft[3]*1e3:.2f} ms ({n_upd} launches) "
        f"update {upd_tf:.2f} TFLOP/s; " + "; ".join(f"{k} {v['ms']:.2f} ms" for k, v in kern.items()))

    # solves on device-resident right-hand sides
    x0, b = spm.rhs_from_x(A, 1, seed=2024)
    db = pastix.DeviceArray(b)
    stimes = []
    for k in range(args.solves + 2):
        db.upload(b)
        pastix.solve_device(h, db, A.shape[0], 1)
        if k >= 2:
            stimes.append(float(h.dparm[pastix.dparm.solv_time]))
    x = db.download()
    berr64 = berr_inf(A, x[:, 0], b[:, 0])
    p50, p99 = float(np.percentile(stimes, 50)) * 1e3, float(np.percentile(stimes, 99)) * 1e3
    log(f"solve p50 {p50:.3f} ms p99 {p99:.3f} ms berr {berr64:.2e}")
    nnzL = float(h.iparm[pastix.iparm.nnzeros])
    solve_gbs = (2 * (nnzL + A.shape[0]) * 8 + 2 * 2 * A.shape[0] * 8) / (p50 * 1e-3) / 1e9
    db.free()
    pastix.finalize(h)

    fp32 = None
    if not args.no_fp32:
        h32, S32 = setup(pastix, spm, A, grid, prec=spm.SpmFloat)
        pastix.task_numfact(h32, S32)
        t32 = []
        for _ in range(max(2, args.steps // 2)):
            pastix.refactorize(h32)
            t32.append(float(h32.dparm[pastix.dparm.fact_time]))
        xr = b.copy(order="F")
        r0 = time.perf_counter()
        pastix.task_solve(h32, S32, xr)
        pastix.task_refine(h32, S32, b, xr)
        t_ref = time.perf_counter() - r0
        b32 = berr_inf(A, xr[:, 0], b[:, 0])
        f32 = float(np.median(t32))
        fp32 = {"numfact_ms": f32 * 1e3, "numfact_gflops": thflops / f32 / 1e9,
                "solve_refine_ms": t_ref * 1e3, "refine_iters": int(h32.iparm[pastix.iparm.nbiter]),
                "berr": b32, "berr_target_met": bool(b32 <= 1e-12),
                "speedup_numfact_vs_fp64": tmed / f32,
                "speedup_numfact_plus_solve_vs_fp64": (tmed + p50 * 1e-3) / (f32 + t_ref)}
        log(f"fp32+IR: numfact {f32*1e3:.2f} ms solve+refine {t_ref*1e3:.1f} ms iters {fp32['refine_iters']} berr {b32:.2e}")
        pastix.finalize(h32)

    cpu = None
    if not args.no_cpu:
        cg = tuple(args.cpu_grid)
        cpu_gf, cpu_t = cpu_baseline(pastix, spm, cg)
        cpu = {"value": cpu_gf, "unit": "GFLOP/s", "cores": 1, "kind": "port",
               "sample": f"oracle sequential LL^t numfact of the {cg[0]}x{cg[1]}x{cg[2]} 7-point Laplacian "
                         f"(same stencil, smaller grid), {cpu_t:.1f} s of CPU work"}
        log(f"cpu baseline {cpu_gf:.2f} GFLOP/s in {cpu_t:.1f}s")

    out = {
        "metric": METRIC, "value": gflops, "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": tmed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"laplacian7_{grid[0]}x{grid[1]}x{grid[2]}_potrf_fp64", "n": int(A.shape[0]),
                   "nnz": int(A.nnz), "stencil": "7pt diag 6 offdiag -1", "ordering": "grid nested dissection",
                   "factorization": "LLt", "nnz_factor": int(nnzL), "thflops": thflops,
                   "cblknbr": int(sym["cblknbr"]), "levels": int(sym["nlevels"])},
        "gflops_pastix_2p30": thflops / tmed / 2 ** 30,
        "wall_ms_per_step": wall / args.steps * 1e3,
        "solve_p50_ms": p50, "solve_p99_ms": p99, "solve_gbs_p50": solve_gbs, "solve_berr": berr64,
        "fp32_ir": fp32,
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": kern[dom]["tflops"],
                     "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": kern[dom]["tflops"] / FP64_MFMA_PEAK_TFLOPS, "traffic": None,
                     "launches": kern[dom]["launches"], "avg_launch_ms": kern[dom]["avg_launch_ms"],
                     "flops_per_factorization": kern[dom]["flops"],
                     "update_family_tflops": upd_tf},
        "numfact_split_ms": {"panel": ft[1] * 1e3, "update": ft[3] * 1e3},
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
