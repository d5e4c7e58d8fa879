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


OVERLAP = 1  # second stream for the updates beside the next panels (--overlap)
NODE_MAX = -1  # widest solve node (--solve-node-max; -1 the library default, 0 one node per cblk)
# level-schedule choices written by pastix_amd/autotune.py (--no-tuning: the plan's rules)
TUNING = os.path.join(ROOT, "pastix_amd", "tuning", "table.txt")
TUNED = {}  # instance label -> whether a table entry applied


def apply_tuning(pastix, h, label):
    TUNED[label] = bool(TUNING and os.path.exists(TUNING) and pastix.load_tuning(h, TUNING))


def setup(pastix, spm, A, grid, prec=None, resident=True, label=None):
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
    pastix.set_amd_iparm(h, pastix.amd_iparm.overlap, OVERLAP)
    pastix.set_amd_iparm(h, pastix.amd_iparm.solve_node_max, NODE_MAX)
    o = pastix.orderGrid(*grid)
    h.order = o
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    if label:
        apply_tuning(pastix, h, label)
    return h, S


def traffic_per_launch(kernel, grid):
    """HBM bytes per launch of a kernel on this workload, from the committed
    rocprofv3 counter summary (lines 'traffic_bytes_per_dispatch <kernel>
    <bytes>'), or None when that summary does not cover it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"laplacian7_{grid[0]}cube_potrf_fp64_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    key = kernel.replace(" ", "")
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_dispatch" and parts[1] == key:
                return float(parts[2])
    return None


def solve_traffic(grid):
    """HBM bytes of one 1-RHS device solve (all its kernels) from the
    committed counter summary (line 'traffic_bytes_per_solve <bytes>'), or
    None when there is none for this workload."""
    path = os.path.join(ROOT, "profiles", f"laplacian7_{grid[0]}cube_solve_fp64_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 2 and parts[0] == "traffic_bytes_per_solve":
                return float(parts[1])
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pastix, spm, grid, threads):
    """The oracle's restatement of the reference static scheduler
    (sopalin/sequential_zpotrf.c:71-120: per-thread task lists, ctrbcnt
    waits, per-cblk locks) on `threads` host threads with a sequential CBLAS
    per worker, on the same matrix as the GPU headline."""
    import oracle
    A = spm.laplacian7(*grid, diag=6.0)
    h, S = setup(pastix, spm, A, grid, resident=False)
    sym = pastix.symbol(h)
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A, sym["permtab"])
    L, _, rc = O.fill(bc)
    blas = oracle.use_blas()
    thr = np.sqrt(np.ldexp(1.0, -53)) * np.linalg.norm(bc["vals"])
    t0 = time.perf_counter()
    rc, _ = O.factor_static(L, None, thr, threads)
    t = time.perf_counter() - t0
    th = O.thflops()
    pastix.finalize(h)
    return th / t / 1e9, t, blas


def cpu_sequential_c1(pastix, spm, grid=(20, 20, 20)):
    """C1: the oracle's sequential scheduler (sopalin/sequential_zpotrf.c:34-68),
    one thread, plain loops, with its backward error."""
    import oracle
    A = spm.laplacian7(*grid, diag=6.0)
    h, S = setup(pastix, spm, A, grid, resident=False)
    sym = pastix.symbol(h)
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A, sym["permtab"])
    L, _, rc = O.fill(bc)
    thr = np.sqrt(np.ldexp(1.0, -53)) * np.linalg.norm(bc["vals"])
    t0 = time.perf_counter()
    O.potrf(L, thr)
    t = time.perf_counter() - t0
    x0, b = spm.rhs_from_x(A, 1, seed=0)
    p = sym["permtab"]
    bp = np.zeros(A.shape[0])
    bp[p] = b[:, 0]
    s0 = time.perf_counter()
    xp = O.solve(L, None, bp.copy())
    ts = time.perf_counter() - s0
    x = xp[p] if xp.ndim == 1 else xp[p, 0]
    out = {"workload": f"laplacian7_{grid[0]}x{grid[1]}x{grid[2]}_potrf_fp64", "scheduler": "sequential",
           "cores": 1, "numfact_s": t, "gflops": O.thflops() / t / 1e9, "solve_ms": ts * 1e3,
           "berr": berr_inf(A, x, b[:, 0])}
    pastix.finalize(h)
    return out


def leg_getrf(pastix, spm, grid, seed=5):
    """Nonsymmetric 27-point convection-diffusion (symmetric pattern), GETRF
    with static pivoting at eps*||A||: FP64 factor vs FP32 factor + FP64
    GMRES to berr <= 1e-12."""
    A = spm.convection_diffusion27(*grid, seed=seed)
    x0, b = spm.rhs_from_x(A, 1, seed=seed + 1)
    out = {"workload": f"convdiff27_{grid[0]}x{grid[1]}x{grid[2]}_getrf", "n": int(A.shape[0])}
    for label, prec in (("fp64", None), ("fp32_gmres", spm.SpmFloat)):
        S = spm.Spm(A, symmetric=False)
        ip, dp = pastix.initParam()
        ip[pastix.iparm.factorization] = pastix.factotype.LU
        ip[pastix.iparm.ordering] = pastix.order.Personal
        ip[pastix.iparm.refinement] = pastix.refine.GMRES
        dp[pastix.dparm.epsilon_magn_ctrl] = np.ldexp(1.0, -52)
        h = pastix.init(ip, dp)
        if prec is not None:
            pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
        pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
        o = pastix.orderGrid(*grid)
        h.order = o
        pastix.subtask_order(h, S, o)
        pastix.subtask_symbfact(h)
        pastix.subtask_reordering(h)
        pastix.subtask_blend(h)
        apply_tuning(pastix, h, "getrf_" + label)
        pastix.task_numfact(h, S)
        times = []
        for _ in range(3):
            pastix.refactorize(h)
            times.append(float(h.dparm[pastix.dparm.fact_time]))
        th = float(pastix.flops(h)[0])
        # right-hand side resident on the device: FP64 plain solve in place;
        # FP32 factor: solve + GMRES on the device (wall clock, several runs)
        db = pastix.DeviceArray(b)
        dx = pastix.DeviceArray(np.zeros_like(b))
        ts = []
        for _ in range(3):
            if prec is None:
                dx.copy_from(db)
                r0 = time.perf_counter()
                pastix.solve_device(h, dx, A.shape[0], 1)
            else:
                r0 = time.perf_counter()
                pastix.solve_refine_device(h, db, dx, A.shape[0], 1)
            ts.append(time.perf_counter() - r0)
        x = dx.download()
        db.free()
        dx.free()
        tm = float(np.median(times))
        t_solve = float(np.median(ts))
        out[label] = {"numfact_ms": tm * 1e3, "numfact_gflops": th / tm / 1e9,
                      "solve_ms" if prec is None else "solve_refine_ms": t_solve * 1e3,
                      "refine_iters": int(h.iparm[pastix.iparm.nbiter]) if prec is not None else 0,
                      "berr": berr_inf(A, x[:, 0], b[:, 0]),
                      "static_pivots": int(h.iparm[pastix.iparm.static_pivoting])}
        pastix.finalize(h)
    f64, f32 = out["fp64"], out["fp32_gmres"]
    out["speedup_numfact_fp32_vs_fp64"] = f64["numfact_ms"] / f32["numfact_ms"]
    out["speedup_numfact_plus_solve_vs_fp64"] = ((f64["numfact_ms"] + f64["solve_ms"]) /
                                                 (f32["numfact_ms"] + f32["solve_refine_ms"]))
    return out


def leg_batch(pastix, spm, nsys, seed=7, sample_alone=16):
    """nsys SPD systems, 2D 5-point Laplacians on k x k grids (k uniform in
    [16, 128]) plus a uniform(0, 1) diagonal shift, 1-8 RHS each: one
    block-diagonal FP32 factor (graph replay) + batched FP64 CG, against plain
    FP64 instances solved one after the other (timed on a sample)."""
    import scipy.sparse as sps
    rng = np.random.default_rng(seed)
    ks = rng.integers(16, 129, size=nsys)
    mats = [spm.laplacian5_2d(int(k)) + rng.uniform(0.0, 1.0) * sps.eye(int(k) * int(k)) for k in ks]
    nrhs = rng.integers(1, 9, size=nsys)
    bs = [rng.uniform(-1.0, 1.0, size=(int(k) * int(k), int(r))) for k, r in zip(ks, nrhs)]
    S = [spm.Spm(A, symmetric=True) for A in mats]
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = pastix.factotype.LLT
    ip[pastix.iparm.refinement] = pastix.refine.CG
    h = pastix.init(ip, dp)
    pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, spm.SpmFloat)
    pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
    pastix.set_amd_iparm(h, pastix.amd_iparm.graph, 1)
    t0 = time.perf_counter()
    pastix.batch_analyze(h, S, [pastix.orderGrid(int(k), int(k), 1) for k in ks])
    t_an = time.perf_counter() - t0
    pastix.batch_numfact(h, S)
    ft = []
    for _ in range(3):
        pastix.refactorize(h)
        ft.append(float(h.dparm[pastix.dparm.fact_time]))
    pastix.batch_solve_refine(h, bs)  # first call sizes the device work blocks
    r0 = time.perf_counter()
    xs = pastix.batch_solve_refine(h, bs)
    t_sr = time.perf_counter() - r0
    worst = max(berr_inf(A, x[:, c], b[:, c]) for A, b, x in zip(mats, bs, xs) for c in range(b.shape[1]))
    iters = int(h.iparm[pastix.iparm.nbiter])
    pastix.finalize(h)
    # plain FP64, one system at a time, on a sample
    samp = list(range(min(sample_alone, nsys)))
    a0 = time.perf_counter()
    t_alone_fs = 0.0  # numfact + solve of the plain loop, without its analyses
    for i in samp:
        ip1, dp1 = pastix.initParam()
        ip1[pastix.iparm.factorization] = pastix.factotype.LLT
        ip1[pastix.iparm.ordering] = pastix.order.Personal
        h1 = pastix.init(ip1, dp1)
        o = pastix.orderGrid(int(ks[i]), int(ks[i]), 1)
        h1.order = o
        pastix.subtask_order(h1, S[i], o)
        pastix.subtask_symbfact(h1)
        pastix.subtask_reordering(h1)
        pastix.subtask_blend(h1)
        f0 = time.perf_counter()
        pastix.task_numfact(h1, S[i])
        xp = bs[i].copy(order="F")
        pastix.task_solve(h1, S[i], xp)
        t_alone_fs += time.perf_counter() - f0
        pastix.finalize(h1)
    t_alone = (time.perf_counter() - a0) / len(samp) * nsys
    t_alone_fs = t_alone_fs / len(samp) * nsys
    return {"workload": f"batch_{nsys}_laplace2d_k16-128", "systems": nsys, "n_total": int(sum(k * k for k in ks)),
            "rhs_total": int(nrhs.sum()), "analysis_s": t_an, "numfact_ms": float(np.median(ft)) * 1e3,
            "solve_refine_ms_wall": t_sr * 1e3, "refine_iters_max": iters, "berr_max": worst,
            "berr_target_met": bool(worst <= 1e-12),
            "plain_fp64_loop_s_wall_est": t_alone,
            "plain_fp64_loop_sample": f"{len(samp)} systems analysed, factored and solved one at a time, scaled to {nsys}",
            "speedup_vs_plain_loop": t_alone / ((float(np.median(ft)) + t_sr) + t_an),
            "plain_fp64_loop_numfact_solve_s_est": t_alone_fs,
            "speedup_numfact_solve_excluding_analysis": t_alone_fs / (float(np.median(ft)) + t_sr)}


def leg_serving(pastix, spm, grid, nreq, seed=11, slots=2, max_rhs=64, load=0.6):
    """Anisotropic 7-point Laplacian (1, 1, 0.01) on FP32 slots + batched FP64
    CG.  First the capacity: a burst of solves (nrhs in {1, 4, 16, 64})
    submitted at once, right-hand sides per second when saturated.  Then a
    seeded Poisson stream at `load` x that capacity: 90 % solves, 10 %
    refactorisations A + shift*I (shift uniform(0, 0.1)); latency p50/p99 and
    throughput, beside the plain FP64 path solving the same kind of requests
    one at a time."""
    A = spm.laplacian7(*grid, coef=(1.0, 1.0, 0.01))
    n = A.shape[0]

    def make(prec):
        S = spm.Spm(A, symmetric=True)
        ip, dp = pastix.initParam()
        ip[pastix.iparm.factorization] = pastix.factotype.LLT
        ip[pastix.iparm.ordering] = pastix.order.Personal
        ip[pastix.iparm.refinement] = pastix.refine.CG
        # CG stops on ||r||_2 / ||b||_2; 1e-13 keeps berr_inf <= 1e-12 on
        # every column at n = 884,736 (1e-12 left columns at 1.35e-12)
        dp[pastix.dparm.epsilon_refinement] = 1e-13
        h = pastix.init(ip, dp)
        if prec is not None:
            pastix.set_amd_iparm(h, pastix.amd_iparm.factor_precision, prec)
        pastix.set_amd_iparm(h, pastix.amd_iparm.resident_input, 1)
        o = pastix.orderGrid(*grid)
        h.order = o
        pastix.subtask_order(h, S, o)
        pastix.subtask_symbfact(h)
        pastix.subtask_reordering(h)
        pastix.subtask_blend(h)
        pastix.task_numfact(h, S)
        return h, S

    hs = [make(spm.SpmFloat) for _ in range(slots)]
    per_slot = pastix.device_bytes(hs[0][0])
    rng = np.random.default_rng(seed)
    choices = np.array([1, 4, 16, 64])
    srv = pastix.Server([h for h, _ in hs], max_rhs=max_rhs)
    # capacity: a saturated burst (the first batch also warms the slots up)
    burst = [int(k) for k in rng.choice(choices, size=24)]
    bufs = [rng.uniform(-1.0, 1.0, size=(n, k)) for k in burst]
    c0 = time.perf_counter()
    tk = [srv.submit_solve(b) for b in bufs]
    for t in tk:
        srv.wait(t)
    cap_rhs = sum(burst) / (time.perf_counter() - c0)
    mean_rhs = float(choices.mean())
    rate = load * cap_rhs / mean_rhs / 0.9  # requests/s, 90 % of them solves
    gaps = rng.exponential(1.0 / rate, size=nreq)
    kinds = rng.uniform(size=nreq) < 0.9
    nrhs = rng.choice(choices, size=nreq)
    tickets = []
    w0 = time.perf_counter()
    t_next = w0
    for i in range(nreq):
        t_next += gaps[i]
        dt = t_next - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        if kinds[i]:
            b = rng.uniform(-1.0, 1.0, size=(n, int(nrhs[i])))
            # column 0 kept to check the result's backward error afterwards
            tickets.append(("solve", srv.submit_solve(b), int(nrhs[i]), b[:, 0].copy()))
        else:
            sh = float(rng.uniform(0.0, 0.1))
            tickets.append(("refactor", srv.submit_refactor(sh), 0, sh))
    lat_s, lat_r, rhs_done = [], [], 0
    shifts, checks = {0: 0.0}, []
    for kind, t, r, extra in tickets:
        x, lat, ver = srv.wait(t)
        if kind == "solve":
            lat_s.append(lat)
            rhs_done += r
            checks.append((ver, extra, x[:, 0].copy()))
        else:
            lat_r.append(lat)
            shifts[ver] = extra
    wall = time.perf_counter() - w0
    # backward error of column 0 of every solve, against the matrix version
    # the request was bound to
    import scipy.sparse as sps
    Ar = A.tocsr()
    anorm = abs(Ar).sum(axis=1).max()
    berr_max = 0.0
    for ver, b0, x0 in checks:
        r = b0 - (Ar @ x0 + shifts[ver] * x0)
        berr_max = max(berr_max, float(np.abs(r).max() / ((anorm + abs(shifts[ver])) * np.abs(x0).max() +
                                                         np.abs(b0).max())))
    st = srv.stats()
    srv.close()
    # plain FP64 path: the same kind of solves one at a time on one FP64 instance
    hp, Sp = make(None)
    nsamp = 8
    p0 = time.perf_counter()
    prhs = 0
    for i in range(nsamp):
        k = int(rng.choice(choices))
        xb = rng.uniform(-1.0, 1.0, size=(n, k)).copy(order="F")
        pastix.task_solve(hp, Sp, xb)
        prhs += k
    t_plain = time.perf_counter() - p0
    for h, _ in hs + [(hp, Sp)]:
        pastix.finalize(h)
    return {"workload": f"serving_laplacian7_aniso_{grid[0]}x{grid[1]}x{grid[2]}", "n": int(n), "requests": nreq,
            "slots": slots, "slot_device_gib": per_slot / 2 ** 30, "max_rhs_per_batch": max_rhs,
            "capacity_rhs_per_s": cap_rhs, "offered_load": load, "poisson_rate_per_s": rate,
            "solves": len(lat_s), "refactorizations": len(lat_r),
            "solve_latency_p50_ms": float(np.percentile(lat_s, 50)) * 1e3 if lat_s else None,
            "solve_latency_p99_ms": float(np.percentile(lat_s, 99)) * 1e3 if lat_s else None,
            "refactor_latency_p50_ms": float(np.percentile(lat_r, 50)) * 1e3 if lat_r else None,
            "solves_per_s": len(lat_s) / wall, "rhs_per_s": rhs_done / wall, "batches": st["batches"],
            "berr_max": berr_max, "berr_target_met": bool(berr_max <= 1e-12),
            "berr_checked": "column 0 of every solve, against the matrix version it was bound to",
            "plain_fp64_rhs_per_s": prhs / t_plain,
            "plain_fp64_sample": f"{nsamp} requests solved one at a time with pastix_task_solve (host arrays, no refinement)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, nargs=3, default=[80, 80, 80])
    ap.add_argument("--solves", type=int, default=100)
    ap.add_argument("--cpu-grid", type=int, nargs=3, default=None,
                    help="grid of the CPU baseline (default: the headline grid)")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="host threads of the CPU baseline (default: min(16, cores))")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-tuning", action="store_true", help="the plan's rules, not the tuning table")
    ap.add_argument("--profile-only", action="store_true", help="only the timed factorisations (for rocprofv3)")
    ap.add_argument("--solve-only", action="store_true",
                    help="the timed factorisations, then --solves device solves (for rocprofv3)")
    ap.add_argument("--legs", default="getrf,batch,serving",
                    help="extra workloads reported beside the headline: getrf, batch, serving (comma list or none)")
    ap.add_argument("--getrf-grid", type=int, nargs=3, default=[64, 64, 64])
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--serving-grid", type=int, nargs=3, default=[96, 96, 96])
    ap.add_argument("--serving-requests", type=int, default=500)
    ap.add_argument("--overlap", type=int, default=1,
                    help="1: updates on a second stream beside the next level's panels; 0: one stream")
    ap.add_argument("--solve-node-max", type=int, default=-1,
                    help="widest solve node in columns (-1: the library default; 0: one node per cblk)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only three device-to-device copies of 256 MiB (reference dispatch for HBM counters)")
    args = ap.parse_args()
    global OVERLAP, TUNING, NODE_MAX
    OVERLAP = args.overlap
    NODE_MAX = args.solve_node_max
    if args.no_tuning:
        TUNING = None
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
    h, S = setup(pastix, spm, A, grid, label="fp64")
    t_an = time.perf_counter() - t0
    sym = pastix.symbol(h)
    fl = pastix.flops(h)
    thflops = float(fl[0])
    log(f"analysis {t_an:.2f}s cblk {sym['cblknbr']} levels {sym['nlevels']} nnz(L) {h.iparm[pastix.iparm.nnzeros]} "
        f"THFLOPS {thflops:.3e}")
    t0 = time.perf_counter()
    pastix.task_numfact(h, S)
    t_first = time.perf_counter() - t0
    log(f"first numfact (fill + upload + factor) {t_first:.2f}s, device {h.dparm[pastix.dparm.fact_time]*1e3:.1f} ms")

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
    if args.solve_only:
        # only device solves of 1 right-hand side after the factorisations
        # above (for rocprofv3 runs of the solve kernels)
        x0, b = spm.rhs_from_x(A, 1, seed=2024)
        db = pastix.DeviceArray(b)
        st = []
        for _ in range(args.solves):
            db.upload(b)
            pastix.solve_device(h, db, A.shape[0], 1)
            st.append(float(h.dparm[pastix.dparm.solv_time]))
        db.free()
        print(json.dumps({"solves": args.solves, "solve_p50_ms": float(np.median(st)) * 1e3}))
        return

    # pastix_task_numfact on new values of the analysed pattern: permuted
    # values by one gather on the host, their upload, the device fill of the
    # tables and the factorisation (wall clock); then the original values back
    refill = {"first_numfact_wall_s": t_first}
    walls = []
    for k, dg in enumerate((6.5, 6.25, 6.0)):
        S.set_values(spm.laplacian7(*grid, diag=dg))
        t0 = time.perf_counter()
        pastix.task_numfact(h, S)
        walls.append(time.perf_counter() - t0)
        log(f"numfact with new values (diag {dg}): wall {walls[-1]*1e3:.1f} ms, device "
            f"{h.dparm[pastix.dparm.fact_time]*1e3:.1f} ms")
    refill["numfact_new_values_wall_ms"] = float(np.median(walls)) * 1e3
    refill["device_numfact_ms"] = float(h.dparm[pastix.dparm.fact_time]) * 1e3

    # per-kernel split of one more factorisation (HIP events around every
    # update launch), on one stream so that no kernel shares the device
    pastix.set_amd_iparm(h, pastix.amd_iparm.profile, 1)
    pastix.set_amd_iparm(h, pastix.amd_iparm.overlap, 0)
    pastix.refactorize(h)
    ft = pastix.fact_times(h)
    kt = pastix.kernel_times(h)
    t_single = float(h.dparm[pastix.dparm.fact_time])
    pastix.set_amd_iparm(h, pastix.amd_iparm.profile, 0)
    pastix.set_amd_iparm(h, pastix.amd_iparm.overlap, OVERLAP)
    pastix.refactorize(h)
    uf = pastix.update_flops(h)
    kern = {}
    for cls, name, fl_ in (("update_direct", "k_update_direct<double,0>", uf["direct"]),
                           ("update_wgemm", "k_update_wgemm<double>", uf["multi"]),
                           ("update_acc", "k_update_seq<double,0>", uf["single"])):
        t_, n_ = kt[cls]
        if n_ > 0:
            kern[name] = {"ms": t_ * 1e3, "launches": n_, "avg_launch_ms": t_ / n_ * 1e3,
                          "flops": fl_, "tflops": fl_ / t_ / 1e12}
    for cls, name in (("panel", "panel steps (k_potrf64 + k_panel_trsm + k_panel_upd) and the solve nodes' inverses"),
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
    solve_bytes = 2 * (nnzL + A.shape[0]) * 8 + 2 * 2 * A.shape[0] * 8
    solve_gbs = solve_bytes / (p50 * 1e-3) / 1e9
    db.free()
    pastix.finalize(h)

    fp32 = None
    if not args.no_fp32:
        h32, S32 = setup(pastix, spm, A, grid, prec=spm.SpmFloat, label="fp32")
        pastix.task_numfact(h32, S32)
        t32 = []
        for _ in range(max(2, args.steps // 2)):
            pastix.refactorize(h32)
            t32.append(float(h32.dparm[pastix.dparm.fact_time]))
        # device-resident right-hand side, as for the FP64 solve: one FP32
        # solve + FP64 CG on the device, timed on the wall clock
        db32 = pastix.DeviceArray(b)
        dx32 = pastix.DeviceArray(np.zeros_like(b))
        tr = []
        for _ in range(max(3, args.solves // 10)):
            r0 = time.perf_counter()
            pastix.solve_refine_device(h32, db32, dx32, A.shape[0], 1)
            tr.append(time.perf_counter() - r0)
        xr = dx32.download()
        db32.free()
        dx32.free()
        t_ref = float(np.median(tr))
        # the same through the host API (copies and permutations on the host)
        xh = b.copy(order="F")
        r0 = time.perf_counter()
        pastix.task_solve(h32, S32, xh)
        pastix.task_refine(h32, S32, b, xh)
        t_host = time.perf_counter() - r0
        b32 = berr_inf(A, xr[:, 0], b[:, 0])
        f32 = float(np.median(t32))
        fp32 = {"numfact_ms": f32 * 1e3, "numfact_gflops": thflops / f32 / 1e9,
                "solve_refine_ms": t_ref * 1e3, "refine_iters": int(h32.iparm[pastix.iparm.nbiter]),
                "solve_refine_ms_host_api": t_host * 1e3,
                "berr": b32, "berr_target_met": bool(b32 <= 1e-12),
                "berr_host_api": berr_inf(A, xh[:, 0], b[:, 0]),
                "speedup_numfact_vs_fp64": tmed / f32,
                "speedup_numfact_plus_solve_vs_fp64": (tmed + p50 * 1e-3) / (f32 + t_ref)}
        log(f"fp32+IR: numfact {f32*1e3:.2f} ms solve+refine {t_ref*1e3:.2f} ms (host API {t_host*1e3:.1f} ms) "
            f"iters {fp32['refine_iters']} berr {b32:.2e}")
        pastix.finalize(h32)

    cpu = None
    if not args.no_cpu:
        cg = tuple(args.cpu_grid) if args.cpu_grid else grid
        nt = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu_gf, cpu_t, blas = cpu_baseline(pastix, spm, cg, nt)
        c1 = cpu_sequential_c1(pastix, spm)
        cpu = {"value": cpu_gf, "unit": "GFLOP/s", "cores": nt, "kind": "port",
               "sample": f"whole LL^t numfact of the {cg[0]}x{cg[1]}x{cg[2]} 7-point Laplacian (the headline "
                         f"matrix) by the oracle's restatement of the reference static scheduler, {nt} threads, "
                         f"{cpu_t:.1f} s",
               "cpu_model": cpu_model(), "blas": blas or "plain loops", "numfact_s": cpu_t,
               "gpu_speedup_numfact": (thflops / tmed) / (cpu_gf * 1e9),
               "c1_sequential": c1}
        log(f"cpu baseline {cpu_gf:.2f} GFLOP/s in {cpu_t:.1f}s on {nt} threads; C1 sequential "
            f"{c1['gflops']:.2f} GFLOP/s berr {c1['berr']:.2e}")

    legs = {}
    for leg in [x for x in args.legs.split(",") if x and x != "none"]:
        t0 = time.perf_counter()
        if leg == "getrf":
            legs["getrf_fp32_gmres"] = leg_getrf(pastix, spm, tuple(args.getrf_grid))
        elif leg == "batch":
            legs["batch"] = leg_batch(pastix, spm, args.batch_systems)
        elif leg == "serving":
            legs["serving"] = leg_serving(pastix, spm, tuple(args.serving_grid), args.serving_requests)
        log(f"leg {leg}: {time.perf_counter() - t0:.1f}s")

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
        "solve_roofline": {"bound": "hbm", "achieved": solve_gbs, "peak": 8000.0, "unit": "GB/s",
                           "frac": solve_gbs / 8000.0, "algorithmic_bytes": solve_bytes,
                           "traffic": solve_traffic(grid),
                           "traffic_unit": "HBM bytes per 1-RHS solve, all solve kernels (rocprofv3 FETCH_SIZE "
                                           "and WRITE_SIZE, copy-calibrated; profiles/)"},
        "fp32_ir": fp32,
        "refill": refill,
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": kern[dom]["tflops"],
                     "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": kern[dom]["tflops"] / FP64_MFMA_PEAK_TFLOPS,
                     "traffic": traffic_per_launch(dom, grid),
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, "
                                     "copy-calibrated; profiles/)",
                     "launches": kern[dom]["launches"], "avg_launch_ms": kern[dom]["avg_launch_ms"],
                     "flops_per_factorization": kern[dom]["flops"],
                     "update_family_tflops": upd_tf},
        "numfact_split_ms": {"panel": ft[1] * 1e3, "update": ft[3] * 1e3,
                             "note": "one-stream profiled run (overlap off)"},
        "overlap": bool(OVERLAP),
        "level_schedule_tuning": {k: ("table entry" if v else "plan rules") for k, v in TUNED.items()},
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    out.update(legs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
