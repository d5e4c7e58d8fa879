/**
 * @file pastix_amd.h
 *
 * Extensions of the MI355X engine that have no counterpart in the reference
 * API.  They live in their own parameter array so that the reference
 * iparm/dparm numbering (include/pastix/api.h) stays untouched.
 */
#ifndef PASTIX_AMD_EXT_H
#define PASTIX_AMD_EXT_H

#include "pastix.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pastix_amd_iparm_e {
    /* Arithmetic of the factor: PastixDouble (default: that of the matrix) or
     * PastixFloat (FP32 factors refined in FP64 by pastix_task_refine).      */
    PXA_IPARM_FACTOR_PRECISION = 0,
    /* HIP device ordinal used by this instance (default 0).                  */
    PXA_IPARM_DEVICE           = 1,
    /* Apply the reference's level reordering of supernodes, as its ordering
     * step always does (default 1; 0 keeps the ordering as computed).      */
    PXA_IPARM_LEVEL_ORDER      = 2,
    /* Keep a device copy of the filled coefficients so that
     * pastix_amd_refactorize() can rerun the factorisation from HBM
     * (default 0; costs one more coefficient table in device memory).      */
    PXA_IPARM_RESIDENT_INPUT   = 3,
    /* Record per-kernel-family device times of numfact (default 0).        */
    PXA_IPARM_PROFILE          = 4,
    /* Capture numfact and device solves in HIP graphs on first use and
     * replay them (default 0).  A graph is re-captured when the static-pivot
     * threshold changes; up to 16 solve graphs are kept, keyed by the
     * right-hand-side pointer, leading dimension and nrhs.  Profiling
     * (PXA_IPARM_PROFILE) runs numfact without the graph.                   */
    PXA_IPARM_GRAPH            = 5,
    /* Run each level's updates into cblks beyond the next level on a second
     * HIP stream, beside the next level's panel steps (default 1; 0: one
     * stream, level after level).  Results are identical either way.       */
    PXA_IPARM_OVERLAP          = 6,
    /* LL^t: factor the leaf region (the levels from the leaves up whose
     * supernodes are at most 64 columns wide) in one persistent launch whose
     * tasks are released by device-side dependency counters, as the
     * reference's dynamic scheduler releases them (default 0: level by
     * level like the levels above; 1 is not validated on the device yet).   */
    PXA_IPARM_LEAF_TASKS       = 7,
    /* Workgroups of that persistent launch (default 0: 512).                */
    PXA_IPARM_LEAF_GRID        = 8,
    /* Widest solve node, in columns (default -1: 1024; 0 or 1: every
     * supernode its own node).  A solve node is a chain of supernodes (each
     * the parent of the one before, coupled to all later ones by full
     * blocks) that the solve treats as one diagonal block with its whole
     * inverse: one level of the solve instead of one per supernode.  Setting
     * it after the analysis rebuilds the schedule.                          */
    PXA_IPARM_SOLVE_NODE_MAX   = 9,
    PXA_IPARM_SIZE             = 10
} pastix_amd_iparm_t;

/* Set / get an engine parameter.  Returns PASTIX_SUCCESS or
 * PASTIX_ERR_BADPARAMETER. */
int pastix_amd_set_iparm(pastix_data_t *pastix_data, int key, pastix_int_t value);
int pastix_amd_get_iparm(const pastix_data_t *pastix_data, int key, pastix_int_t *value);

/* Sizes of the analysed structure:
 * out[0]=n, [1]=cblknbr, [2]=bloknbr, [3]=coefnbr (entries of one panel
 * table), [4]=browtab size, [5]=number of dependency levels,
 * [6]=update tiles, [7]=update entries. */
int pastix_amd_symbol_sizes(const pastix_data_t *pastix_data, int64_t out[8]);

/* Copy the analysed structure out (any pointer may be NULL).  Row/column
 * indices are in the permuted numbering; permtab maps user -> permuted.
 * cblk_* have cblknbr+1 entries (fcol, bloknum, coef) or cblknbr (stride);
 * blok_* have bloknbr entries. */
int pastix_amd_symbol_export(const pastix_data_t *pastix_data,
                             int64_t *cblk_fcol, int64_t *cblk_bloknum,
                             int64_t *cblk_stride, int64_t *cblk_coef,
                             int64_t *blok_frow, int64_t *blok_lrow,
                             int64_t *blok_fcblk, int64_t *blok_roff,
                             int64_t *permtab, int64_t *peritab);

/* Copy the device factor to host buffers of coefnbr entries each, in the
 * factor precision (U only for LU). */
int pastix_amd_coeftab_export(const pastix_data_t *pastix_data, void *L, void *U);

/* Solve with right-hand sides already resident in device memory, user
 * ordering, in the factor precision; ldb >= n. */
int pastix_amd_solve_device(pastix_data_t *pastix_data, pastix_int_t nrhs,
                            void *d_b, pastix_int_t ldb);

/* Refinement on FP64 device vectors in the user ordering (ldb, ldx >= n):
 * the reference's pastix_task_refine (refinement/pastix_task_refine.c:115-288)
 * without the host copies.  x holds the initial guesses.  Every right-hand
 * side runs its own recurrence (CG: all columns advance together with
 * per-column scalars; GMRES: one column after the other).  IPARM_NBITER,
 * DPARM_RELATIVE_ERROR and DPARM_REFINE_TIME (wall clock) as in the
 * reference. */
int pastix_amd_refine_device(pastix_data_t *pastix_data, pastix_int_t nrhs,
                             const double *d_b, pastix_int_t ldb,
                             double *d_x, pastix_int_t ldx);
/* x = M^{-1} b (one solve in the factor precision) followed by the same
 * refinement: pastix_task_solve + pastix_task_refine on device vectors. */
int pastix_amd_solve_refine_device(pastix_data_t *pastix_data, pastix_int_t nrhs,
                                   const double *d_b, pastix_int_t ldb,
                                   double *d_x, pastix_int_t ldx);

/* Normwise backward error max|b-Ax| / (||A||_inf max|x| + max|b|) for host
 * FP64 vectors in the user ordering, computed on the device. */
int pastix_amd_backward_error(pastix_data_t *pastix_data, const double *b,
                              const double *x, double *berr);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pastix_amd_backward_error on nrhs FP64 device columns
 * d_b (ldb) and d_x (ldx) in the user ordering, results into the device array
 * d_berr[nrhs].  Allocates nothing: the partial maxima use a library-static
 * device array of 768 KiB per GPU (512 blocks x 3 maxima x 64 columns),
 * held until the process ends, so nrhs is at most 64 (larger values return
 * PASTIX_ERR_BADPARAMETER).  Calls that overlap on one GPU share that array
 * and race; the call returns once the results are written. */
int pastix_amd_backward_error_device_noalloc(pastix_data_t *pastix_data, pastix_int_t nrhs,
                                             const double *d_b, pastix_int_t ldb,
                                             const double *d_x, pastix_int_t ldx, double *d_berr);

/* ---- Residual SpMV and vector kernels of the refinement ------------------
 * The operations the reference's refinement drivers call through their
 * z_solver table (refinement/z_refine_functions.c:115-148), on FP64 device
 * vectors and on the instance's HIP stream; every call returns after its
 * work is done.  Reductions use a fixed grid and order: bitwise reproducible.
 *
 * y = alpha op(A) x + beta y with A the permuted matrix P A P^T of the last
 * numfact (x, y of length n in the permuted ordering): the reference's
 * bcsc_zspmv (bcsc/bcsc_zspmv.c:625-646).  op(A) = A for PastixNoTrans, A^T
 * for PastixTrans and PastixConjTrans (real arithmetic). */
int pastix_amd_bcsc_spmv(pastix_data_t *pastix_data, pastix_trans_t trans, double alpha, const double *d_x,
                         double beta, double *d_y);
/* bvec_zdotc / bvec_znrm2 (scaled sum of squares) / bvec_zaxpy / bvec_zscal /
 * bvec_zcopy / bvec_zgemv (bcsc/bvec_zcompute.c:67-120,245-280,384-420,
 * 547-600,1168-1210) on vectors of length n (gemv: A is m x n column-major,
 * y = alpha A x + beta y, y not read when beta = 0).  The scalar results go
 * to host memory. */
int pastix_amd_bvec_dot(pastix_data_t *pastix_data, pastix_int_t n, const double *d_x, const double *d_y,
                        double *result);
int pastix_amd_bvec_nrm2(pastix_data_t *pastix_data, pastix_int_t n, const double *d_x, double *result);
int pastix_amd_bvec_axpy(pastix_data_t *pastix_data, pastix_int_t n, double alpha, const double *d_x, double *d_y);
int pastix_amd_bvec_scal(pastix_data_t *pastix_data, pastix_int_t n, double alpha, double *d_x);
int pastix_amd_bvec_copy(pastix_data_t *pastix_data, pastix_int_t n, const double *d_x, double *d_y);
int pastix_amd_bvec_gemv(pastix_data_t *pastix_data, pastix_int_t m, pastix_int_t n, double alpha, const double *d_A,
                         pastix_int_t lda, const double *d_x, double beta, double *d_y);

/* Device time of the last numerical factorisation split by kernel family:
 * out[0]=total, [1]=panel steps (diagonal blocks, panel TRSM, trailing
 * panel updates), [2]=0, [3]=Schur update (seconds), and the number of
 * timed spans of each family out[4..6]; the split is filled only when
 * PXA_IPARM_PROFILE is set. */
int pastix_amd_fact_times(const pastix_data_t *pastix_data, double out[8]);

/* Device time of the last numerical factorisation per kernel class, filled
 * when PXA_IPARM_PROFILE is set: out[2k] seconds and out[2k+1] timed spans
 * for k = 0 panel steps (diagonal blocks, panel TRSM and trailing panel
 * update of one level per span), 1 single-source update tiles
 * (k_update_direct), 2 multi-source product tiles (k_update_wgemm), 3
 * multi-source ordered sums (k_update_gather), 4 single-pass multi-source
 * update (k_update_acc, used when a level's products exceed the workspace),
 * 5 the leaf region's persistent task launch (PXA_IPARM_LEAF_TASKS).
 * Every span of classes 1-4 is exactly one kernel launch (LU runs the L
 * and U sides as separate launches). */
int pastix_amd_kernel_times(const pastix_data_t *pastix_data, double out[12]);

/* Schur-update flops by kernel, counting only the (p, q) pairs the
 * reference's per-block GEMMs compute: out[0] single-source tiles and out[1]
 * multi-source tiles of two-phase levels (product tiles, then gather) on the
 * L side (block(p) >= block(q)), out[2], out[3] the same for the U side of
 * LU (block(p) > block(q)), out[4] the flops of the multi-source product
 * tiles as computed (one side), out[5], out[6] the multi-source tiles of
 * single-pass levels (k_update_acc), L and U side, out[7] the updates inside
 * the leaf region's persistent launch.  With that launch on, out[0..6] count
 * what the level kernels run: the region's levels leave them and its
 * updates into the supernodes above it join out[0] and out[5]. */
int pastix_amd_update_flops(const pastix_data_t *pastix_data, double out[8]);

/* ---- Level-schedule tuning (no reference counterpart) --------------------
 * The level schedule makes two choices per dependency level: whether its
 * multi-source update tiles are summed in one pass (1) or through the
 * product workspace (0), and whether its panels are updated right-looking
 * (1) or left-looking (0).  The plan's rules pick them unless per-level
 * values are given (-1 keeps the rule).  pastix_amd/autotune.py times both
 * choices per level on the device and writes a tuning table; a table is read
 * by pastix_amd_load_tuning, or at analysis from the file PASTIX_AMD_TUNING
 * names, for an entry matching the analysed partition and factorisation. */
int pastix_amd_set_level_knobs(pastix_data_t *pastix_data, int nlevels, const int8_t *single_pass,
                               const int8_t *panel_right);
int pastix_amd_load_tuning(pastix_data_t *pastix_data, const char *path, int *applied);
int pastix_amd_save_tuning(const pastix_data_t *pastix_data, const char *path);
/* Per level: 4 ints (cblks, multi-source tiles, single-pass, right-looking);
 * returns the number of levels written (-1 before the analysis). */
int pastix_amd_level_info(const pastix_data_t *pastix_data, int32_t *out, int maxlevels);
/* Solve structure after the analysis: out[0] solve nodes, [1] nodes of
 * several supernodes, [2] solve levels, [3] merge rounds of the nodes'
 * diagonal-block inverses (summed over the batches the factorisation issues
 * them in), [4] their 64x64 tiles, [5] entries of the
 * inverses of the nodes of several supernodes, [6] flops of the merge rounds
 * (one triangular factor), [7] factorisation levels.  No reference
 * counterpart (the solve nodes are this engine's). */
int pastix_amd_solve_info(const pastix_data_t *pastix_data, double out[8]);
/* Per level of the last numfact run with PXA_IPARM_PROFILE: 2 doubles (panel
 * steps, Schur updates, seconds); returns the number of levels written. */
int pastix_amd_level_times(const pastix_data_t *pastix_data, double *out, int maxlevels);

/* Algorithmic flops of the analysed factorisation: out[0]=DPARM_FACT_THFLOPS,
 * [1]=diagonal factorisations, [2]=panel TRSMs, [3]=Schur updates counted
 * as the reference's per-block GEMMs (kernels/core_zgemmsp.c:1352-1395). */
int pastix_amd_flops(const pastix_data_t *pastix_data, double out[4]);

/* Update-schedule statistics per dependency level (up to maxlevels levels),
 * 8 doubles each: single-source tiles, their flops, multi-source tiles,
 * their entries, their flops, flops issued on 32-row quadrants (all tiles),
 * sum of source widths over entries, 0.  Returns the number of levels
 * written or -1 before analysis. */
int pastix_amd_plan_stats(const pastix_data_t *pastix_data, double *out, int maxlevels);

/* ---- Batches of independent systems (no counterpart in the reference) ----
 * nsys systems are analysed, factored and solved as their block-diagonal
 * union: every kernel launch (or, with PXA_IPARM_GRAPH, every graph launch)
 * covers the supernodes of all systems at one dependency level.  System s
 * takes rows [off[s], off[s+1]) of the union.  Orders may be NULL (or hold
 * NULL entries): the system is then ordered by nested dissection of its own
 * graph.  All systems share mtxtype; values are held in FP64.  The factor
 * precision, graph mode etc. are the instance's. */
int pastix_amd_batch_analyze(pastix_data_t *pastix_data, pastix_int_t nsys, const spmatrix_t *const *spms,
                             const pastix_order_t *const *orders);
/* Numerical factorisation of the batch (same patterns as analysed). */
int pastix_amd_batch_numfact(pastix_data_t *pastix_data, const spmatrix_t *const *spms);
/* Solves in place: system s has nrhs[s] FP64 right-hand sides at b[s]
 * (column-major, leading dimension ldb[s]); all go through one multi-RHS
 * solve of max(nrhs) columns. */
int pastix_amd_batch_solve(pastix_data_t *pastix_data, const pastix_int_t *nrhs, void *const *b,
                           const pastix_int_t *ldb);
/* Batched preconditioned CG (IPARM_REFINEMENT must be PastixRefineCG): every
 * (system, right-hand side) pair runs the reference's CG recurrence
 * (refinement/z_refine_grad.c) with its own scalars and stopping test
 * ||r||/||b|| <= DPARM_EPSILON_REFINEMENT; M^{-1} is one multi-RHS solve of
 * the union per iteration.  x[s] holds the initial guesses on entry.
 * IPARM_NBITER receives the largest iteration count, DPARM_RELATIVE_ERROR the
 * largest final ||r||/||b||. */
int pastix_amd_batch_refine(pastix_data_t *pastix_data, const pastix_int_t *nrhs, void *const *b,
                            const pastix_int_t *ldb, void *const *x, const pastix_int_t *ldx);
/* Both in one: x[s] receives the solution of one multi-RHS solve in the
 * factor precision refined by the batched CG above (b[s] is left as is);
 * the right-hand sides cross to the device once. */
int pastix_amd_batch_solve_refine(pastix_data_t *pastix_data, const pastix_int_t *nrhs, void *const *b,
                                  const pastix_int_t *ldb, void *const *x, const pastix_int_t *ldx);
/* Offsets of the systems in the union (nsys+1 entries, may be NULL);
 * returns nsys or PASTIX_ERR_BADPARAMETER. */
int pastix_amd_batch_offsets(const pastix_data_t *pastix_data, int64_t *off);

/* Rerun the numerical factorisation from the device copy of the filled
 * coefficients (needs PXA_IPARM_RESIDENT_INPUT before pastix_task_numfact);
 * updates DPARM_FACT_TIME / DPARM_FACT_FLOPS like pastix_subtask_sopalin. */
int pastix_amd_refactorize(pastix_data_t *pastix_data);

/* Refactorise A + delta*I (same pattern, no counterpart in the reference):
 * the filled coefficients are restored from their device copy (needs
 * PXA_IPARM_RESIDENT_INPUT), delta is added to their diagonal and to the
 * diagonal of the device matrix used by the refinement, DPARM_A_NORM becomes
 * ||A + delta I||_F (static-pivot threshold), and the factorisation reruns.
 * delta is relative to the analysed matrix, not cumulative; every diagonal
 * entry must be stored.  pastix_amd_refactorize() returns to delta = 0. */
int pastix_amd_refactorize_shift(pastix_data_t *pastix_data, double delta);

/* Device memory held by the instance's engine, in bytes (factor, inverses,
 * plan, work space, input copy). */
int pastix_amd_device_bytes(const pastix_data_t *pastix_data, int64_t *bytes);

/* ---- Serving loop (no counterpart in the reference) ----------------------
 * A server drives nslots instances analysed and factored (LL^t) on the same
 * pattern with PXA_IPARM_RESIDENT_INPUT set; nslots is what the HBM budget
 * admits (pastix_amd_device_bytes gives one instance's share).  Slot 0
 * serves the matrix it was factored on (version 0).  Requests:
 *   solve: nrhs FP64 right-hand sides at b (user ordering, ldb >= n),
 *          overwritten with the solution of the newest matrix version
 *          submitted before the request; queued solves of one version are
 *          merged in arrival order, up to max_rhs columns, into one solve in
 *          the factor precision followed by batched FP64 CG
 *          (DPARM_EPSILON_REFINEMENT, IPARM_ITERMAX of the slot);
 *   refactor: the next version, A + shift*I (A the analysed matrix); it
 *          runs on a slot whose version has no solve left while other slots
 *          keep solving, each slot on its own HIP stream.
 * One worker thread per slot.  b must stay valid until the request is
 * waited for. */
typedef struct pastix_amd_server_s pastix_amd_server_t;
int pastix_amd_server_create(pastix_amd_server_t **srv, pastix_data_t **slots, int nslots, pastix_int_t max_rhs);
int pastix_amd_server_submit_solve(pastix_amd_server_t *srv, pastix_int_t nrhs, double *b, pastix_int_t ldb,
                                   int64_t *ticket);
int pastix_amd_server_submit_refactor(pastix_amd_server_t *srv, double shift, int64_t *ticket);
/* Blocks until the request is done; latency = completion - submission
 * (seconds), version = matrix version it used or created.  Returns the
 * request's status. */
int pastix_amd_server_wait(pastix_amd_server_t *srv, int64_t ticket, double *latency, int64_t *version);
/* out[0] solve batches run, [1] right-hand sides they carried, [2] newest
 * version, [3] requests not finished. */
int pastix_amd_server_stats(pastix_amd_server_t *srv, int64_t out[4]);
/* HBM-budget admission: the device memory of the merged solves' work blocks
 * (right-hand sides, solutions, refinement vectors) that all slots together
 * may hold; by default IPARM_GPU_MEMORY_PERCENTAGE (of the first slot) of the
 * device memory free when the server is created.  A batch is admitted only
 * while it fits, shorter if need be; the oldest request otherwise waits
 * while idle slots give their blocks back, and a request that cannot fit
 * alone ends with PASTIX_ERR_OUTOFMEMORY. */
int pastix_amd_server_set_budget(pastix_amd_server_t *srv, int64_t bytes);
/* out[0] budget, [1] bytes the slots hold now, [2] the most they held at
 * once, [3] times the oldest request waited for memory. */
int pastix_amd_server_memory(pastix_amd_server_t *srv, int64_t out[4]);
/* Finishes every admitted request, stops the workers, frees the server. */
int pastix_amd_server_destroy(pastix_amd_server_t **srv);

#ifdef __cplusplus
}
#endif

#endif
