/*
 * CPU oracle for the MI355X engine -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference PaStiX numerical path, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker and
 * CPU baseline.  It is never linked into or called by libpastix_amd.so.
 *
 * It runs on the same block structure as the engine (1D panel layout: cblk c
 * owns coef[cblk_coef[c] .. + stride*width), column-major, ld = stride) and
 * follows, function by function:
 *   ora_fill      kernels/cpucblk_zinit.c:229-309      (cpucblk_zfillin_fr)
 *   potf2/potrf   kernels/core_zpotrfsp.c:66-189        (64-blocked, static pivot)
 *   getf2/getrf   kernels/core_zgetrfsp.c:63-259, core_zgetmo.c:48-60, core_zgeadd.c
 *   trsm panel    kernels/core_ztrsmsp.c:147-181,357-398
 *   gemmsp        kernels/core_zgemmsp.c:396-465        (2d2d loop, 1D addressing)
 *   sequential    sopalin/sequential_zpotrf.c:34-68, sequential_zgetrf.c (same shape)
 *   solve         kernels/solve_ztrsmsp.c:67-100,332-557, sopalin/sequential_ztrsm.c:34-103
 *   spmv          bcsc/bcsc_zspmv.c:55-202
 *   thflops       symbol/symbol_cost.c:53-88, symbol_cost_flops.c, common/flops.h
 *   cg            refinement/z_refine_grad.c:46-156
 *   sr, bicgstab  refinement/z_refine_pivot.c:26-119, z_refine_bicgstab.c:25-185
 *   gmres         refinement/z_refine_gmres.c:46-352
 *
 * Parity status of each function is recorded in the repository design notes;
 * every function here is pinned by a test in tests/ (dense factorisations,
 * closed-form flop counts, scipy SpMV, dense solves).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define BLK 64 /* MAXSIZEOFBLOCKS, kernels/core_zpotrfsp.c:25 */

static inline int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }
static inline int64_t width(const ora_symbol_t *s, int64_t c) { return s->cblk_fcol[c + 1] - s->cblk_fcol[c]; }
static inline int64_t brows(const ora_symbol_t *s, int64_t b) { return s->blok_lrow[b] - s->blok_frow[b] + 1; }

/* C(m x n, ldc) += alpha * A(m x k, lda) * op(B), op(B)(l, j) = B[j + l*ldb] (B^T, NT) */
static void gemm_nt(int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda, const double *B,
                    int64_t ldb, double *C, int64_t ldc) {
    for (int64_t j = 0; j < n; j++)
        for (int64_t l = 0; l < k; l++) {
            const double b = alpha * B[j + l * ldb];
            if (b == 0.0) continue;
            const double *a = A + l * lda;
            double *c = C + j * ldc;
            for (int64_t i = 0; i < m; i++) c[i] += a[i] * b;
        }
}
/* C += alpha * A * B with B(l, j) = B[l + j*ldb] (NN) */
static void gemm_nn(int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda, const double *B,
                    int64_t ldb, double *C, int64_t ldc) {
    for (int64_t j = 0; j < n; j++)
        for (int64_t l = 0; l < k; l++) {
            const double b = alpha * B[l + j * ldb];
            if (b == 0.0) continue;
            const double *a = A + l * lda;
            double *c = C + j * ldc;
            for (int64_t i = 0; i < m; i++) c[i] += a[i] * b;
        }
}

/* X * L^T = B in place, L lower non-unit n x n: dtrsm(Right, Lower, Trans, NonUnit) */
static void trsm_rltn(int64_t m, int64_t n, const double *L, int64_t ldl, double *X, int64_t ldx) {
    for (int64_t j = 0; j < n; j++) {
        double *xj = X + j * ldx;
        for (int64_t k = 0; k < j; k++) {
            const double l = L[j + k * ldl];
            if (l == 0.0) continue;
            const double *xk = X + k * ldx;
            for (int64_t i = 0; i < m; i++) xj[i] -= l * xk[i];
        }
        const double d = 1.0 / L[j + j * ldl];
        for (int64_t i = 0; i < m; i++) xj[i] *= d;
    }
}
/* X * U = B in place, U upper n x n stored in A: dtrsm(Right, Upper, NoTrans, diag) */
static void trsm_run(int64_t m, int64_t n, const double *U, int64_t ldu, int unit, double *X, int64_t ldx) {
    for (int64_t j = 0; j < n; j++) {
        double *xj = X + j * ldx;
        for (int64_t k = 0; k < j; k++) {
            const double u = U[k + j * ldu];
            if (u == 0.0) continue;
            const double *xk = X + k * ldx;
            for (int64_t i = 0; i < m; i++) xj[i] -= u * xk[i];
        }
        if (!unit) {
            const double d = 1.0 / U[j + j * ldu];
            for (int64_t i = 0; i < m; i++) xj[i] *= d;
        }
    }
}
/* L X = B in place, L unit lower n x n: dtrsm(Left, Lower, NoTrans, Unit), B n x m */
static void trsm_llnu(int64_t n, int64_t m, const double *L, int64_t ldl, double *X, int64_t ldx) {
    for (int64_t j = 0; j < m; j++) {
        double *x = X + j * ldx;
        for (int64_t k = 0; k < n; k++) {
            const double xk = x[k];
            if (xk == 0.0) continue;
            for (int64_t i = k + 1; i < n; i++) x[i] -= xk * L[i + k * ldl];
        }
    }
}

/* ---------------------------------------------------------------- fill --- */
int ora_fill(const ora_symbol_t *s, const int64_t *colptr, const int64_t *rows, const double *vals,
             const int64_t *tcolptr, const int64_t *trows, const double *tvals, double *L, double *U) {
    memset(L, 0, s->coefnbr * sizeof(double));
    if (U) memset(U, 0, s->coefnbr * sizeof(double));
    int64_t dropped = 0;
    for (int64_t c = 0; c < s->cblknbr; c++) {
        const int64_t fc = s->cblk_fcol[c], lc = s->cblk_fcol[c + 1], st = s->cblk_stride[c];
        for (int64_t j = fc; j < lc; j++) {
            int64_t b = s->cblk_bloknum[c];
            for (int64_t k = colptr[j]; k < colptr[j + 1]; k++) {
                const int64_t i = rows[k];
                if (i < j) continue;
                while (b < s->cblk_bloknum[c + 1] && s->blok_lrow[b] < i) b++;
                if (b >= s->cblk_bloknum[c + 1] || s->blok_frow[b] > i) { dropped++; continue; }
                L[s->cblk_coef[c] + s->blok_roff[b] + (i - s->blok_frow[b]) + (j - fc) * st] = vals[k];
            }
            if (U) {
                int64_t bu = s->cblk_bloknum[c];
                for (int64_t k = tcolptr[j]; k < tcolptr[j + 1]; k++) {
                    const int64_t i = trows[k];
                    if (i <= j) continue;
                    while (bu < s->cblk_bloknum[c + 1] && s->blok_lrow[bu] < i) bu++;
                    if (bu >= s->cblk_bloknum[c + 1] || s->blok_frow[bu] > i) { dropped++; continue; }
                    U[s->cblk_coef[c] + s->blok_roff[bu] + (i - s->blok_frow[bu]) + (j - fc) * st] = tvals[k];
                }
            }
        }
    }
    return (int)(dropped > 0);
}

/* ------------------------------------------------------- diagonal LLt --- */
static int potf2(int64_t n, double *A, int64_t lda, int64_t *npiv, double crit) {
    for (int64_t k = 0; k < n; k++) {
        double *akk = A + k * (lda + 1);
        if (fabs(*akk) < crit) { *akk = crit; (*npiv)++; }
        if (*akk < 0.0) return -1;
        *akk = sqrt(*akk);
        const double alpha = 1.0 / *akk;
        for (int64_t i = k + 1; i < n; i++) A[i + k * lda] *= alpha;
        /* dsyr lower: A[k+1:, k+1:] -= a a^T */
        for (int64_t j = k + 1; j < n; j++) {
            const double aj = A[j + k * lda];
            for (int64_t i = j; i < n; i++) A[i + j * lda] -= A[i + k * lda] * aj;
        }
    }
    return 0;
}

static int potrfsp(int64_t n, double *A, int64_t lda, int64_t *npiv, double crit) {
    for (int64_t k0 = 0; k0 < n; k0 += BLK) {
        const int64_t bs = imin(BLK, n - k0);
        double *akk = A + k0 * (lda + 1);
        if (potf2(bs, akk, lda, npiv, crit)) return -1;
        const int64_t rest = n - k0 - bs;
        if (rest > 0) {
            double *a1 = akk + bs, *a2 = a1 + bs * lda;
            trsm_rltn(rest, bs, akk, lda, a1, lda);
            /* dsyrk lower: a2 -= a1 a1^T */
            for (int64_t j = 0; j < rest; j++)
                for (int64_t l = 0; l < bs; l++) {
                    const double b = a1[j + l * lda];
                    for (int64_t i = j; i < rest; i++) a2[i + j * lda] -= a1[i + l * lda] * b;
                }
        }
    }
    return 0;
}

/* -------------------------------------------------------- diagonal LU --- */
static void getf2(int64_t m, int64_t n, double *A, int64_t lda, int64_t *npiv, double crit) {
    const int64_t mn = imin(m, n);
    for (int64_t k = 0; k < mn; k++) {
        double *akk = A + k * (lda + 1);
        if (fabs(*akk) < crit) { *akk = (*akk < 0.0) ? -crit : crit; (*npiv)++; }
        const double alpha = 1.0 / *akk;
        for (int64_t i = k + 1; i < m; i++) A[i + k * lda] *= alpha;
        if (k + 1 < mn)
            for (int64_t j = k + 1; j < n; j++) {
                const double u = A[k + j * lda];
                for (int64_t i = k + 1; i < m; i++) A[i + j * lda] -= A[i + k * lda] * u;
            }
    }
}

static void getrfsp(int64_t n, double *A, int64_t lda, int64_t *npiv, double crit) {
    for (int64_t k0 = 0; k0 < n; k0 += BLK) {
        const int64_t tm = n - k0, bs = imin(BLK, tm);
        double *akk = A + k0 * (lda + 1);
        getf2(tm, bs, akk, lda, npiv, crit);
        const int64_t rest = tm - bs;
        if (rest > 0) {
            double *lik = akk + bs, *ukj = akk + bs * lda, *aij = ukj + bs;
            trsm_llnu(bs, rest, akk, lda, ukj, lda);
            gemm_nn(rest, rest, bs, -1.0, lik, lda, ukj, lda, aij, lda);
        }
    }
}

/* ------------------------------------------ optional CBLAS (CPU timing) --- */
/* The static scheduler below times the reference's CPU path, which calls a
 * sequential CBLAS (README.md:76-77) for its TRSMs and GEMMs; ora_use_blas()
 * loads one (the cblas_dgemm / cblas_dtrsm of a given shared library, with its
 * own threading set to one thread).  The sequential drivers above never use
 * it: they stay plain loops. */
#include <dlfcn.h>
enum { CB_COL = 102, CB_NOTRANS = 111, CB_TRANS = 112, CB_UPPER = 121, CB_LOWER = 122, CB_NONUNIT = 131,
       CB_UNIT = 132, CB_RIGHT = 142 };
typedef void (*cblas_dgemm_t)(int, int, int, int, int, int, double, const double *, int, const double *, int,
                              double, double *, int);
typedef void (*cblas_dtrsm_t)(int, int, int, int, int, int, int, double, const double *, int, double *, int);
static cblas_dgemm_t blas_dgemm = NULL;
static cblas_dtrsm_t blas_dtrsm = NULL;

int ora_use_blas(const char *path, const char *prefix) {
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    char name[128];
    snprintf(name, sizeof name, "%scblas_dgemm", prefix);
    cblas_dgemm_t g = (cblas_dgemm_t)dlsym(h, name);
    snprintf(name, sizeof name, "%scblas_dtrsm", prefix);
    cblas_dtrsm_t t = (cblas_dtrsm_t)dlsym(h, name);
    snprintf(name, sizeof name, "%sopenblas_set_num_threads", prefix);
    void (*nt)(int) = (void (*)(int))dlsym(h, name);
    if (!g || !t) return -1;
    if (nt) nt(1);
    blas_dgemm = g;
    blas_dtrsm = t;
    return 0;
}

/* --------------------------------------------- Schur update (gemmsp) --- */
/* shift = 0: iterblok >= blok (L side), 1: iterblok > blok (U side) */
static void gemmsp_x(const ora_symbol_t *s, int shift, int64_t c, int64_t b, const double *A, const double *B,
                     double *C, int blas);
static void gemmsp(const ora_symbol_t *s, int shift, int64_t c, int64_t b, const double *A, const double *B,
                   double *C) {
    gemmsp_x(s, shift, c, b, A, B, C, 0);
}
static void gemmsp_x(const ora_symbol_t *s, int shift, int64_t c, int64_t b, const double *A, const double *B,
                     double *C, int blas) {
    const int64_t f = s->blok_fcblk[b];
    const int64_t K = width(s, c), st = s->cblk_stride[c], stf = s->cblk_stride[f];
    const int64_t N = brows(s, b);
    const double *Bb = B + s->cblk_coef[c] + s->blok_roff[b];
    int64_t fb = s->cblk_bloknum[f];
    for (int64_t i = b + shift; i < s->cblk_bloknum[c + 1]; i++) {
        while (!(s->blok_frow[fb] <= s->blok_frow[i] && s->blok_lrow[i] <= s->blok_lrow[fb])) fb++;
        const int64_t M = brows(s, i);
        const double *Ai = A + s->cblk_coef[c] + s->blok_roff[i];
        double *Ci = C + s->cblk_coef[f] + s->blok_roff[fb] + (s->blok_frow[i] - s->blok_frow[fb]) +
                     (s->blok_frow[b] - s->cblk_fcol[f]) * stf;
        if (blas && blas_dgemm)
            blas_dgemm(CB_COL, CB_NOTRANS, CB_TRANS, (int)M, (int)N, (int)K, -1.0, Ai, (int)st, Bb, (int)st, 1.0, Ci,
                       (int)stf);
        else
            gemm_nt(M, N, K, -1.0, Ai, st, Bb, st, Ci, stf);
    }
}

/* ------------------------------------------------ sequential numfact --- */
int ora_potrf(const ora_symbol_t *s, double *L, double crit, int64_t *npiv) {
    *npiv = 0;
    for (int64_t c = 0; c < s->cblknbr; c++) {
        const int64_t K = width(s, c), st = s->cblk_stride[c];
        double *P = L + s->cblk_coef[c];
        if (potrfsp(K, P, st, npiv, crit)) return -1;
        if (st > K) trsm_rltn(st - K, K, P, st, P + K, st);
        for (int64_t b = s->cblk_bloknum[c] + 1; b < s->cblk_bloknum[c + 1]; b++) gemmsp(s, 0, c, b, L, L, L);
    }
    return 0;
}

int ora_getrf(const ora_symbol_t *s, double *L, double *U, double crit, int64_t *npiv) {
    *npiv = 0;
    for (int64_t c = 0; c < s->cblknbr; c++) {
        const int64_t K = width(s, c), st = s->cblk_stride[c];
        double *P = L + s->cblk_coef[c], *Q = U + s->cblk_coef[c];
        for (int64_t j = 0; j < K; j++) /* core_zgeadd(Trans): L += U^T */
            for (int64_t i = 0; i < K; i++) P[i + j * st] += Q[j + i * st];
        getrfsp(K, P, st, npiv, crit);
        for (int64_t j = 0; j < K; j++) /* core_zgetmo: U_diag = L_diag^T */
            for (int64_t i = 0; i < K; i++) Q[i + j * st] = P[j + i * st];
        if (st > K) {
            trsm_run(st - K, K, P, st, 0, P + K, st);
            trsm_run(st - K, K, Q, st, 1, Q + K, st);
        }
        for (int64_t b = s->cblk_bloknum[c] + 1; b < s->cblk_bloknum[c + 1]; b++) {
            gemmsp(s, 0, c, b, L, U, L);
            if (b + 1 < s->cblk_bloknum[c + 1]) gemmsp(s, 1, c, b, U, L, U);
        }
    }
    return 0;
}

/* ----------------------------------------------------------- solve ---- */
int ora_solve(const ora_symbol_t *s, int lu, const double *L, const double *U, double *x, int64_t ldx, int nrhs) {
    const double *Bk = lu ? U : L;
    /* browtab for the backward sweep */
    int64_t *cnt = calloc(s->cblknbr + 1, sizeof(int64_t));
    int64_t *brow = malloc((s->bloknbr + 1) * sizeof(int64_t));
    if (!cnt || !brow) { free(cnt); free(brow); return -1; }
    for (int64_t c = 0; c < s->cblknbr; c++)
        for (int64_t b = s->cblk_bloknum[c] + 1; b < s->cblk_bloknum[c + 1]; b++) cnt[s->blok_fcblk[b] + 1]++;
    for (int64_t c = 0; c < s->cblknbr; c++) cnt[c + 1] += cnt[c];
    int64_t *pos = malloc((s->cblknbr + 1) * sizeof(int64_t));
    int64_t *owner = malloc((s->bloknbr + 1) * sizeof(int64_t));
    memcpy(pos, cnt, (s->cblknbr + 1) * sizeof(int64_t));
    for (int64_t c = 0; c < s->cblknbr; c++)
        for (int64_t b = s->cblk_bloknum[c]; b < s->cblk_bloknum[c + 1]; b++) {
            owner[b] = c;
            if (b > s->cblk_bloknum[c]) brow[pos[s->blok_fcblk[b]]++] = b;
        }
    for (int r = 0; r < nrhs; r++) {
        double *xr = x + r * ldx;
        /* forward: solve_cblk_ztrsmsp_forward, Left Lower NoTrans (Unit for LU) */
        for (int64_t c = 0; c < s->cblknbr; c++) {
            const int64_t K = width(s, c), st = s->cblk_stride[c], fc = s->cblk_fcol[c];
            const double *P = L + s->cblk_coef[c];
            double *xc = xr + fc;
            for (int64_t k = 0; k < K; k++) {
                if (!lu) xc[k] /= P[k + k * st];
                const double v = xc[k];
                for (int64_t i = k + 1; i < K; i++) xc[i] -= P[i + k * st] * v;
            }
            for (int64_t b = s->cblk_bloknum[c] + 1; b < s->cblk_bloknum[c + 1]; b++) {
                const int64_t nr = brows(s, b);
                double *xf = xr + s->blok_frow[b];
                const double *A = P + s->blok_roff[b];
                for (int64_t k = 0; k < K; k++) {
                    const double v = xc[k];
                    for (int64_t i = 0; i < nr; i++) xf[i] -= A[i + k * st] * v;
                }
            }
        }
        /* backward: solve_cblk_ztrsmsp_backward, Left Lower Trans on L (or on U^T for LU) */
        for (int64_t c = s->cblknbr - 1; c >= 0; c--) {
            const int64_t K = width(s, c), st = s->cblk_stride[c], fc = s->cblk_fcol[c];
            const double *P = Bk + s->cblk_coef[c];
            double *xc = xr + fc;
            for (int64_t k = K - 1; k >= 0; k--) {
                double v = xc[k];
                for (int64_t i = k + 1; i < K; i++) v -= P[i + k * st] * xc[i];
                xc[k] = v / P[k + k * st];
            }
            for (int64_t e = cnt[c + 1] - 1; e >= cnt[c]; e--) {
                const int64_t b = brow[e], sc = owner[b];
                const int64_t Ks = width(s, sc), sts = s->cblk_stride[sc], nr = brows(s, b);
                const double *A = Bk + s->cblk_coef[sc] + s->blok_roff[b];
                double *xs = xr + s->cblk_fcol[sc];
                const double *xb = xr + s->blok_frow[b];
                for (int64_t k = 0; k < Ks; k++) {
                    double v = 0.0;
                    for (int64_t i = 0; i < nr; i++) v += A[i + k * sts] * xb[i];
                    xs[k] -= v;
                }
            }
        }
    }
    free(cnt); free(brow); free(pos); free(owner);
    return 0;
}

/* ------------------------------------------------------------ spmv ---- */
void ora_spmv(int64_t n, const int64_t *rowptr, const int64_t *cols, const double *vals, double alpha,
              const double *x, double beta, double *y) {
    for (int64_t i = 0; i < n; i++) {
        double v = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) v += vals[k] * x[cols[k]];
        y[i] = (beta == 0.0) ? alpha * v : alpha * v + beta * y[i];
    }
}

/* --------------------------------------------------------- thflops ---- */
double ora_thflops(const ora_symbol_t *s, int lu) {
    double total = 0.0;
    for (int64_t c = 0; c < s->cblknbr; c++) {
        const double N = (double)width(s, c);
        double M = 0.0;
        for (int64_t b = s->cblk_bloknum[c] + 1; b < s->cblk_bloknum[c + 1]; b++) M += (double)brows(s, b);
        if (lu) {
            /* FLOPS_DGETRF(N,N) (m >= n branch) + 2 FLOPS_DTRSM(Right,M,N) + FLOPS_DGEMM(M,M,N) */
            const double mul = 0.5 * N * (N * (N - N / 3.0 - 1.0) + N) + (2.0 / 3.0) * N;
            const double add = 0.5 * N * (N * (N - N / 3.0) - N) + (1.0 / 6.0) * N;
            total += mul + add;
            if (M > 0) total += 2.0 * (M * N * (N + 1.0)) + 2.0 * M * M * N;
        } else {
            /* FLOPS_DPOTRF(N) + FLOPS_DTRSM(Right,M,N) + FLOPS_DSYRK(N,M) */
            total += N * ((N / 6.0 + 0.5) * N + 1.0 / 3.0) + N * ((N / 6.0) * N - 1.0 / 6.0);
            if (M > 0) total += M * N * (N + 1.0) + N * M * (M + 1.0);
        }
    }
    return total;
}

/* ------------------------------------------------------ refinement ---- */
static double dot(int64_t n, const double *a, const double *b) {
    double v = 0.0;
    for (int64_t i = 0; i < n; i++) v += a[i] * b[i];
    return v;
}

int ora_cg(const ora_symbol_t *s, const double *L, int64_t n, const int64_t *rowptr, const int64_t *cols,
           const double *vals, const double *b, double *x, int itermax, double eps, double *resid) {
    double *r = malloc(4 * n * sizeof(double));
    if (!r) return -1;
    double *p = r + n, *z = r + 2 * n, *Ap = r + 3 * n;
    const double normb = sqrt(dot(n, b, b)), normx = sqrt(dot(n, x, x));
    memcpy(r, b, n * sizeof(double));
    if (normx > 0.0) ora_spmv(n, rowptr, cols, vals, -1.0, x, 1.0, r);
    double rb = sqrt(dot(n, r, r)) / normb;
    memcpy(z, r, n * sizeof(double));
    ora_solve(s, 0, L, NULL, z, n, 1);
    memcpy(p, z, n * sizeof(double));
    int it = 0;
    while (rb > eps && it < itermax) {
        it++;
        ora_spmv(n, rowptr, cols, vals, 1.0, p, 0.0, Ap);
        const double rz = dot(n, r, z), alpha = rz / dot(n, Ap, p);
        for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        memcpy(z, r, n * sizeof(double));
        ora_solve(s, 0, L, NULL, z, n, 1);
        const double beta = dot(n, r, z) / rz;
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rb = sqrt(dot(n, r, r)) / normb;
    }
    free(r);
    if (resid) *resid = rb;
    return it;
}

/* Simple iterative refinement, reference refinement/z_refine_pivot.c:26-119.
 * Each pass: r = b - A x, berr = ||r||_2 / ||b||_2; while iter < itermax,
 * berr > eps and berr <= last_berr / 2 (the first pass takes last_berr =
 * 3 berr), x += M^{-1} r with M^{-1} the supernodal solve.  Returns the number
 * of passes, the stopping one included. */
int ora_sr(const ora_symbol_t *s, int lu, const double *L, const double *U, int64_t n, const int64_t *rowptr,
           const int64_t *cols, const double *vals, const double *b, double *x, int itermax, double eps,
           double *resid) {
    double *r = malloc(2 * n * sizeof(double));
    if (!r) return -1;
    double *dx = r + n;
    const double normb = sqrt(dot(n, b, b));
    double berr = 0.0, last = 0.0;
    int it = 0;
    for (;;) {
        memcpy(r, b, n * sizeof(double));
        ora_spmv(n, rowptr, cols, vals, -1.0, x, 1.0, r);
        berr = sqrt(dot(n, r, r)) / normb;
        if (it == 0) last = 3.0 * berr;
        const int go = it < itermax && berr > eps && berr <= last / 2.0;
        it++;
        if (!go) break;
        memcpy(dx, r, n * sizeof(double));
        ora_solve(s, lu, L, U, dx, n, 1);
        for (int64_t i = 0; i < n; i++) x[i] += dx[i];
        last = berr;
    }
    free(r);
    if (resid) *resid = berr;
    return it;
}

/* Preconditioned BiCGSTAB, reference refinement/z_refine_bicgstab.c:25-185:
 * y = M^{-1} p, v = A y, alpha = (r, r2) / (v, r2), s = r - alpha v,
 * z = M^{-1} s, t = A z, w = (M^{-1} t, z) / (M^{-1} t, M^{-1} t),
 * x += alpha y + w z, r = s - w t, beta = (r, r2) / (r_old, r2) * alpha / w,
 * p = r + beta (p - w v); stops when ||r|| / ||b|| <= eps or after itermax
 * iterations.  Returns the iterations. */
int ora_bicgstab(const ora_symbol_t *s, int lu, const double *L, const double *U, int64_t n, const int64_t *rowptr,
                 const int64_t *cols, const double *vals, const double *b, double *x, int itermax, double eps,
                 double *resid) {
    double *r = malloc(9 * n * sizeof(double));
    if (!r) return -1;
    double *r2 = r + n, *p = r2 + n, *y = p + n, *v = y + n, *sv = v + n, *z = sv + n, *t = z + n, *mt = t + n;
    const double normb = sqrt(dot(n, b, b)), normx = sqrt(dot(n, x, x));
    memcpy(r, b, n * sizeof(double));
    if (normx > 0.0) ora_spmv(n, rowptr, cols, vals, -1.0, x, 1.0, r);
    double rb = sqrt(dot(n, r, r)) / normb;
    memcpy(r2, r, n * sizeof(double));
    memcpy(p, r, n * sizeof(double));
    int it = 0;
    while (rb > eps && it < itermax) {
        it++;
        memcpy(y, p, n * sizeof(double));
        ora_solve(s, lu, L, U, y, n, 1);
        ora_spmv(n, rowptr, cols, vals, 1.0, y, 0.0, v);
        const double rr2 = dot(n, r, r2);
        const double alpha = rr2 / dot(n, v, r2);
        for (int64_t i = 0; i < n; i++) sv[i] = r[i] - alpha * v[i];
        memcpy(z, sv, n * sizeof(double));
        ora_solve(s, lu, L, U, z, n, 1);
        ora_spmv(n, rowptr, cols, vals, 1.0, z, 0.0, t);
        memcpy(mt, t, n * sizeof(double));
        ora_solve(s, lu, L, U, mt, n, 1);
        const double w = dot(n, mt, z) / dot(n, mt, mt);
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * y[i];
            x[i] += w * z[i];
            r[i] = sv[i] - w * t[i];
        }
        const double beta = dot(n, r, r2) / rr2 * (alpha / w);
        for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - w * v[i]);
        rb = sqrt(dot(n, r, r)) / normb;
    }
    free(r);
    if (resid) *resid = rb;
    return it;
}

/* Right-preconditioned restarted GMRES, reference refinement/z_refine_gmres.c:46-352
 * (Saad, Iterative Methods, 2nd ed., Alg. 9.5):
 *   cycle: v_0 = b - A x with the current x, resid =
 *          ||v_0||; leave when resid / ||b|| <= eps (:190-205); g_0 = resid;
 *   step i: w_i = M^{-1} v_i, v_{i+1} = A w_i, then for j = 0..i in turn
 *          h_{j,i} = (v_{i+1}, v_j) and v_{i+1} -= h_{j,i} v_j (:240-249);
 *          h_{i+1,i} = ||v_{i+1}||, v_{i+1} /= h_{i+1,i} only when it exceeds
 *          1e-50 (:252-260); the earlier rotations on column i (:263-273); the
 *          new rotation from t = sqrt(h_ii^2 + h_{i+1,i}^2), t replaced by eps
 *          when t <= eps (:281-289); g_{i+1} = -s_i g_i, g_i = c_i g_i;
 *          resid = |g_{i+1}|; the step count grows; the cycle ends at i+1 = im,
 *          resid / ||b|| <= eps or the step limit (:300-309);
 *   end of cycle: y = H^{-1} g (upper triangular, i+1 columns), x += W y
 *          (:338-348); leave when the step count reached itermax (:357-360).
 * Returns the step count; resid gets the last resid / ||b||. */
int ora_gmres(const ora_symbol_t *s, int lu, const double *L, const double *U, int64_t n, const int64_t *rowptr,
              const int64_t *cols, const double *vals, const double *b, double *x, int itermax, int im, double eps,
              double *resid) {
    const int im1 = im + 1;
    double *V = malloc((size_t)n * im1 * sizeof(double));
    double *W = malloc((size_t)n * im * sizeof(double));
    double *H = calloc((size_t)im * im1, sizeof(double));
    double *cs = malloc(im * sizeof(double)), *sn = malloc(im * sizeof(double)), *g = malloc(im1 * sizeof(double));
    if (!V || !W || !H || !cs || !sn || !g) return -1;
    const double normb = sqrt(dot(n, b, b));
    double res = 0.0, rb = 0.0;
    int iters = 0;
    for (;;) {
        /* the reference tests ||x|| once, before the first cycle (:145,185),
         * and with a zero initial guess would take r = b on every restart;
         * the product is taken with the current x on every cycle (the same
         * vector on the first cycle, since b - A 0 = b exactly) */
        memcpy(V, b, n * sizeof(double));
        ora_spmv(n, rowptr, cols, vals, -1.0, x, 1.0, V);
        res = sqrt(dot(n, V, V));
        rb = res / normb;
        if (rb <= eps) break;
        for (int64_t k = 0; k < n; k++) V[k] *= 1.0 / res;
        g[0] = res;
        int i = -1;
        for (;;) {
            i++;
            double *Hi = H + (size_t)i * im1;
            double *vi = V + (size_t)i * n, *wi = W + (size_t)i * n, *vn = V + (size_t)(i + 1) * n;
            memcpy(wi, vi, n * sizeof(double));
            ora_solve(s, lu, L, U, wi, n, 1);
            ora_spmv(n, rowptr, cols, vals, 1.0, wi, 0.0, vn);
            for (int j = 0; j <= i; j++) {
                const double *vj = V + (size_t)j * n;
                Hi[j] = dot(n, vn, vj);
                for (int64_t k = 0; k < n; k++) vn[k] += -1.0 * Hi[j] * vj[k];
            }
            const double nrm = sqrt(dot(n, vn, vn));
            Hi[i + 1] = nrm;
            if (nrm > 1e-50)
                for (int64_t k = 0; k < n; k++) vn[k] *= 1.0 / nrm;
            for (int j = 0; j < i; j++) {
                const double t = Hi[j];
                Hi[j] = cs[j] * t + sn[j] * Hi[j + 1];
                Hi[j + 1] = cs[j] * Hi[j + 1] - sn[j] * t;
            }
            double t = sqrt(Hi[i] * Hi[i] + Hi[i + 1] * Hi[i + 1]);
            if (fabs(t) <= eps) t = eps;
            cs[i] = Hi[i] / t;
            sn[i] = Hi[i + 1] / t;
            g[i + 1] = -sn[i] * g[i];
            g[i] = cs[i] * g[i];
            Hi[i] = cs[i] * Hi[i] + sn[i] * Hi[i + 1];
            res = fabs(g[i + 1]);
            rb = res / normb;
            iters++;
            if (i + 1 >= im || rb <= eps || iters >= itermax) break;
        }
        /* y = H^{-1} g: upper triangular (i+1) x (i+1), ld im1 (dtrsv) */
        for (int k = i; k >= 0; k--) {
            double v = g[k];
            for (int l = k + 1; l <= i; l++) v -= H[k + (size_t)l * im1] * g[l];
            g[k] = v / H[k + (size_t)k * im1];
        }
        for (int l = 0; l <= i; l++) {
            const double *wl = W + (size_t)l * n;
            for (int64_t k = 0; k < n; k++) x[k] += g[l] * wl[k];
        }
        if (iters >= itermax) break;
    }
    free(V); free(W); free(H); free(cs); free(sn); free(g);
    if (resid) *resid = rb;
    return iters;
}

/* ------------------------------------------------ static scheduler --- */
/* Restatement of the reference static scheduler (sopalin/sequential_zpotrf.c:
 * 71-120 thread_zpotrf_static, sequential_zgetrf.c the same shape) on
 * nthreads POSIX threads:
 *   - ctrbcnt[c] starts at the number of blocks facing c (blend/
 *     solver_backup.c:95-101);
 *   - each thread walks its own task list (one task per cblk, increasing
 *     cblk order); before a task it waits until ctrbcnt reaches zero
 *     (cpucblk_zincoming_deps, kernels/cpucblk_zmpi_coeftab.c:482-515),
 *     factors the panel (diagonal block + TRSM), then for every off-diagonal
 *     block applies the Schur update into the facing cblk under that cblk's
 *     lock (cpucblk_zgemmsp with pastix_cblk_lock) and decrements its counter
 *     (cpucblk_zrelease_deps :541-562).
 * The task lists come from a proportional mapping of the cblk tree (parent =
 * the cblk facing the first off-diagonal block) by flop cost, standing in for
 * blend's proportional mapping and simulation (blend/propmap.c,
 * blend/simu_run.c): a subtree whose share falls to one thread goes to that
 * thread whole; a cblk shared by several candidates goes to the candidate
 * with the least work so far.  Every task waits only on lower-numbered cblks,
 * so the threads cannot deadlock; a failing panel (negative pivot in LL^t)
 * raises a flag that ends every wait. */
#include <pthread.h>
#include <sched.h>

typedef struct {
    const ora_symbol_t *s;
    double *L, *U;
    double crit;
    int lu;
    int64_t **tasks, *ntasks;
    int32_t *ctrb;
    pthread_mutex_t *locks;
    int64_t *npiv;
    volatile int err;
    int rank;
} static_ctx_t;

typedef struct {
    static_ctx_t *g;
    int rank;
} static_arg_t;

static double cblk_cost(const ora_symbol_t *s, int64_t c, int lu) {
    const double N = (double)width(s, c);
    double M = 0.0;
    for (int64_t b = s->cblk_bloknum[c] + 1; b < s->cblk_bloknum[c + 1]; b++) M += (double)brows(s, b);
    return (lu ? 2.0 : 1.0) * (N * N * N / 3.0 + M * N * N + M * M * N) + 1.0;
}

/* thread ranges [t0, t1) of the subtree rooted at c, children first */
static void propmap(const ora_symbol_t *s, int64_t c, double t0, double t1, const int64_t *cptr,
                    const int64_t *clist, const double *scost, const double *ccost, double *load, int nthreads,
                    int64_t *owner) {
    const int a = (int)t0, b = (int)(t1 - 1e-9);
    if (b <= a) {
        /* the whole subtree on one thread: iterate it explicitly */
        int64_t stack_sz = 1, *stack = malloc((size_t)(s->cblknbr + 1) * sizeof(int64_t));
        stack[0] = c;
        while (stack_sz) {
            const int64_t x = stack[--stack_sz];
            owner[x] = a < nthreads ? a : nthreads - 1;
            load[owner[x]] += ccost[x];
            for (int64_t k = cptr[x]; k < cptr[x + 1]; k++) stack[stack_sz++] = clist[k];
        }
        free(stack);
        return;
    }
    int best = a;
    for (int t = a + 1; t <= b && t < nthreads; t++)
        if (load[t] < load[best]) best = t;
    owner[c] = best;
    load[best] += ccost[c];
    const double sub = scost[c] - ccost[c];
    double at = t0;
    for (int64_t k = cptr[c]; k < cptr[c + 1]; k++) {
        const int64_t ch = clist[k];
        const double w = sub > 0 ? (t1 - t0) * scost[ch] / sub : 0.0;
        propmap(s, ch, at, at + w, cptr, clist, scost, ccost, load, nthreads, owner);
        at += w;
    }
}

static void *static_worker(void *argp) {
    static_arg_t *arg = (static_arg_t *)argp;
    static_ctx_t *g = arg->g;
    const ora_symbol_t *s = g->s;
    const int rank = arg->rank;
    int64_t npiv = 0;
    for (int64_t ii = 0; ii < g->ntasks[rank]; ii++) {
        const int64_t c = g->tasks[rank][ii];
        /* incoming dependencies */
        while (__atomic_load_n(&g->ctrb[c], __ATOMIC_ACQUIRE) > 0) {
            if (g->err) return NULL;
            sched_yield();
        }
        if (g->err) return NULL;
        const int64_t K = width(s, c), st = s->cblk_stride[c];
        double *P = g->L + s->cblk_coef[c];
        if (!g->lu) {
            if (potrfsp(K, P, st, &npiv, g->crit)) {
                g->err = 1;
                return NULL;
            }
            if (st > K) {
                if (blas_dtrsm)
                    blas_dtrsm(CB_COL, CB_RIGHT, CB_LOWER, CB_TRANS, CB_NONUNIT, (int)(st - K), (int)K, 1.0, P,
                               (int)st, P + K, (int)st);
                else
                    trsm_rltn(st - K, K, P, st, P + K, st);
            }
        } else {
            double *Q = g->U + s->cblk_coef[c];
            for (int64_t j = 0; j < K; j++)
                for (int64_t i = 0; i < K; i++) P[i + j * st] += Q[j + i * st];
            getrfsp(K, P, st, &npiv, g->crit);
            for (int64_t j = 0; j < K; j++)
                for (int64_t i = 0; i < K; i++) Q[i + j * st] = P[j + i * st];
            if (st > K) {
                if (blas_dtrsm) {
                    blas_dtrsm(CB_COL, CB_RIGHT, CB_UPPER, CB_NOTRANS, CB_NONUNIT, (int)(st - K), (int)K, 1.0, P,
                               (int)st, P + K, (int)st);
                    blas_dtrsm(CB_COL, CB_RIGHT, CB_UPPER, CB_NOTRANS, CB_UNIT, (int)(st - K), (int)K, 1.0, Q,
                               (int)st, Q + K, (int)st);
                } else {
                    trsm_run(st - K, K, P, st, 0, P + K, st);
                    trsm_run(st - K, K, Q, st, 1, Q + K, st);
                }
            }
        }
        for (int64_t b = s->cblk_bloknum[c] + 1; b < s->cblk_bloknum[c + 1]; b++) {
            const int64_t f = s->blok_fcblk[b];
            pthread_mutex_lock(&g->locks[f]);
            if (!g->lu) {
                gemmsp_x(s, 0, c, b, g->L, g->L, g->L, 1);
            } else {
                gemmsp_x(s, 0, c, b, g->L, g->U, g->L, 1);
                if (b + 1 < s->cblk_bloknum[c + 1]) gemmsp_x(s, 1, c, b, g->U, g->L, g->U, 1);
            }
            pthread_mutex_unlock(&g->locks[f]);
            __atomic_sub_fetch(&g->ctrb[f], 1, __ATOMIC_RELEASE);
        }
    }
    g->npiv[rank] = npiv;
    return NULL;
}

int ora_factor_static(const ora_symbol_t *s, int lu, double *L, double *U, double crit, int nthreads,
                      int64_t *npiv) {
    const int64_t cb = s->cblknbr;
    if (nthreads < 1) nthreads = 1;
    *npiv = 0;
    /* cblk tree: parent = cblk facing the first off-diagonal block */
    int64_t *parent = malloc(cb * sizeof(int64_t)), *cptr = calloc(cb + 2, sizeof(int64_t));
    int64_t *clist = malloc((cb + 1) * sizeof(int64_t)), *owner = malloc(cb * sizeof(int64_t));
    double *ccost = malloc(cb * sizeof(double)), *scost = malloc(cb * sizeof(double));
    double *load = calloc(nthreads, sizeof(double));
    int32_t *ctrb = calloc(cb, sizeof(int32_t));
    pthread_mutex_t *locks = malloc(cb * sizeof(pthread_mutex_t));
    int64_t *ntasks = calloc(nthreads, sizeof(int64_t)), **tasks = malloc(nthreads * sizeof(int64_t *));
    int64_t *tnpiv = calloc(nthreads, sizeof(int64_t));
    for (int64_t c = 0; c < cb; c++) {
        const int64_t b = s->cblk_bloknum[c] + 1;
        parent[c] = (b < s->cblk_bloknum[c + 1]) ? s->blok_fcblk[b] : -1;
        ccost[c] = cblk_cost(s, c, lu);
        scost[c] = ccost[c];
        pthread_mutex_init(&locks[c], NULL);
        for (int64_t k = s->cblk_bloknum[c] + 1; k < s->cblk_bloknum[c + 1]; k++) ctrb[s->blok_fcblk[k]]++;
    }
    for (int64_t c = 0; c < cb; c++) /* parents have higher numbers: accumulate upward */
        if (parent[c] >= 0) scost[parent[c]] += scost[c];
    for (int64_t c = 0; c < cb; c++) cptr[(parent[c] >= 0 ? parent[c] : cb) + 1]++;
    for (int64_t c = 0; c <= cb; c++) cptr[c + 1] += cptr[c];
    {
        int64_t *next = malloc((cb + 1) * sizeof(int64_t));
        for (int64_t c = 0; c <= cb; c++) next[c] = cptr[c];
        for (int64_t c = 0; c < cb; c++) clist[next[parent[c] >= 0 ? parent[c] : cb]++] = c;
        free(next);
    }
    /* roots share the threads by cost */
    double rtot = 0.0;
    for (int64_t k = cptr[cb]; k < cptr[cb + 1]; k++) rtot += scost[clist[k]];
    double at = 0.0;
    for (int64_t k = cptr[cb]; k < cptr[cb + 1]; k++) {
        const double w = rtot > 0 ? nthreads * scost[clist[k]] / rtot : 0.0;
        propmap(s, clist[k], at, at + w, cptr, clist, scost, ccost, load, nthreads, owner);
        at += w;
    }
    for (int64_t c = 0; c < cb; c++) ntasks[owner[c]]++;
    for (int t = 0; t < nthreads; t++) {
        tasks[t] = malloc((ntasks[t] + 1) * sizeof(int64_t));
        ntasks[t] = 0;
    }
    for (int64_t c = 0; c < cb; c++) tasks[owner[c]][ntasks[owner[c]]++] = c;
    static_ctx_t g = {s, L, U, crit, lu, tasks, ntasks, ctrb, locks, tnpiv, 0, 0};
    pthread_t *th = malloc(nthreads * sizeof(pthread_t));
    static_arg_t *args = malloc(nthreads * sizeof(static_arg_t));
    for (int t = 0; t < nthreads; t++) {
        args[t].g = &g;
        args[t].rank = t;
        pthread_create(&th[t], NULL, static_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    if (getenv("ORA_STATIC_DEBUG"))
        for (int t = 0; t < nthreads; t++) fprintf(stderr, "thread %d: %ld tasks, cost %.3e\n", t, (long)ntasks[t], load[t]);
    for (int t = 0; t < nthreads; t++) *npiv += tnpiv[t];
    const int rc = g.err ? -1 : 0;
    for (int64_t c = 0; c < cb; c++) pthread_mutex_destroy(&locks[c]);
    for (int t = 0; t < nthreads; t++) free(tasks[t]);
    free(th); free(args); free(tasks); free(ntasks); free(tnpiv); free(locks); free(ctrb); free(load);
    free(scost); free(ccost); free(owner); free(clist); free(cptr); free(parent);
    return rc;
}
