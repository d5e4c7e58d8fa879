/* CPU oracle of the MI355X engine (test infrastructure only, see oracle.c). */
#ifndef PASTIX_AMD_ORACLE_H
#define PASTIX_AMD_ORACLE_H
#include <stdint.h>

typedef struct {
    int64_t n, cblknbr, bloknbr, coefnbr;
    const int64_t *cblk_fcol;    /* cblknbr+1 */
    const int64_t *cblk_bloknum; /* cblknbr+1 */
    const int64_t *cblk_stride;  /* cblknbr   */
    const int64_t *cblk_coef;    /* cblknbr+1 */
    const int64_t *blok_frow, *blok_lrow, *blok_fcblk, *blok_roff; /* bloknbr */
} ora_symbol_t;

int    ora_fill(const ora_symbol_t *s, const int64_t *colptr, const int64_t *rows, const double *vals,
                const int64_t *tcolptr, const int64_t *trows, const double *tvals, double *L, double *U);
int    ora_potrf(const ora_symbol_t *s, double *L, double crit, int64_t *npiv);
int    ora_getrf(const ora_symbol_t *s, double *L, double *U, double crit, int64_t *npiv);
int    ora_solve(const ora_symbol_t *s, int lu, const double *L, const double *U, double *x, int64_t ldx, int nrhs);
void   ora_spmv(int64_t n, const int64_t *rowptr, const int64_t *cols, const double *vals, double alpha,
                const double *x, double beta, double *y);
double ora_thflops(const ora_symbol_t *s, int lu);
int    ora_cg(const ora_symbol_t *s, const double *L, int64_t n, const int64_t *rowptr, const int64_t *cols,
              const double *vals, const double *b, double *x, int itermax, double eps, double *resid);
int    ora_sr(const ora_symbol_t *s, int lu, const double *L, const double *U, int64_t n, const int64_t *rowptr,
              const int64_t *cols, const double *vals, const double *b, double *x, int itermax, double eps,
              double *resid);
int    ora_bicgstab(const ora_symbol_t *s, int lu, const double *L, const double *U, int64_t n,
                    const int64_t *rowptr, const int64_t *cols, const double *vals, const double *b, double *x,
                    int itermax, double eps, double *resid);
int    ora_gmres(const ora_symbol_t *s, int lu, const double *L, const double *U, int64_t n,
                 const int64_t *rowptr, const int64_t *cols, const double *vals, const double *b, double *x,
                 int itermax, int im, double eps, double *resid);
/* load cblas_dgemm / cblas_dtrsm (names prefixed by prefix) from a shared
 * library for ora_factor_static; 0 on success */
int    ora_use_blas(const char *path, const char *prefix);
/* the reference static scheduler on nthreads POSIX threads (lu: GETRF with U) */
int    ora_factor_static(const ora_symbol_t *s, int lu, double *L, double *U, double crit, int nthreads,
                         int64_t *npiv);
#endif
