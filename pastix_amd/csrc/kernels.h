// Host-callable launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "dev_types.h"

namespace pxa {

// direct: every item has exactly one entry (no LDS accumulation).
template <typename T>
void launch_update(const DevSymbol &S, int mode, bool direct, const UpdItem *items, int n, const UpdEntry *ents,
                   const int32_t *shift, const T *A, const T *B, T *C, hipStream_t st);
// Two-phase update of the multi-source tiles: dense product tiles into the
// level workspace W, then the per-target-tile ordered sum.
template <typename T>
void launch_update_wgemm(const DevSymbol &S, const WTile *items, int n, const T *A, const T *B, T *W, hipStream_t st);
template <typename T>
void launch_update_gather(const DevSymbol &S, int mode, const UpdItem *items, int n, const UpdEntry *ents,
                          const int32_t *shift, const WGroup *grp, const T *W, T *C, hipStream_t st);

// Blocked panel factorisation (64-column steps).
template <typename T>
void launch_potrf64(const DevSymbol &S, const int32_t *cblks, int n, int kb, T *coef, T *winv, T *winvt, double thr,
                    int *stat, hipStream_t st);
template <typename T>
void launch_panel_trsm(const DevSymbol &S, const PanelItem *items, int n, int kb, T *coef, const T *winv,
                       hipStream_t st);
// Both sides of an LU step's TRSM in one launch: (coef, winv), (coef2, winv2)
template <typename T>
void launch_panel_trsm2(const DevSymbol &S, const PanelItem *items, int n, int kb, T *coef, const T *winv, T *coef2,
                        const T *winv2, hipStream_t st);
template <typename T>
void launch_panel_upd(const DevSymbol &S, const PanelItem *items, int n, int kb, bool right, T *C, const T *X,
                      const T *Y, bool strict, hipStream_t st);
// Both sides of an LU step's in-cblk update in one launch: (C, X, Y) = (L, L,
// U) and (U, U, L), the U side strict
template <typename T>
void launch_panel_upd_lu(const DevSymbol &S, const PanelItem *items, int n, int kb, bool right, T *L, T *U,
                         hipStream_t st);
// An LL^t step's in-cblk update and the diagonal factorisation waiting on it
// in one launch (right-looking: update of step kb, factorisation of block
// kb+1; left-looking: update and factorisation of block kb).
template <typename T>
void launch_panel_upd_potrf(const DevSymbol &S, const int32_t *cblks, int npot, const PanelItem *items, int n, int kb,
                            bool right, T *coef, T *winv, T *winvt, double thr, int *stat, hipStream_t st);
// LU: diagonal 64-block with inverses; final layout of the square parts.
template <typename T>
void launch_getrf64(const DevSymbol &S, const int32_t *cblks, int n, int kb, T *lcoef, T *ucoef, T *w1, T *w2, T *wb,
                    double thr, int *stat, hipStream_t st);
template <typename T>
void launch_lu_square_merge(const DevSymbol &S, const int32_t *cblks, int n, T *lcoef, T *ucoef, hipStream_t st);
template <typename T>
void launch_lu_square_layout(const DevSymbol &S, const int32_t *cblks, int n, T *lcoef, T *ucoef, hipStream_t st);
// Inverses of the solve nodes' diagonal blocks from the 64-column inverses
// the diagonal-block kernels leave in w (and wt): the leaves of nodes of
// several cblks copied into the nodes' inverses, then nrounds merge rounds,
// round r the tiles [round_ptr[r], round_ptr[r+1]) (round_ptr a host array),
// reading the factor in coef.  wt may be null (no transposed copy); tbuf
// holds the largest round's products.
template <typename T>
void launch_winv_nodes(const DevSymbol &S, const WinvLeaf *leaves, int nleaves, const WinvTile *tiles,
                       const int32_t *round_ptr, int nrounds, const T *coef, T *w, T *wt, T *tbuf, hipStream_t st);

// Supernodal triangular solves (reference kernels/solve_ztrsmsp.c:332-557).
// nrhs <= 8 per launch; the forward sweep writes z, the backward sweep x.
// On levels of wide nodes the forward diagonal step leaves b minus the
// pending products in b's node rows, the backward one z minus the
// partials in z.
template <typename T>
void launch_fwd_diag(const DevSymbol &S, const DiagItem *items, int n, const YSlices &ys, const T *ybuf,
                     int64_t ldy, T *b, int64_t ldb, T *z, int64_t ldz, int nrhs, const T *w1, const T *w1t,
                     int maxwidth, hipStream_t st);
template <typename T>
void launch_fwd_prod(const DevSymbol &S, const FwdItem *items, int n, const T *coef, const T *z, int64_t ldz,
                     int nrhs, const int32_t *ydst, T *ybuf, int64_t ldy, int maxwidth, hipStream_t st);
template <typename T>
void launch_bwd_prod(const DevSymbol &S, const BwdItem *items, int n, const T *coef, const T *x, int64_t ldx,
                     int nrhs, T *part, int maxwidth, hipStream_t st);
template <typename T>
void launch_bwd_diag(const DevSymbol &S, const DiagItem *items, int n, const BwdItem *bitems, T *z,
                     int64_t ldz, T *x, int64_t ldx, int nrhs, const T *part, const T *wb, int maxwidth,
                     hipStream_t st);

// Sparse matrix-vector product on the permuted block CSC
// (reference bcsc/bcsc_zspmv.c:55-202): y = alpha*A*x + beta*y.
template <typename T>
void launch_spmv(int64_t n, const int64_t *colptr, const int32_t *rows, const T *vals, T alpha, const T *x, T beta,
                 T *y, hipStream_t st);

// BLAS-1 on device vectors (reference bcsc/bvec_zcompute.c).
void launch_dot(int64_t n, const double *x, const double *y, double *partials, double *out, hipStream_t st);
void launch_nrm2sq(int64_t n, const double *x, double *partials, double *out, hipStream_t st);
void launch_axpy(int64_t n, double alpha, const double *x, double *y, hipStream_t st);
void launch_axpy_dev(int64_t n, const double *alpha, double sign, const double *x, double *y, hipStream_t st);
void launch_scal(int64_t n, double alpha, double *x, hipStream_t st);
void launch_copy(int64_t n, const double *x, double *y, hipStream_t st);
// Schur rows [r0, r1) of a solve: mode 0 z = x, 1 z = x - pending forward
// products, 2 x = 0, 3 x = z
template <typename T>
void launch_schur_rows(int64_t r0, int64_t r1, int nrhs, int mode, const YSlices &ys, const T *ybuf, int64_t ldy,
                       T *x, int64_t ldx, T *z, int64_t ldz, hipStream_t st);
// out[0] = ||x||_2 by scaled sums of squares (partials: 2 * 512 doubles)
void launch_nrm2(int64_t n, const double *x, double *partials, double *out, hipStream_t st);
// y = alpha A x + beta y, A m x n column-major
void launch_gemv(int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x, double beta,
                 double *y, hipStream_t st);
// y = alpha A^T x + beta y from the transposed pattern of the CSR of A
void launch_spmv_t(int64_t n, const int64_t *trowptr, const int32_t *tcols, const int64_t *tpos, const double *vals,
                   double alpha, const double *x, double beta, double *y, hipStream_t st);
// Segmented BLAS-1 over an n x R block whose rows are split into nseg
// consecutive segments (off: nseg+1 offsets, seg: segment of every row);
// per-(segment, column) scalars are stored at [s + c*nseg].
// chunk length of the segmented dot products (entries per first-pass block)
int seg_dot_chunk();
// out[s + c*nseg] = x[:,c] . y[:,c] over rows [off[s], off[s+1]); the chunks
// of segment s are [coff[s], coff[s+1]) (chunk k in segment cseg[k], nchunk
// chunks in all, part: nchunk x R scratch)
void launch_seg_dot(const int64_t *off, const int32_t *coff, const int32_t *cseg, int nseg, int nchunk, int R,
                    int64_t ld, const double *x, const double *y, double *part, double *out, hipStream_t st);
void launch_seg_axpy(int64_t n, int R, int64_t ld, const int32_t *seg, int nseg, const double *a, double sign,
                     const double *x, double *y, hipStream_t st);
void launch_seg_xpby(int64_t n, int R, int64_t ld, const int32_t *seg, int nseg, const double *b, const double *z,
                     double *p, hipStream_t st);
// coef[pos[i]] += delta (diagonal of the coefficient table);
// vals[pos[i]] = d0[i] + delta where pos[i] >= 0 (diagonal of the CSR).
template <typename T>
void launch_add_diag(int64_t n, const int64_t *pos, double delta, T *coef, hipStream_t st);
void launch_set_csr_diag(int64_t n, const int64_t *pos, const double *d0, double delta, double *vals, hipStream_t st);
// Zeroes n ints with a kernel (usable inside a captured graph).
// out[i] = coef[pos[i]] for i < n (the factor's diagonal, as FP64)
template <typename T> void launch_gather_diag(int64_t n, const int64_t *pos, const T *coef, double *out, hipStream_t st);
// Coefficient tables (zeroed beforehand) from the FP64 CSR values of the
// permuted matrix: L[cmap[e]] for cmap[e] >= 0, U[-2 - cmap[e]] for
// cmap[e] <= -2 (when U is given), nothing at -1.
template <typename T>
void launch_fill_coef(int64_t nnz, const int64_t *cmap, const double *vals, T *L, T *U, hipStream_t st);
template <typename T> void launch_copy2d(int64_t n, int nr, const T *x, int64_t ldx, T *y, int64_t ldy, hipStream_t st);
template <typename T>
void launch_diag_scale(int64_t n, int nr, const int64_t *pos, const T *coef, bool divide, T *x, int64_t ldx,
                       hipStream_t st);
void launch_zero_int(int n, int *p, hipStream_t st);
void launch_d2s(int64_t n, const double *x, float *y, hipStream_t st);
void launch_s2d(int64_t n, const float *x, double *y, hipStream_t st);
void launch_gemv_t(int64_t n, int m, const double *V, int64_t ldv, const double *x, double *partials, double *out,
                   hipStream_t st);
void launch_gemv_n(int64_t n, int m, const double *V, int64_t ldv, const double *y, double alpha, double *x,
                   hipStream_t st);
void launch_berr_parts(int64_t n, const int64_t *colptr, const int32_t *rows, const double *vals, const double *x,
                       const double *b, double *partials, double *out, hipStream_t st);
// Leaf region: counters and queue set up, then the persistent task launch
// (grid workgroups at most; a wait gives up after maxwait ticks of the
// device's constant-rate wall clock and sets stat[2]).
void launch_leaf_init(const LeafDev &LP, int32_t *ctr, int32_t *queue, int32_t *sync, hipStream_t st);
template <typename T>
void launch_leaf_tasks(const DevSymbol &S, const LeafDev &LP, const UpdEntry *ents, const int32_t *shift, T *coef,
                       T *winv, T *winvt, double thr, int *stat, int32_t *ctr, int32_t *queue, int32_t *sync,
                       int grid, long long maxwait, hipStream_t st);

// Backward error without allocation: partial maxima in a library-static
// device array.  NOALLOC_MAX_RHS columns at most; noalloc_scratch_init
// resolves the array's address on the current device (call it once per
// device with that device current), noalloc_scratch returns it (nullptr
// before the init).  partials: NOALLOC_SCRATCH_DOUBLES doubles.
constexpr int NOALLOC_MAX_RHS = 64;
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)512 * 3 * NOALLOC_MAX_RHS;
int noalloc_scratch_init(int device);
double *noalloc_scratch(int device);
void launch_berr_user(int64_t n, const int64_t *rowptr, const int32_t *cols, const double *vals,
                      const int32_t *peri, const double *x, int64_t ldx, const double *b, int64_t ldb, int nrhs,
                      double normInf, double *partials, double *berr, hipStream_t st);
// Permute rows of a column-major m x nrhs block: y[perm[i]] = x[i] (scatter) or y[i] = x[perm[i]] (gather).
template <typename T>
void launch_permute(int64_t m, int nrhs, const int32_t *perm, bool scatter, const T *x, int64_t ldx, T *y,
                    int64_t ldy, hipStream_t st);

}  // namespace pxa
