// Numerical factorisation kernels for CDNA4 (gfx950).
//
//  k_potrf_diag  : Cholesky of the diagonal block of every cblk of a level,
//                  64-column blocking with static pivoting, plus the explicit
//                  inverse of the lower-triangular diagonal block
//                  (reference kernels/core_zpotrfsp.c:66-189).
//  k_getrf_diag  : LU without pivoting of the diagonal block, static pivoting,
//                  explicit inverses of unit-L and of U
//                  (reference kernels/core_zgetrfsp.c:63-259, core_zgetmo.c).
//  k_trsm_panel  : off-diagonal panel solve as a GEMM against the inverted
//                  diagonal block (reference kernels/core_ztrsmsp.c:147-181).
// The Schur-complement update lives in kernels_update.hip and the blocked
// panel factorisation used for LL^t in kernels_panel.hip.
#include "dev_common.h"
#include "dev_types.h"
#include "kernels.h"

namespace pxa {

#define QUAD_FOR(...)                                                    \
    for (int m_ = 0; m_ < 2; m_++)                                       \
        for (int n_ = 0; n_ < 2; n_++)                                   \
            for (int r_ = 0; r_ < 4; r_++) {                             \
                const int qrow = m_ * 16 + Mfma<T>::drow(lane, r_);       \
                const int qcol = n_ * 16 + (lane & 15);                   \
                __VA_ARGS__                                              \
            }

template <typename T> __device__ __forceinline__ T tabs(T v) { return v < (T)0 ? -v : v; }

// In-place inverse of the n x n lower-triangular matrix held in Sm (LAPACK
// trti2 order: columns from last to first).  UNIT: implicit unit diagonal.
template <typename T, bool UNIT>
__device__ void lds_trtri_lower(T (*Sm)[65], T *vec, int n, int tid) {
    for (int j = n - 1; j >= 0; j--) {
        T a = UNIT ? (T)1 : Sm[j][j];
        if (tid > j && tid < n) {
            T s = 0;
            for (int k = j + 1; k <= tid; k++) s += (UNIT && k == tid ? (T)1 : Sm[tid][k]) * Sm[k][j];
            vec[tid] = s;
        }
        __syncthreads();
        T ia = (T)1 / a;
        if (tid > j && tid < n) Sm[tid][j] = -ia * vec[tid];
        if (!UNIT && tid == j) Sm[j][j] = ia;
        __syncthreads();
    }
}

// In-place inverse of the n x n upper-triangular matrix held in Sm.
template <typename T>
__device__ void lds_trtri_upper(T (*Sm)[65], T *vec, int n, int tid) {
    for (int j = 0; j < n; j++) {
        // column j above the diagonal: U = [[U11 u],[0 a]] -> [[U11^-1, -U11^-1 u / a],[0, 1/a]]
        T a = Sm[j][j];
        if (tid < j) {
            T s = 0;
            for (int k = tid; k < j; k++) s += Sm[tid][k] * Sm[k][j];
            vec[tid] = s;
        }
        __syncthreads();
        T ia = (T)1 / a;
        if (tid < j) Sm[tid][j] = -ia * vec[tid];
        if (tid == j) Sm[j][j] = ia;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Off-diagonal blocks of the inverse of a lower-triangular K x K matrix whose
// diagonal 64-blocks of the inverse are already in W (ld = K):
//   W[ib,jb] = -W[ib,ib] * sum_{kb=jb}^{ib-1} L[ib,kb] W[kb,jb]
// L(r,k) is read through the view Lv (absolute row/col indices).
// ---------------------------------------------------------------------------
template <typename T>
__device__ void wg_lower_inverse_offdiag(View<T> Lv, T *W, int K, T (*Sm)[65], int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    for (int jb0 = 0; jb0 < K; jb0 += 64) {
        const int nj = min(64, K - jb0);
        for (int ib0 = jb0 + 64; ib0 < K; ib0 += 64) {
            const int ni = min(64, K - ib0);
            Quad<T> q;
            q.zero();
            if (wr < ni && wc < nj) {
                View<T> A{Lv.p + (int64_t)(ib0 + wr) * Lv.rs + (int64_t)jb0 * Lv.cs, Lv.rs, Lv.cs};
                View<T> B{W + jb0 + (int64_t)(jb0 + wc) * K, K, 1};
                quad_gemm_nt(q, A, ni - wr, B, nj - wc, ib0 - jb0, lane);
            }
            QUAD_FOR(Sm[wr + qrow][wc + qcol] = q.c[m_][n_][r_];)
            __syncthreads();
            Quad<T> q2;
            q2.zero();
            if (wr < ni && wc < nj) {
                View<T> A{W + (ib0 + wr) + (int64_t)ib0 * K, 1, K};
                View<T> B{&Sm[0][wc], 1, 65};
                quad_gemm_nt(q2, A, ni - wr, B, nj - wc, ni, lane);
                QUAD_FOR(if (qrow < ni - wr && qcol < nj - wc)
                             W[(ib0 + wr + qrow) + (int64_t)(jb0 + wc + qcol) * K] = -q2.c[m_][n_][r_];)
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal Cholesky + inverse.  One workgroup per cblk of the level.
// stat[0] += static pivots, stat[1] |= 1 on a negative diagonal.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(DevSymbol S, const int32_t *__restrict__ cblks, T *__restrict__ coef,
                                                    T *__restrict__ winv, double thr, int *stat) {
    __shared__ T Sm[64][65];
    __shared__ T vec[64];
    __shared__ int npiv_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    T *P = coef + S.cblk_coef[c];
    T *W = winv + S.cblk_winv[c];
    const T crit = (T)thr;
    if (tid == 0) npiv_s = 0;
    for (int k0 = 0; k0 < K; k0 += 64) {
        const int nb = min(64, K - k0);
        for (int e = tid; e < 64 * 64; e += 256) {
            int i = e & 63, j = e >> 6;
            T v = 0;
            if (i < nb && j < nb && i >= j) v = P[(k0 + i) + (int64_t)(k0 + j) * ld];
            Sm[i][j] = v;
        }
        __syncthreads();
        // unblocked Cholesky with static pivoting (core_zpotf2sp)
        for (int k = 0; k < nb; k++) {
            if (tid == 0) {
                T d = Sm[k][k];
                if (tabs(d) < crit) {
                    d = crit;
                    npiv_s++;
                }
                if (d < (T)0) {
                    atomicOr(&stat[1], 1);
                    d = -d;
                }
                Sm[k][k] = sqrt(d);
            }
            __syncthreads();
            const T alpha = (T)1 / Sm[k][k];
            if (tid > k && tid < nb) Sm[tid][k] *= alpha;
            __syncthreads();
            const int m = nb - k - 1;
            for (int e = tid; e < m * m; e += 256) {
                int ii = e % m, jj = e / m;
                if (ii >= jj) Sm[k + 1 + ii][k + 1 + jj] -= Sm[k + 1 + ii][k] * Sm[k + 1 + jj][k];
            }
            __syncthreads();
        }
        for (int e = tid; e < 64 * 64; e += 256) {
            int i = e & 63, j = e >> 6;
            if (i < nb && j < nb && i >= j) P[(k0 + i) + (int64_t)(k0 + j) * ld] = Sm[i][j];
        }
        __syncthreads();
        lds_trtri_lower<T, false>(Sm, vec, nb, tid);
        for (int e = tid; e < 64 * 64; e += 256) {
            int i = e & 63, j = e >> 6;
            if (i < nb && j < nb && i >= j) W[(k0 + i) + (int64_t)(k0 + j) * K] = Sm[i][j];
        }
        // rows below inside the diagonal block: X = A * Lkk^{-T}
        for (int rt = k0 + nb; rt < K; rt += 64) {
            const int mr = min(64, K - rt);
            Quad<T> q;
            q.zero();
            if (wr < mr && wc < nb) {
                View<T> A{P + (rt + wr) + (int64_t)k0 * ld, 1, ld};
                View<T> B{&Sm[wc][0], 65, 1};
                quad_gemm_nt(q, A, mr - wr, B, nb - wc, nb, lane);
            }
            __syncthreads();
            if (wr < mr && wc < nb) {
                QUAD_FOR(if (qrow < mr - wr && qcol < nb - wc)
                             P[(rt + wr + qrow) + (int64_t)(k0 + wc + qcol) * ld] = q.c[m_][n_][r_];)
            }
        }
        __syncthreads();
        // trailing update of the diagonal block (lower tiles)
        for (int rt = k0 + nb; rt < K; rt += 64) {
            for (int ct = k0 + nb; ct <= rt; ct += 64) {
                const int mr = min(64, K - rt), mc = min(64, K - ct);
                if (wr < mr && wc < mc) {
                    Quad<T> q;
                    q.zero();
                    View<T> A{P + (rt + wr) + (int64_t)k0 * ld, 1, ld};
                    View<T> B{P + (ct + wc) + (int64_t)k0 * ld, 1, ld};
                    quad_gemm_nt(q, A, mr - wr, B, mc - wc, nb, lane);
                    QUAD_FOR(if (qrow < mr - wr && qcol < mc - wc)
                                 P[(rt + wr + qrow) + (int64_t)(ct + wc + qcol) * ld] -= q.c[m_][n_][r_];)
                }
            }
        }
        __syncthreads();
    }
    wg_lower_inverse_offdiag<T>(View<T>{P, 1, ld}, W, K, Sm, tid);
    if (tid == 0 && npiv_s) atomicAdd(&stat[0], npiv_s);
}

// ---------------------------------------------------------------------------
// Diagonal LU (no pivoting, static pivoting) + inverses.  L holds the full
// block on exit (unit L strictly below, U on and above the diagonal); U's
// diagonal block receives (LU)^T as in the reference.  W1 = inv(unit L),
// W2 = inv(U)^T, both lower, ld = K.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_getrf_diag(DevSymbol S, const int32_t *__restrict__ cblks, T *__restrict__ lcoef,
                                                    T *__restrict__ ucoef, T *__restrict__ w1, T *__restrict__ w2,
                                                    double thr, int *stat) {
    __shared__ T Sm[64][65];
    __shared__ T vec[64];
    __shared__ int npiv_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    T *P = lcoef + S.cblk_coef[c];
    T *U = ucoef + S.cblk_coef[c];
    T *W1 = w1 + S.cblk_winv[c];
    T *W2 = w2 + S.cblk_winv[c];
    const T crit = (T)thr;
    if (tid == 0) npiv_s = 0;
    // L_diag += U_diag^T (core_zgeadd Trans)
    for (int e = tid; e < K * K; e += 256) {
        int i = e % K, j = e / K;
        P[i + (int64_t)j * ld] += U[j + (int64_t)i * ld];
    }
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += 64) {
        const int nb = min(64, K - k0);
        for (int e = tid; e < 64 * 64; e += 256) {
            int i = e & 63, j = e >> 6;
            Sm[i][j] = (i < nb && j < nb) ? P[(k0 + i) + (int64_t)(k0 + j) * ld] : (T)0;
        }
        __syncthreads();
        // core_zgetf2sp on the square block
        for (int k = 0; k < nb; k++) {
            if (tid == 0) {
                T d = Sm[k][k];
                if (tabs(d) < crit) {
                    d = (d < (T)0) ? -crit : crit;
                    npiv_s++;
                }
                Sm[k][k] = d;
            }
            __syncthreads();
            const T alpha = (T)1 / Sm[k][k];
            if (tid > k && tid < nb) Sm[tid][k] *= alpha;
            __syncthreads();
            const int m = nb - k - 1;
            for (int e = tid; e < m * m; e += 256) {
                int ii = e % m, jj = e / m;
                Sm[k + 1 + ii][k + 1 + jj] -= Sm[k + 1 + ii][k] * Sm[k][k + 1 + jj];
            }
            __syncthreads();
        }
        for (int e = tid; e < 64 * 64; e += 256) {
            int i = e & 63, j = e >> 6;
            if (i < nb && j < nb) P[(k0 + i) + (int64_t)(k0 + j) * ld] = Sm[i][j];
        }
        __syncthreads();
        lds_trtri_lower<T, true>(Sm, vec, nb, tid);
        lds_trtri_upper<T>(Sm, vec, nb, tid);
        for (int e = tid; e < 64 * 64; e += 256) {
            int i = e & 63, j = e >> 6;
            if (i < nb && j < nb && i >= j) {
                W1[(k0 + i) + (int64_t)(k0 + j) * K] = (i == j) ? (T)1 : Sm[i][j];
                W2[(k0 + i) + (int64_t)(k0 + j) * K] = Sm[j][i];  // inv(U)^T
            }
        }
        __syncthreads();
        // rows below: Lik = Aik * inv(Ukk) = Aik * W2kk^T
        for (int rt = k0 + nb; rt < K; rt += 64) {
            const int mr = min(64, K - rt);
            Quad<T> q;
            q.zero();
            if (wr < mr && wc < nb) {
                View<T> A{P + (rt + wr) + (int64_t)k0 * ld, 1, ld};
                View<T> B{W2 + (k0 + wc) + (int64_t)k0 * K, 1, K};
                quad_gemm_nt(q, A, mr - wr, B, nb - wc, nb, lane);
            }
            __syncthreads();
            if (wr < mr && wc < nb) {
                QUAD_FOR(if (qrow < mr - wr && qcol < nb - wc)
                             P[(rt + wr + qrow) + (int64_t)(k0 + wc + qcol) * ld] = q.c[m_][n_][r_];)
            }
        }
        // columns right: Ukj = inv(Lkk) * Akj ; as (col, row) tiles: Y(j, i) = sum_k Akj(k, j) W1(i, k)
        for (int ct = k0 + nb; ct < K; ct += 64) {
            const int mc = min(64, K - ct);
            Quad<T> q;
            q.zero();
            if (wr < nb && wc < mc) {
                View<T> A{W1 + (k0 + wr) + (int64_t)k0 * K, 1, K};            // (i, k) = W1[k0+wr+i, k0+k]
                View<T> B{P + k0 + (int64_t)(ct + wc) * ld, ld, 1};           // (j, k) = A[k0+k, ct+wc+j]
                quad_gemm_nt(q, A, nb - wr, B, mc - wc, nb, lane);
            }
            __syncthreads();
            if (wr < nb && wc < mc) {
                QUAD_FOR(if (qrow < nb - wr && qcol < mc - wc)
                             P[(k0 + wr + qrow) + (int64_t)(ct + wc + qcol) * ld] = q.c[m_][n_][r_];)
            }
        }
        __syncthreads();
        // trailing update of the square: A -= Lik * Ukj
        for (int rt = k0 + nb; rt < K; rt += 64) {
            for (int ct = k0 + nb; ct < K; ct += 64) {
                const int mr = min(64, K - rt), mc = min(64, K - ct);
                if (wr < mr && wc < mc) {
                    Quad<T> q;
                    q.zero();
                    View<T> A{P + (rt + wr) + (int64_t)k0 * ld, 1, ld};
                    View<T> B{P + k0 + (int64_t)(ct + wc) * ld, ld, 1};
                    quad_gemm_nt(q, A, mr - wr, B, mc - wc, nb, lane);
                    QUAD_FOR(if (qrow < mr - wr && qcol < mc - wc)
                                 P[(rt + wr + qrow) + (int64_t)(ct + wc + qcol) * ld] -= q.c[m_][n_][r_];)
                }
            }
        }
        __syncthreads();
    }
    // full inverses: W1 from the strictly-lower part of L, W2 = inv(U^T) from U^T
    wg_lower_inverse_offdiag<T>(View<T>{P, 1, ld}, W1, K, Sm, tid);
    wg_lower_inverse_offdiag<T>(View<T>{P, ld, 1}, W2, K, Sm, tid);
    // U diagonal block <- (LU)^T (core_zgetmo)
    for (int e = tid; e < K * K; e += 256) {
        int i = e % K, j = e / K;
        U[i + (int64_t)j * ld] = P[j + (int64_t)i * ld];
    }
    if (tid == 0 && npiv_s) atomicAdd(&stat[0], npiv_s);
}

// ---------------------------------------------------------------------------
// Panel solve on 64-row tiles: X = A * W^T with W lower K x K.  Column blocks
// are produced from the last to the first so the product can overwrite A.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_panel(DevSymbol S, const TrsmItem *__restrict__ items, int nitems,
                                                    T *__restrict__ coef, const T *__restrict__ winv) {
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const TrsmItem it = items[id];
    const int c = it.c, r0 = it.r0;
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int stride = S.cblk_stride[c];
    const int64_t ld = stride;
    T *P = coef + S.cblk_coef[c];
    const T *W = winv + S.cblk_winv[c];
    const int mr = min(64, stride - r0);
    const int nkb = (K + 63) / 64;
    for (int jb = nkb - 1; jb >= 0; jb--) {
        const int j0 = jb * 64, nj = min(64, K - j0);
        Quad<T> q;
        q.zero();
        if (wr < mr && wc < nj) {
            View<T> A{P + (r0 + wr), 1, ld};
            View<T> B{W + (j0 + wc), 1, K};
            quad_gemm_nt(q, A, mr - wr, B, nj - wc, j0 + nj, lane);
        }
        __syncthreads();
        if (wr < mr && wc < nj) {
            QUAD_FOR(if (qrow < mr - wr && qcol < nj - wc)
                         P[(r0 + wr + qrow) + (int64_t)(j0 + wc + qcol) * ld] = q.c[m_][n_][r_];)
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline int grid8(int n) { return ((n + 7) / 8) * 8; }

template <typename T>
void launch_potrf_diag(const DevSymbol &S, const int32_t *cblks, int n, T *coef, T *winv, double thr, int *stat,
                       hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_potrf_diag<T>, dim3(n), dim3(256), 0, st, S, cblks, coef, winv, thr, stat);
}
template <typename T>
void launch_getrf_diag(const DevSymbol &S, const int32_t *cblks, int n, T *lcoef, T *ucoef, T *w1, T *w2, double thr,
                       int *stat, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_getrf_diag<T>, dim3(n), dim3(256), 0, st, S, cblks, lcoef, ucoef, w1, w2, thr, stat);
}
template <typename T>
void launch_trsm_panel(const DevSymbol &S, const TrsmItem *items, int n, T *coef, const T *winv, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_trsm_panel<T>, dim3(grid8(n)), dim3(256), 0, st, S, items, n, coef, winv);
}
#define INST(T)                                                                                              \
    template void launch_potrf_diag<T>(const DevSymbol &, const int32_t *, int, T *, T *, double, int *,      \
                                       hipStream_t);                                                         \
    template void launch_getrf_diag<T>(const DevSymbol &, const int32_t *, int, T *, T *, T *, T *, double,   \
                                       int *, hipStream_t);                                                  \
    template void launch_trsm_panel<T>(const DevSymbol &, const TrsmItem *, int, T *, const T *, hipStream_t);
INST(double)
INST(float)

}  // namespace pxa
