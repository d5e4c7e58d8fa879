// Blocked right-looking panel factorisation for CDNA4 (gfx950).
//
// A cblk's panel (stride x K, column-major) is factored by 64-column steps,
// each step being (at most) three launches over all cblks of the level:
//   k_potrf64   : Cholesky of the 64x64 diagonal sub-block with static
//                 pivoting (reference kernels/core_zpotrfsp.c:66-108), the
//                 block in registers, plus its inverse;
//   k_panel_trsm: every panel row below the sub-block times inv(L_kk)^T
//                 (core_zpotrfsp.c:166-176 and core_ztrsmsp.c:147-181);
//   k_panel_upd : before the step, left-looking update of the step's
//                 column block by the blocks before it (core_zpotrfsp.c:178-185
//                 applied column block by column block).
// k_winv_offdiag completes the inverse of the whole diagonal block for the
// solve once the panel is factored.  Inverses are stored twice: W (lower,
// column-major) for the forward sweep and WT = W^T for the backward sweep.
#include "dev_common.h"
#include "dev_types.h"
#include "kernels.h"

namespace pxa {

#define QFOR(...)                                                        \
    for (int m_ = 0; m_ < 2; m_++)                                       \
        for (int n_ = 0; n_ < 2; n_++)                                   \
            for (int r_ = 0; r_ < 4; r_++) {                             \
                const int qrow = m_ * 16 + Mfma<T>::drow(lane, r_);       \
                const int qcol = n_ * 16 + (lane & 15);                   \
                __VA_ARGS__                                              \
            }

template <typename T> __device__ __forceinline__ T pabs(T v) { return v < (T)0 ? -v : v; }

// Square root and its reciprocal of a positive pivot, and reciprocals, from
// the hardware estimates refined by Newton steps (y <- y + y(1/2 - d y^2/2)
// and y <- y + y(1 - d y)), each step doubling the correct bits, until about
// one ulp: fewer dependent instructions than the correctly rounded sqrt and
// division on the diagonal block's critical path.
__device__ __forceinline__ void sqrt_rsqrt(double d, double &L, double &r) {
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double h = 0.5 * d * y;
        y = __builtin_fma(y, __builtin_fma(-h, y, 0.5), y);
    }
    r = y;
    L = d * y;
}
__device__ __forceinline__ void sqrt_rsqrt(float d, float &L, float &r) {
    float y = __builtin_amdgcn_rsqf(d);
    const float h = 0.5f * d * y;
    y = __builtin_fmaf(y, __builtin_fmaf(-h, y, 0.5f), y);
    r = y;
    L = d * y;
}
__device__ __forceinline__ double recip(double d) {
    double y = __builtin_amdgcn_rcp(d);
    y = __builtin_fma(y, __builtin_fma(-d, y, 1.0), y);
    return __builtin_fma(y, __builtin_fma(-d, y, 1.0), y);
}
__device__ __forceinline__ float recip(float d) {
    const float y = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(y, __builtin_fmaf(-d, y, 1.0f), y);
}

// Shared by the inverses: column j of inv(M) for the lower-triangular 64x64
// M held in Ls (M(r, k) = Ls[r][k] when TRANS is false, Ls[k][r] when true;
// diagonal 1 when UNIT), by the four lanes 4*jl+q of a wave, x[kk] =
// inv(M)(4*kk + q, j).
template <typename T, bool UNIT, bool TRANS>
__device__ __forceinline__ void quad_lower_inverse(T (*Ls)[65], const T *dinv, int j, int q, T (&x)[16]) {
#pragma unroll
    for (int kk = 0; kk < 16; kk++) x[kk] = (T)0;
#pragma unroll
    for (int r = 0; r < 64; r++) {
        T s = (T)0;
#pragma unroll
        for (int kk = 0; 4 * kk < r; kk++)
            if (4 * kk + q < r) s += (TRANS ? Ls[4 * kk + q][r] : Ls[r][4 * kk + q]) * x[kk];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        const T dr = UNIT ? (T)1 : dinv[r];
        const T xr = (r < j) ? (T)0 : ((r == j) ? dr : -s * dr);
        if (q == (r & 3)) x[r >> 2] = xr;
    }
}

// The same inverse by blocks of 4 rows: per block, the four lanes of a
// column form the partial dot products of the 4 rows with the known part of
// x, combine them with lane shuffles, and solve the 4x4 diagonal block
// redundantly; 16 rounds instead of 64.
template <typename T, bool UNIT, bool TRANS>
__device__ __forceinline__ void quad_lower_inverse4(T (*Ls)[65], const T *dinv, int j, int q, T (&x)[16]) {
#pragma unroll
    for (int kk = 0; kk < 16; kk++) x[kk] = (T)0;
    auto M = [&](int r, int k) -> T { return TRANS ? Ls[k][r] : Ls[r][k]; };
#pragma unroll
    for (int rb = 0; rb < 16; rb++) {
        T s[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
        for (int kk = 0; kk < rb; kk++)
#pragma unroll
            for (int t = 0; t < 4; t++) s[t] += M(4 * rb + t, 4 * kk + q) * x[kk];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            s[t] += __shfl_xor(s[t], 1);
            s[t] += __shfl_xor(s[t], 2);
        }
        T y[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int r = 4 * rb + t;
            T acc = s[t];
#pragma unroll
            for (int u = 0; u < t; u++) acc += M(r, 4 * rb + u) * y[u];
            const T dr = UNIT ? (T)1 : dinv[r];
            y[t] = (r < j) ? (T)0 : ((r == j) ? dr : -acc * dr);
        }
        x[rb] = (q == 0) ? y[0] : (q == 1) ? y[1] : (q == 2) ? y[2] : y[3];
    }
}

// Cholesky and inverse of one 64-column diagonal sub-block.
//
// Thread (i, w) = (tid % 64, tid / 64) keeps the entries (i, j), j = w mod 4,
// of row i in registers.  Columns are taken 4 at a time: each wave publishes
// one of them through LDS (two buffers by parity), every thread factors the
// 4x4 diagonal block, replacing pivots below the threshold (reference
// kernels/core_zpotrfsp.c:78-81), and forms its row's 4 multipliers, which
// are published for the rank-4 update of the remaining columns.  Rows beyond
// the block width are padded with the identity, which the updates leave
// untouched.
//
// The inverse X = inv(L) is formed by columns: four lanes of one wave share
// column j and split the forward substitution L x = e_j by k mod 4; their
// partial sums are combined with lane shuffles (no barrier).  W receives X
// (lower, column-major, ld = K) and WT its transpose, both with explicit
// zeros in the other triangle of the sub-block.
template <typename T>
__global__ __launch_bounds__(256) void k_potrf64(DevSymbol S, const int32_t *__restrict__ cblks, int kb,
                                                 T *__restrict__ coef, T *__restrict__ winv, T *__restrict__ winvt,
                                                 double thr, int *stat) {
    __builtin_amdgcn_s_setprio(3);  // panel steps lead the critical path
    __shared__ T colb[2][4][64];
    __shared__ T lb[4][64];
    __shared__ T Ls[64][65];
    __shared__ T dinv[64];
    const int tid = threadIdx.x, i = tid & 63, w = tid >> 6;
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int k0 = kb * 64;
    const int nb = min(64, K - k0);
    const int64_t ld = S.cblk_stride[c];
    T *P = coef + S.cblk_coef[c] + k0 + (int64_t)k0 * ld;
    T a[16];
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
        const int j = 4 * jj + w;
        T v = (i == j) ? (T)1 : (T)0;
        if (i < nb && j < nb && j <= i) v = P[i + (int64_t)j * ld];
        a[jj] = v;
    }
    const T crit = (T)thr;
    int npiv = 0, neg = 0;
    // blocks of 4 columns k..k+3, one column per wave: publish, factor the
    // 4x4 diagonal block redundantly, form the row's multipliers, publish
    // them, rank-4 update of the row's remaining columns (two barriers per
    // block)
    auto piv = [&](T d, int col) -> T {
        if (col < nb) {
            if (pabs(d) < crit) {
                d = crit;
                npiv++;
            }
            if (d < (T)0) {
                neg = 1;
                d = -d;
            }
        }
        return d;
    };
#pragma unroll
    for (int b4 = 0; b4 < 16; b4++) {
        const int k = 4 * b4;
        T(*cb)[64] = colb[b4 & 1];
        cb[w][i] = a[b4];
        __syncthreads();
        // 4x4 Cholesky of the diagonal block (same in every thread)
        const T d0 = piv(cb[0][k], k);
        T L00, i00;
        sqrt_rsqrt(d0, L00, i00);
        const T L10 = cb[0][k + 1] * i00, L20 = cb[0][k + 2] * i00, L30 = cb[0][k + 3] * i00;
        const T d1 = piv(cb[1][k + 1] - L10 * L10, k + 1);
        T L11, i11;
        sqrt_rsqrt(d1, L11, i11);
        const T L21 = (cb[1][k + 2] - L20 * L10) * i11, L31 = (cb[1][k + 3] - L30 * L10) * i11;
        const T d2 = piv(cb[2][k + 2] - L20 * L20 - L21 * L21, k + 2);
        T L22, i22;
        sqrt_rsqrt(d2, L22, i22);
        const T L32 = (cb[2][k + 3] - L30 * L20 - L31 * L21) * i22;
        const T d3 = piv(cb[3][k + 3] - L30 * L30 - L31 * L31 - L32 * L32, k + 3);
        T L33, i33;
        sqrt_rsqrt(d3, L33, i33);
        // multipliers of row i
        T l0, l1, l2, l3;
        if (i > k + 3) {
            l0 = cb[0][i] * i00;
            l1 = (cb[1][i] - l0 * L10) * i11;
            l2 = (cb[2][i] - l0 * L20 - l1 * L21) * i22;
            l3 = (cb[3][i] - l0 * L30 - l1 * L31 - l2 * L32) * i33;
        } else {
            const int t = i - k;  // rows of the block (t < 0: rows above)
            l0 = (t == 0) ? L00 : (t == 1) ? L10 : (t == 2) ? L20 : L30;
            l1 = (t == 1) ? L11 : (t == 2) ? L21 : L31;
            l2 = (t == 2) ? L22 : L32;
            l3 = L33;
        }
        lb[0][i] = l0;
        lb[1][i] = l1;
        lb[2][i] = l2;
        lb[3][i] = l3;
        if (i >= k + w) a[b4] = (w == 0) ? l0 : (w == 1) ? l1 : (w == 2) ? l2 : l3;
        __syncthreads();
        if (i > k + 3) {
#pragma unroll
            for (int jj = b4 + 1; jj < 16; jj++) {
                const int j = 4 * jj + w;
                a[jj] -= l0 * lb[0][j] + l1 * lb[1][j] + l2 * lb[2][j] + l3 * lb[3][j];
            }
        }
    }
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
        const int j = 4 * jj + w;
        if (j <= i) Ls[i][j] = a[jj];
        if (i < nb && j < nb && j <= i) P[i + (int64_t)j * ld] = a[jj];
    }
    if (tid == 0) {
        if (npiv) atomicAdd(&stat[0], npiv);
        if (neg) atomicOr(&stat[1], 1);
    }
    __syncthreads();
    if (tid < 64) dinv[tid] = recip(Ls[tid][tid]);
    __syncthreads();
    // column j of inv(L): lanes 4*jl + q of wave w, x[kk] = X(4*kk + q, j)
    const int q = tid & 3, j = 16 * w + ((tid >> 2) & 15);
    T x[16];
    quad_lower_inverse4<T, false, false>(Ls, dinv, j, q, x);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) Ls[4 * kk + q][j] = x[kk];
    __syncthreads();
    T *W = winv + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    T *WT = winvt + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int lo = e & 63, hi = e >> 6;
        if (lo < nb && hi < nb) {
            W[lo + (int64_t)hi * K] = Ls[lo][hi];
            WT[lo + (int64_t)hi * K] = Ls[hi][lo];
        }
    }
}

// rows [r0, r0+64) below the diagonal sub-block kb: X = A * W_kk^T
template <typename T>
__global__ __launch_bounds__(256) void k_panel_trsm(DevSymbol S, const PanelItem *__restrict__ items, int nitems,
                                                    int kb, T *__restrict__ coef, const T *__restrict__ winv) {
    __builtin_amdgcn_s_setprio(3);
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const PanelItem it = items[id];
    const int c = it.c, r0 = it.r0;
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int stride = S.cblk_stride[c];
    const int64_t ld = stride;
    const int k0 = kb * 64, nb = min(64, K - k0);
    T *P = coef + S.cblk_coef[c];
    const T *W = winv + S.cblk_winv[c];
    const int mr = min(64, stride - r0);
    __shared__ StageBuf<T, 16> sb;
    Quad<T> q;
    q.zero();
    // the row tile is read completely (staged through LDS) before it is overwritten
    wg_gemm_lds<T, 16>(q, P + r0 + (int64_t)k0 * ld, mr, ld, W + k0 + (int64_t)k0 * K, nb, K, nb, sb, tid);
    if (wr < mr && wc < nb) {
        QFOR(if (qrow < mr - wr && qcol < nb - wc) P[(r0 + wr + qrow) + (int64_t)(k0 + wc + qcol) * ld] = q.c[m_][n_][r_];)
    }
}

// Left-looking update of the panel's column block kb (columns [k0, k0+64),
// k0 = 64*kb) by all the blocks before it, before that block is factored:
//   C[r, k0+j] -= sum_{k < k0} X[r, k] . Y[k0+j, k]   for the 64-row tile r0
// (C, X, Y panels of the same cblk: all L for LL^t; L, L, U and U, U, L for
// the two sides of LU; the reference's blocked potrf/getrf applies the same
// products step by step, kernels/core_zpotrfsp.c:178-185).  Every element of
// the block is read and written once, with K = k0 staged through LDS.  On the
// diagonal tile (r0 == k0) only r >= c (strict = 0) or r > c (strict = 1) is
// written: the L side of LU keeps the lower part and the diagonal, the U side
// (transposed) the strict upper part.
template <typename T>
__global__ __launch_bounds__(256) void k_panel_upd(DevSymbol S, const PanelItem *__restrict__ items, int nitems,
                                                   int kb, T *__restrict__ ccoef, const T *__restrict__ xcoef,
                                                   const T *__restrict__ ycoef, int strict) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ StageBuf<T, 16> sb;
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const PanelItem it = items[id];
    const int c = it.c, r0 = it.r0;
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int stride = S.cblk_stride[c];
    const int64_t ld = stride;
    const int k0 = kb * 64;
    T *P = ccoef + S.cblk_coef[c];
    const T *X = xcoef + S.cblk_coef[c], *Y = ycoef + S.cblk_coef[c];
    const int mr = min(64, stride - r0), mc = min(64, K - k0);
    Quad<T> q;
    q.zero();
    wg_gemm_lds<T, 16>(q, X + r0, mr, ld, Y + k0, mc, ld, k0, sb, tid);
    if (wr < mr && wc < mc) {
        const bool diag = r0 == k0;
        QFOR(if (qrow < mr - wr && qcol < mc - wc && (!diag || wr + qrow > wc + qcol - (strict ? 0 : 1)))
                 P[(r0 + wr + qrow) + (int64_t)(k0 + wc + qcol) * ld] -= q.c[m_][n_][r_];)
    }
}

// Off-diagonal 64-blocks of inv(L_cc) (and of its transpose) for the solve:
//   W[ib,jb] = -W[ib,ib] * sum_{kb=jb}^{ib-1} L[ib,kb] W[kb,jb]
template <typename T>
__global__ __launch_bounds__(256) void k_winv_offdiag(DevSymbol S, const int32_t *__restrict__ cblks, int lstride,
                                                      int lrs, int lcs_is_ld, const T *__restrict__ coef,
                                                      T *__restrict__ winv, T *__restrict__ winvt) {
    __shared__ T Sm[64][65];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    const T *P = coef + S.cblk_coef[c];
    // L(r, k) = P[r + k*ld] (lower factor) or P[k + r*ld] (transposed storage)
    const int64_t rs = lcs_is_ld ? ld : 1, cs = lcs_is_ld ? 1 : ld;
    (void)lstride;
    (void)lrs;
    T *W = winv + S.cblk_winv[c];
    T *WT = winvt ? winvt + S.cblk_winv[c] : nullptr;
    for (int jb0 = 0; jb0 < K; jb0 += 64) {
        const int nj = min(64, K - jb0);
        for (int ib0 = jb0 + 64; ib0 < K; ib0 += 64) {
            const int ni = min(64, K - ib0);
            Quad<T> q;
            q.zero();
            if (wr < ni && wc < nj) {
                View<T> A{P + (int64_t)(ib0 + wr) * rs + (int64_t)jb0 * cs, rs, cs};
                View<T> B{W + jb0 + (int64_t)(jb0 + wc) * K, K, 1};
                quad_gemm_nt(q, A, ni - wr, B, nj - wc, ib0 - jb0, lane);
            }
            QFOR(Sm[wr + qrow][wc + qcol] = q.c[m_][n_][r_];)
            __syncthreads();
            Quad<T> q2;
            q2.zero();
            if (wr < ni && wc < nj) {
                View<T> A{W + (ib0 + wr) + (int64_t)ib0 * K, 1, K};
                View<T> B{&Sm[0][wc], 1, 65};
                quad_gemm_nt(q2, A, ni - wr, B, nj - wc, ni, lane);
                QFOR(if (qrow < ni - wr && qcol < nj - wc) {
                    const int gi = ib0 + wr + qrow, gj = jb0 + wc + qcol;
                    const T v = -q2.c[m_][n_][r_];
                    W[gi + (int64_t)gj * K] = v;
                    if (winvt) WT[gj + (int64_t)gi * K] = v;
                })
            }
            __syncthreads();
        }
    }
}

// LU without pivoting of one 64-column diagonal sub-block (reference
// kernels/core_zgetrfsp.c:63-106, getf2 with static pivoting +-thr).  During
// the factorisation the square part of a cblk keeps its lower triangle
// (diagonal included) in the L panel and its strict upper triangle,
// transposed, in the U panel.  Thread (i, w) keeps the entries (i, j),
// j = w mod 4, of row i of the full sub-block; the elimination runs by
// blocks of four columns with one barrier each.  Outputs: L (unit, strict lower) and the diagonal of U
// in the L panel, the strict upper part of U transposed in the U panel,
// W1 = inv(L) (lower, unit diagonal), W2 = inv(U)^T (lower) and WB = inv(U)
// (upper), with explicit zeros in the other triangles of the sub-block.
template <typename T>
__global__ __launch_bounds__(256) void k_getrf64(DevSymbol S, const int32_t *__restrict__ cblks, int kb,
                                                 T *__restrict__ lcoef, T *__restrict__ ucoef, T *__restrict__ w1,
                                                 T *__restrict__ w2, T *__restrict__ wb, double thr, int *stat) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ T colb[2][4][64], rowb[2][4][64];
    __shared__ T Ls[64][65];
    __shared__ T dinv[64];
    const int tid = threadIdx.x, i = tid & 63, w = tid >> 6;
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int k0 = kb * 64;
    const int nb = min(64, K - k0);
    const int64_t ld = S.cblk_stride[c];
    T *PL = lcoef + S.cblk_coef[c] + k0 + (int64_t)k0 * ld;
    T *PU = ucoef + S.cblk_coef[c] + k0 + (int64_t)k0 * ld;
    T a[16];
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
        const int j = 4 * jj + w;
        T v = (i == j) ? (T)1 : (T)0;
        if (i < nb && j < nb) v = (j <= i) ? PL[i + (int64_t)j * ld] : PU[j + (int64_t)i * ld];
        a[jj] = v;
    }
    const T crit = (T)thr;
    int npiv = 0;
    auto piv = [&](T d, int col) -> T {
        if (col < nb && pabs(d) < crit) {
            d = (d < (T)0) ? -crit : crit;
            npiv++;
        }
        return d;
    };
    // blocks of 4 columns k..k+3: publish the block's 4 columns and 4 rows,
    // eliminate the 4x4 diagonal block redundantly in every thread, then row
    // i forms its 4 multipliers and column j its 4 rows of U from the
    // published values, and the rank-4 update needs no second barrier
#pragma unroll
    for (int b4 = 0; b4 < 16; b4++) {
        const int k = 4 * b4;
        T(*cb)[64] = colb[b4 & 1];
        T(*rb)[64] = rowb[b4 & 1];
        cb[w][i] = a[b4];
        if (i >= k && i < k + 4) {
#pragma unroll
            for (int jj = b4 + 1; jj < 16; jj++) rb[i - k][4 * jj + w] = a[jj];
        }
        __syncthreads();
        // D(t, u) = cb[u][k + t]
        const T u00 = piv(cb[0][k], k), r0 = recip(u00);
        const T l10 = cb[0][k + 1] * r0, l20 = cb[0][k + 2] * r0, l30 = cb[0][k + 3] * r0;
        const T u01 = cb[1][k], u02 = cb[2][k], u03 = cb[3][k];
        const T u11 = piv(cb[1][k + 1] - l10 * u01, k + 1), r1 = recip(u11);
        const T u12 = cb[2][k + 1] - l10 * u02, u13 = cb[3][k + 1] - l10 * u03;
        const T l21 = (cb[1][k + 2] - l20 * u01) * r1, l31 = (cb[1][k + 3] - l30 * u01) * r1;
        const T u22 = piv(cb[2][k + 2] - l20 * u02 - l21 * u12, k + 2), r2 = recip(u22);
        const T u23 = cb[3][k + 2] - l20 * u03 - l21 * u13;
        const T l32 = (cb[2][k + 3] - l30 * u02 - l31 * u12) * r2;
        const T u33 = piv(cb[3][k + 3] - l30 * u03 - l31 * u13 - l32 * u23, k + 3), r3 = recip(u33);
        if (i > k + 3) {
            const T m0 = cb[0][i] * r0;
            const T m1 = (cb[1][i] - m0 * u01) * r1;
            const T m2 = (cb[2][i] - m0 * u02 - m1 * u12) * r2;
            const T m3 = (cb[3][i] - m0 * u03 - m1 * u13 - m2 * u23) * r3;
            a[b4] = (w == 0) ? m0 : (w == 1) ? m1 : (w == 2) ? m2 : m3;
#pragma unroll
            for (int jj = b4 + 1; jj < 16; jj++) {
                const int j = 4 * jj + w;
                const T v0 = rb[0][j];
                const T v1 = rb[1][j] - l10 * v0;
                const T v2 = rb[2][j] - l20 * v0 - l21 * v1;
                const T v3 = rb[3][j] - l30 * v0 - l31 * v1 - l32 * v2;
                a[jj] -= m0 * v0 + m1 * v1 + m2 * v2 + m3 * v3;
            }
        } else if (i >= k) {
            const int t = i - k;
            // row t of the block: L strictly left of the diagonal, U from it
            T e;
            if (w == 0) e = (t == 0) ? u00 : (t == 1) ? l10 : (t == 2) ? l20 : l30;
            else if (w == 1) e = (t == 0) ? u01 : (t == 1) ? u11 : (t == 2) ? l21 : l31;
            else if (w == 2) e = (t == 0) ? u02 : (t == 1) ? u12 : (t == 2) ? u22 : l32;
            else e = (t == 0) ? u03 : (t == 1) ? u13 : (t == 2) ? u23 : u33;
            a[b4] = e;
#pragma unroll
            for (int jj = b4 + 1; jj < 16; jj++) {
                const int j = 4 * jj + w;
                const T v0 = rb[0][j];
                const T v1 = rb[1][j] - l10 * v0;
                const T v2 = rb[2][j] - l20 * v0 - l21 * v1;
                const T v3 = rb[3][j] - l30 * v0 - l31 * v1 - l32 * v2;
                a[jj] = (t == 0) ? v0 : (t == 1) ? v1 : (t == 2) ? v2 : v3;
            }
        }
    }
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
        const int j = 4 * jj + w;
        Ls[i][j] = a[jj];
        if (i < nb && j < nb) {
            if (j <= i) PL[i + (int64_t)j * ld] = a[jj];
            else PU[j + (int64_t)i * ld] = a[jj];
        }
    }
    if (tid == 0 && npiv) atomicAdd(&stat[0], npiv);
    __syncthreads();
    if (tid < 64) dinv[tid] = recip(Ls[tid][tid]);
    __syncthreads();
    const int q = tid & 3, j = 16 * w + ((tid >> 2) & 15);
    T x1[16], x2[16];
    quad_lower_inverse<T, true, false>(Ls, dinv, j, q, x1);   // inv(L), L unit lower
    quad_lower_inverse<T, false, true>(Ls, dinv, j, q, x2);   // inv(U^T)
    T *W1 = w1 + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    T *W2 = w2 + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    T *WB = wb + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) Ls[4 * kk + q][j] = x1[kk];
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
        const int lo = e & 63, hi = e >> 6;
        if (lo < nb && hi < nb) W1[lo + (int64_t)hi * K] = Ls[lo][hi];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) Ls[4 * kk + q][j] = x2[kk];
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
        const int lo = e & 63, hi = e >> 6;
        if (lo < nb && hi < nb) {
            W2[lo + (int64_t)hi * K] = Ls[lo][hi];
            WB[lo + (int64_t)hi * K] = Ls[hi][lo];
        }
    }
}

// Before the LU panel of a cblk: the square part's upper triangle arrives
// split between the L panel (same-block Schur contributions, written above
// the diagonal as the reference does) and the U panel (transposed): the
// reference adds them (core_zgeadd, kernels/core_zgetrfsp.c:240-242); here
// the sum goes to the U panel and the L panel keeps the lower triangle only.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_square_merge(DevSymbol S, const int32_t *__restrict__ cblks,
                                                         T *__restrict__ lcoef, T *__restrict__ ucoef) {
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    T *PL = lcoef + S.cblk_coef[c];
    T *PU = ucoef + S.cblk_coef[c];
    for (int64_t e = threadIdx.x; e < (int64_t)K * K; e += 256) {
        const int r = (int)(e % K), cc = (int)(e / K);
        if (r < cc) {
            PU[cc + (int64_t)r * ld] += PL[r + (int64_t)cc * ld];
            PL[r + (int64_t)cc * ld] = (T)0;
        }
    }
}

// Final layout of the LU diagonal blocks (reference core_zgetrfsp.c:240-251):
// the L panel's square receives the whole LU (strict upper part from the U
// panel, where it was kept transposed) and the U panel's square its
// transpose (core_zgetmo).  One workgroup per cblk.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_square_layout(DevSymbol S, const int32_t *__restrict__ cblks,
                                                          T *__restrict__ lcoef, T *__restrict__ ucoef) {
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    T *PL = lcoef + S.cblk_coef[c];
    T *PU = ucoef + S.cblk_coef[c];
    for (int64_t e = threadIdx.x; e < (int64_t)K * K; e += 256) {
        const int r = (int)(e % K), cc = (int)(e / K);
        if (r < cc) PL[r + (int64_t)cc * ld] = PU[cc + (int64_t)r * ld];
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < (int64_t)K * K; e += 256) {
        const int r = (int)(e % K), cc = (int)(e / K);
        PU[r + (int64_t)cc * ld] = PL[cc + (int64_t)r * ld];
    }
}

// ------------------------------- launchers ---------------------------------
static inline int grid8(int n) { return ((n + 7) / 8) * 8; }

template <typename T>
void launch_potrf64(const DevSymbol &S, const int32_t *cblks, int n, int kb, T *coef, T *winv, T *winvt, double thr,
                    int *stat, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_potrf64<T>, dim3(n), dim3(256), 0, st, S, cblks, kb, coef, winv, winvt, thr, stat);
}
template <typename T>
void launch_getrf64(const DevSymbol &S, const int32_t *cblks, int n, int kb, T *lcoef, T *ucoef, T *w1, T *w2, T *wb,
                    double thr, int *stat, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_getrf64<T>, dim3(n), dim3(256), 0, st, S, cblks, kb, lcoef, ucoef, w1, w2, wb, thr, stat);
}
template <typename T>
void launch_lu_square_merge(const DevSymbol &S, const int32_t *cblks, int n, T *lcoef, T *ucoef, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lu_square_merge<T>, dim3(n), dim3(256), 0, st, S, cblks, lcoef, ucoef);
}
template <typename T>
void launch_lu_square_layout(const DevSymbol &S, const int32_t *cblks, int n, T *lcoef, T *ucoef, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lu_square_layout<T>, dim3(n), dim3(256), 0, st, S, cblks, lcoef, ucoef);
}
template <typename T>
void launch_panel_trsm(const DevSymbol &S, const PanelItem *items, int n, int kb, T *coef, const T *winv,
                       hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_panel_trsm<T>, dim3(grid8(n)), dim3(256), 0, st, S, items, n, kb, coef, winv);
}
template <typename T>
void launch_panel_upd(const DevSymbol &S, const PanelItem *items, int n, int kb, T *C, const T *X, const T *Y,
                      bool strict, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_panel_upd<T>, dim3(grid8(n)), dim3(256), 0, st, S, items, n, kb, C, X, Y, strict ? 1 : 0);
}
template <typename T>
void launch_winv_offdiag(const DevSymbol &S, const int32_t *cblks, int n, bool transposed_factor, const T *coef,
                         T *winv, T *winvt, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_winv_offdiag<T>, dim3(n), dim3(256), 0, st, S, cblks, 0, 0, transposed_factor ? 1 : 0, coef,
                       winv, winvt);
}

#define INST(T)                                                                                                    \
    template void launch_getrf64<T>(const DevSymbol &, const int32_t *, int, int, T *, T *, T *, T *, T *, double, \
                                    int *, hipStream_t);                                                           \
    template void launch_lu_square_layout<T>(const DevSymbol &, const int32_t *, int, T *, T *, hipStream_t);      \
    template void launch_lu_square_merge<T>(const DevSymbol &, const int32_t *, int, T *, T *, hipStream_t);       \
    template void launch_potrf64<T>(const DevSymbol &, const int32_t *, int, int, T *, T *, T *, double, int *,    \
                                    hipStream_t);                                                                  \
    template void launch_panel_trsm<T>(const DevSymbol &, const PanelItem *, int, int, T *, const T *, hipStream_t); \
    template void launch_panel_upd<T>(const DevSymbol &, const PanelItem *, int, int, T *, const T *, const T *,  \
                                      bool, hipStream_t);                                                        \
    template void launch_winv_offdiag<T>(const DevSymbol &, const int32_t *, int, bool, const T *, T *, T *,       \
                                         hipStream_t);
INST(double)
INST(float)

}  // namespace pxa
