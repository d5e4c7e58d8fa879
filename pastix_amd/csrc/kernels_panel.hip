// Blocked right-looking panel factorisation for CDNA4 (gfx950).
//
// A cblk's panel (stride x K, column-major) is factored by 64-column steps,
// each step being three launches over all cblks of the level:
//   k_potrf64   : Cholesky of the 64x64 diagonal sub-block with static
//                 pivoting (reference kernels/core_zpotrfsp.c:66-108), the
//                 block in registers, plus its inverse;
//   k_panel_trsm: every panel row below the sub-block times inv(L_kk)^T
//                 (core_zpotrfsp.c:166-176 and core_ztrsmsp.c:147-181);
//   k_panel_upd : the in-cblk products (core_zpotrfsp.c:178-185 applied
//                 column block by column block): before the step, of the
//                 blocks before it into the step's block (left-looking), or,
//                 on levels with few rows, after the step, of the step's block
//                 into every later block (right-looking).
// k_winv_merge completes the inverses of the solve nodes' diagonal blocks
// once the panels are factored.  Inverses are stored twice: W (lower,
// column-major) for the forward sweep and WT = W^T for the backward sweep.
#include <algorithm>

#include "dev_common.h"
#include "dev_types.h"
#include "dev_update.h"
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
// the hardware estimates refined until about one ulp: for the FP64 inverse
// square root one third-order step, y <- y + y e (1/2 + 3/8 e) with
// e = 1 - d y^2, which triples the correct bits of the estimate; Newton steps
// y <- y + y(1/2 - d y^2/2) and y <- y + y(1 - d y), each doubling them,
// elsewhere.  Fewer dependent instructions than the correctly rounded sqrt
// and division on the diagonal block's critical path.
__device__ __forceinline__ void sqrt_rsqrt(double d, double &L, double &r) {
    double y = __builtin_amdgcn_rsq(d);
    const double e = __builtin_fma(-(d * y), y, 1.0);
    y = __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
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

__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmax_(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float fmax_(float a, float b) { return __builtin_fmaxf(a, b); }

// Static pivoting of one Cholesky pivot (reference
// kernels/core_zpotrfsp.c:78-81: |d| < crit is replaced by crit, a negative
// pivot is an error): max(|d|, crit), one instruction on the dependency
// chain; the counts are kept off the chain.  Identity padding columns (pivot
// exactly 1) pass through too: they stay decoupled from the real columns, and
// the caller removes them from the count when crit > 1.
template <typename T> __device__ __forceinline__ T piv_max(T d, T crit, int &npiv, int &neg) {
    npiv += (pabs(d) < crit) ? 1 : 0;
    neg |= (d < (T)0 && d <= -crit) ? 1 : 0;
    return fmax_(pabs(d), crit);
}

// Value of v in lane `lane` (wave-uniform) of the calling wave.
__device__ __forceinline__ double rlane(double v, int lane) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float rlane(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// One joining level of the recursive-doubling inverse below, on the vector
// units: for every pair of inverted diagonal blocks of size B (X11 at
// [base, base+B), X22 at [base+B, base+2B)), X21 = -X22 (M21 X11).
template <typename T, int B, bool TRANS>
__device__ __forceinline__ void inv_join_valu(T (*Ls)[65], T (*Xs)[65], T *Ts, int tid) {
    constexpr int NO = 32 * B;  // outputs of the level: 64/(2B) pairs of B x B
    auto M = [&](int r, int k) -> T { return TRANS ? Ls[k][r] : Ls[r][k]; };
    const int o = tid;
    const int p = o / (B * B), rem = o % (B * B), r = rem / B, cc = rem % B, base = 2 * B * p;
    if (o < NO) {
        T s = (T)0;
#pragma unroll
        for (int m = 0; m < B; m++) s = fma_(M(base + B + r, base + m), Xs[base + m][base + cc], s);
        Ts[o] = s;
    }
    __syncthreads();
    if (o < NO) {
        T s = (T)0;
#pragma unroll
        for (int m = 0; m < B; m++) s = fma_(Xs[base + B + r][base + B + m], Ts[p * B * B + m * B + cc], s);
        Xs[base + B + r][base + cc] = -s;
    }
    __syncthreads();
}

// The same on the matrix cores for B = 16, 32: one 16x16 tile per wave.
template <typename T, int B, bool TRANS>
__device__ __forceinline__ void inv_join_mfma(T (*Ls)[65], T (*Xs)[65], T *Ts, int lane, int w) {
    constexpr int TPP = (B / 16) * (B / 16), NT = (64 / (2 * B)) * TPP;
    auto M = [&](int r, int k) -> T { return TRANS ? Ls[k][r] : Ls[r][k]; };
    const int i = lane & 15, kq = lane >> 4;
    const int p = w / TPP, ti = (w % TPP) / (B / 16), tj = w % (B / 16), base = 2 * B * p;
    if (w < NT) {
        typename Mfma<T>::v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < B / 4; s++) {
            const int k = 4 * s + kq;
            acc = Mfma<T>::mma(M(base + B + ti * 16 + i, base + k), Xs[base + k][base + tj * 16 + i], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) Ts[p * B * B + (ti * 16 + Mfma<T>::drow(lane, r)) * B + tj * 16 + i] = acc[r];
    }
    __syncthreads();
    if (w < NT) {
        typename Mfma<T>::v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < B / 4; s++) {
            const int k = 4 * s + kq;
            acc = Mfma<T>::mma(Xs[base + B + ti * 16 + i][base + B + k], Ts[p * B * B + k * B + tj * 16 + i], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) Xs[base + B + ti * 16 + Mfma<T>::drow(lane, r)][base + tj * 16 + i] = -acc[r];
    }
    __syncthreads();
}

// X = inv(M) for the lower-triangular 64x64 M held in Ls (M(r, k) = Ls[r][k],
// or Ls[k][r] when TRANS; unit diagonal when UNIT, else 1/M(r, r) = dinv[r])
// by recursive doubling: the sixteen 4x4 diagonal blocks are inverted
// directly, then pairs of inverted blocks of size 4, 8, 16 and 32 are joined
// (inv([M11 0; M21 M22]) = [X11 0; -X22 M21 X11 X22]): 4 levels of small
// products instead of a 64-step substitution.  Xs receives X with zeros above
// the diagonal; Ts (1024 entries) is scratch.  With INPLACE (not TRANS), Xs
// is Ls, whose upper triangle the caller has zeroed: every join overwrites
// the block M21 it has consumed, and only M21 blocks are read from M after
// the 4x4 blocks.  Only X[0:nb, 0:nb] is complete: levels that would join
// blocks of the identity padding only are skipped.  All 256 threads (nb
// uniform), barriers inside; Ls and dinv must be visible to all threads on
// entry.
template <typename T, bool UNIT, bool TRANS, bool INPLACE>
__device__ __forceinline__ void lower_inverse(T (*Ls)[65], T (*Xs)[65], T *Ts, const T *dinv, int nb, int tid,
                                              int w) {
    static_assert(!(INPLACE && TRANS), "in place only from the lower triangle");
    auto M = [&](int r, int k) -> T { return TRANS ? Ls[k][r] : Ls[r][k]; };
    if (!INPLACE) {
        for (int e = tid; e < 64 * 64; e += 256) Xs[e >> 6][e & 63] = (T)0;
        __syncthreads();
    }
    if (tid < 64) {
        // one wave; the block's strict lower entries are read before any write
        const int b0 = 4 * (tid >> 2), cc = tid & 3;
        const T m10 = M(b0 + 1, b0), m20 = M(b0 + 2, b0), m21 = M(b0 + 2, b0 + 1);
        const T m30 = M(b0 + 3, b0), m31 = M(b0 + 3, b0 + 1), m32 = M(b0 + 3, b0 + 2);
        const T d0 = UNIT ? (T)1 : dinv[b0], d1 = UNIT ? (T)1 : dinv[b0 + 1];
        const T d2 = UNIT ? (T)1 : dinv[b0 + 2], d3 = UNIT ? (T)1 : dinv[b0 + 3];
        const T x0 = (cc == 0) ? d0 : (T)0;
        const T x1 = (cc > 1) ? (T)0 : (cc == 1) ? d1 : -(m10 * x0) * d1;
        const T x2 = (cc > 2) ? (T)0 : (cc == 2) ? d2 : -fma_(m21, x1, m20 * x0) * d2;
        const T x3 = (cc == 3) ? d3 : -fma_(m32, x2, fma_(m31, x1, m30 * x0)) * d3;
        Xs[b0][b0 + cc] = x0;
        Xs[b0 + 1][b0 + cc] = x1;
        Xs[b0 + 2][b0 + cc] = x2;
        Xs[b0 + 3][b0 + cc] = x3;
    }
    __syncthreads();
    // a level whose joined blocks start at or beyond nb only builds padding
    if (nb > 4) inv_join_valu<T, 4, TRANS>(Ls, Xs, Ts, tid);
    if (nb > 8) inv_join_valu<T, 8, TRANS>(Ls, Xs, Ts, tid);
    if (nb > 16) inv_join_mfma<T, 16, TRANS>(Ls, Xs, Ts, tid & 63, w);
    if (nb > 32) inv_join_mfma<T, 32, TRANS>(Ls, Xs, Ts, tid & 63, w);
}

// Cholesky and inverse of one 64-column diagonal sub-block.
//
// Thread (i, w) = (tid % 64, tid / 64) keeps the entries (i, j), j = w mod 4,
// of row i in registers.  Columns are taken 4 at a time with one barrier per
// block: the block's 4 columns are published through LDS (two buffers by
// parity), every thread factors the 4x4 diagonal block, replacing pivots
// below the threshold (reference kernels/core_zpotrfsp.c:78-81), and forms
// its row's 4 multipliers.  The multipliers of the rows a wave's columns
// belong to are held by lanes of the same wave (every wave forms all rows'
// multipliers), so the rank-4 update reads them across lanes (readlane)
// rather than through LDS.  Only the next block's columns are updated before
// they are published; the update of the columns after them is deferred into
// the next block, where it overlaps that block's 4x4 dependency chain.  Rows
// beyond the block width are padded with the identity, which the updates
// leave untouched.
//
// The inverse X = inv(L) comes from lower_inverse (recursive doubling, in
// place of L in LDS).  W receives X (lower, column-major, ld = K) and WT its
// transpose, both with explicit zeros in the other triangle of the sub-block.
// LDS of one diagonal-block factorisation: the published columns (then the
// inverse's scratch), the factor and the pivots' reciprocals.
template <typename T> struct PotrfLds {
    T fac[1024];
    T Ls[64][65];
    T dinv[64];
};

template <typename T>
__device__ __forceinline__ void potrf64_block(const DevSymbol &S, const int c, int kb, T *__restrict__ coef,
                                              T *__restrict__ winv, T *__restrict__ winvt, double thr, int *stat,
                                              PotrfLds<T> &lds) {
    T *fac = lds.fac;
    T(*colb)[4][64] = reinterpret_cast<T(*)[4][64]>(fac);  // [parity][w][row]
    T(*Ls)[65] = lds.Ls;
    T *dinv = lds.dinv;
    T *Ts = fac;  // inverse scratch once the factor is done
    const int tid = threadIdx.x, i = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
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
    T pl0 = 0, pl1 = 0, pl2 = 0, pl3 = 0;  // this row's multipliers of the last block (0 above it)
    colb[0][w][i] = a[0];
#pragma unroll
    for (int b4 = 0; b4 < 16; b4++) {
        const int k = 4 * b4;
        if (k >= nb) break;  // uniform: the padding needs no elimination
        __syncthreads();
        // the row index as an opaque value: keeps the compiler from hoisting
        // the 16 blocks' row masks out of the loop (they would hold registers
        // for the whole factorisation)
        int ii = i;
        asm volatile("" : "+v"(ii));
        T(*cb)[64] = colb[b4 & 1];
        const T c00 = cb[0][k], c10 = cb[0][k + 1], c20 = cb[0][k + 2], c30 = cb[0][k + 3];
        const T c11 = cb[1][k + 1], c21 = cb[1][k + 2], c31 = cb[1][k + 3];
        const T c22 = cb[2][k + 2], c32 = cb[2][k + 3], c33 = cb[3][k + 3];
        const T r0 = cb[0][ii], r1 = cb[1][ii], r2 = cb[2][ii], r3 = cb[3][ii];
        // deferred rank-4 update by the previous block of the columns after this block's
        if (b4 >= 1) {
#pragma unroll
            for (int jj = b4 + 1; jj < 16; jj++) {
                const int lj = 4 * jj + w;
                a[jj] = fma_(-pl3, rlane(pl3, lj), fma_(-pl2, rlane(pl2, lj),
                        fma_(-pl1, rlane(pl1, lj), fma_(-pl0, rlane(pl0, lj), a[jj]))));
            }
        }
        // 4x4 Cholesky of the diagonal block (same in every thread)
        T L00, i00, L11, i11, L22, i22, L33, i33;
        sqrt_rsqrt(piv_max(c00, crit, npiv, neg), L00, i00);
        const T L10 = c10 * i00, L20 = c20 * i00, L30 = c30 * i00;
        sqrt_rsqrt(piv_max(fma_(-L10, L10, c11), crit, npiv, neg), L11, i11);
        const T L21 = fma_(-L20, L10, c21) * i11, L31 = fma_(-L30, L10, c31) * i11;
        sqrt_rsqrt(piv_max(fma_(-L21, L21, fma_(-L20, L20, c22)), crit, npiv, neg), L22, i22);
        const T L32 = fma_(-L31, L21, fma_(-L30, L20, c32)) * i22;
        sqrt_rsqrt(piv_max(fma_(-L32, L32, fma_(-L31, L31, fma_(-L30, L30, c33))), crit, npiv, neg), L33, i33);
        // take the counts now: left to the end, they would keep every pivot in registers
        asm volatile("" : "+v"(npiv), "+v"(neg));
        // multipliers of row i (rows of the block take the block's L)
        const T m0 = r0 * i00;
        const T m1 = (r1 - m0 * L10) * i11;
        const T m2 = (r2 - m0 * L20 - m1 * L21) * i22;
        const T m3 = (r3 - m0 * L30 - m1 * L31 - m2 * L32) * i33;
        const int t = ii - k;
        const bool below = t > 3;
        const T l0 = below ? m0 : (t == 0) ? L00 : (t == 1) ? L10 : (t == 2) ? L20 : L30;
        const T l1 = below ? m1 : (t == 1) ? L11 : (t == 2) ? L21 : L31;
        const T l2 = below ? m2 : (t == 2) ? L22 : L32;
        const T l3 = below ? m3 : L33;
        const T lw = (w == 0) ? l0 : (w == 1) ? l1 : (w == 2) ? l2 : l3;
        if (t >= w) a[b4] = lw;
        pl0 = below ? l0 : (T)0;
        pl1 = below ? l1 : (T)0;
        pl2 = below ? l2 : (T)0;
        pl3 = below ? l3 : (T)0;
        if (b4 < 15) {
            const int jn = b4 + 1, lj = 4 * jn + w;
            a[jn] = fma_(-pl3, rlane(pl3, lj), fma_(-pl2, rlane(pl2, lj),
                    fma_(-pl1, rlane(pl1, lj), fma_(-pl0, rlane(pl0, lj), a[jn]))));
            colb[jn & 1][w][ii] = a[jn];
        }
    }
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
        const int j = 4 * jj + w;
        Ls[i][j] = (j <= i) ? a[jj] : (T)0;
        if (i < nb && j < nb && j <= i) P[i + (int64_t)j * ld] = a[jj];
    }
    if (crit > (T)1) npiv -= ((nb + 3) & ~3) - nb;  // the identity padding's pivots
    if (tid == 0) {
        if (npiv) atomicAdd(&stat[0], npiv);
        if (neg) atomicOr(&stat[1], 1);
    }
    __syncthreads();
    if (tid < 64) dinv[tid] = recip(Ls[tid][tid]);
    __syncthreads();
    lower_inverse<T, false, false, true>(Ls, Ls, Ts, dinv, nb, tid, w);
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

template <typename T>
__global__ __launch_bounds__(256) void k_potrf64(DevSymbol S, const int32_t *__restrict__ cblks, int kb,
                                                 T *__restrict__ coef, T *__restrict__ winv, T *__restrict__ winvt,
                                                 double thr, int *stat) {
    __builtin_amdgcn_s_setprio(3);  // panel steps lead the critical path
    __shared__ PotrfLds<T> lds;
    potrf64_block<T>(S, cblks[blockIdx.x], kb, coef, winv, winvt, thr, stat, lds);
}

// rows [r0, r0+64) of cblk c below the diagonal sub-block kb: X = A * W_kk^T
// (all 256 threads; contains barriers)
template <typename T>
__device__ __forceinline__ void panel_trsm_tile(const DevSymbol &S, const int c, const int r0, int kb,
                                                T *__restrict__ coef, const T *__restrict__ winv,
                                                StageBuf<T, 16> &sb) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int stride = S.cblk_stride[c];
    const int64_t ld = stride;
    const int k0 = kb * 64, nb = min(64, K - k0);
    T *P = coef + S.cblk_coef[c];
    const T *W = winv + S.cblk_winv[c];
    const int mr = min(64, stride - r0);
    Quad<T> q;
    q.zero();
    // the row tile is read completely (staged through LDS) before it is overwritten
    wg_gemm_lds<T, 16>(q, P + r0 + (int64_t)k0 * ld, mr, ld, W + k0 + (int64_t)k0 * K, nb, K, nb, sb, tid);
    if (wr < mr && wc < nb) {
        QFOR(if (qrow < mr - wr && qcol < nb - wc) P[(r0 + wr + qrow) + (int64_t)(k0 + wc + qcol) * ld] = q.c[m_][n_][r_];)
    }
}

// blockIdx.y selects the side: (coef, winv), or (coef2, winv2) for the
// second panel of LU (both sides of a step in one launch: they write
// different panels and read only the step's diagonal inverses)
template <typename T>
__global__ __launch_bounds__(256) void k_panel_trsm(DevSymbol S, const PanelItem *__restrict__ items, int nitems,
                                                    int kb, T *__restrict__ coef, const T *__restrict__ winv,
                                                    T *__restrict__ coef2, const T *__restrict__ winv2) {
    __builtin_amdgcn_s_setprio(3);
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    __shared__ StageBuf<T, 16> sb;
    const PanelItem it = items[id];
    if (blockIdx.y == 0) panel_trsm_tile<T>(S, it.c, it.r0, kb, coef, winv, sb);
    else panel_trsm_tile<T>(S, it.c, it.r0, kb, coef2, winv2, sb);
}

// In-cblk update of column block c0 of the panel (item), 64-row tile r0:
//   C[r, c0+j] -= sum_{k in Kr} X[r, k] . Y[c0+j, k]
// left-looking (right = 0: c0 is the step's block kb, Kr = [0, c0), before
// the step) or right-looking (right = 1: c0 is a later block, Kr = the
// columns of block kb, after the step).  C, X, Y are panels of the same cblk:
// all L for LL^t; L, L, U and U, U, L for the two sides of LU; the
// reference's blocked potrf/getrf applies the same products,
// kernels/core_zpotrfsp.c:178-185.  On a block's diagonal tile (r0 == c0)
// only r >= c (strict = 0) or r > c (strict = 1) is written: the L side of LU
// keeps the lower part and the diagonal, the U side (transposed) the strict
// upper part.
template <typename T>
__device__ __forceinline__ void panel_upd_tile(const DevSymbol &S, const PanelItem it, int kb, int right,
                                               T *__restrict__ ccoef, const T *__restrict__ xcoef,
                                               const T *__restrict__ ycoef, int strict) {
    __shared__ StageBuf<T, 16> sb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int c = it.c, r0 = it.r0, c0 = it.c0;
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int stride = S.cblk_stride[c];
    const int64_t ld = stride;
    const int ks = right ? kb * 64 : 0, kn = right ? min(64, K - ks) : c0;
    T *P = ccoef + S.cblk_coef[c];
    const T *X = xcoef + S.cblk_coef[c] + (int64_t)ks * ld, *Y = ycoef + S.cblk_coef[c] + (int64_t)ks * ld;
    const int mr = min(64, stride - r0), mc = min(64, K - c0);
    Quad<T> q;
    q.zero();
    wg_gemm_lds<T, 16>(q, X + r0, mr, ld, Y + c0, mc, ld, kn, sb, tid);
    if (wr < mr && wc < mc) {
        const bool diag = r0 == c0;
        QFOR(if (qrow < mr - wr && qcol < mc - wc && (!diag || wr + qrow > wc + qcol - (strict ? 0 : 1)))
                 P[(r0 + wr + qrow) + (int64_t)(c0 + wc + qcol) * ld] -= q.c[m_][n_][r_];)
    }
}

// blockIdx.y = 1: the other side of LU, C = Y = ycoef's panel, X = xcoef's,
// strict (the two sides write different panels and read only the blocks
// the step does not write)
template <typename T>
__global__ __launch_bounds__(256) void k_panel_upd(DevSymbol S, const PanelItem *__restrict__ items, int nitems,
                                                   int kb, int right, T *__restrict__ ccoef,
                                                   const T *__restrict__ xcoef, T *__restrict__ ycoef,
                                                   int strict) {
    __builtin_amdgcn_s_setprio(3);
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    if (blockIdx.y == 0) panel_upd_tile<T>(S, items[id], kb, right, ccoef, xcoef, ycoef, strict);
    else panel_upd_tile<T>(S, items[id], kb, right, ycoef, ycoef, xcoef, 1);
}

// One launch for an LL^t panel step's in-cblk update and the diagonal
// factorisation that waits on it.  Right-looking (right = 1): step kb's
// update of the later blocks, and the factorisation of block kb+1;
// left-looking (right = 0): the update of block kb by the blocks before it,
// and the factorisation of block kb.  Workgroups [0, npot) update the
// diagonal tile of the block to factor for cblk cblks[i] and factor it at
// once (potrf64_block); the others apply the rest of the update (items; the
// diagonal tiles of that block are skipped, their workgroup above has them).
// The factorisation no longer waits for the whole update, only for its tile.
template <typename T>
__global__ __launch_bounds__(256) void k_panel_upd_potrf(DevSymbol S, const int32_t *__restrict__ cblks, int npot,
                                                         const PanelItem *__restrict__ items, int nitems, int kb,
                                                         int right, T *__restrict__ coef, T *__restrict__ winv,
                                                         T *__restrict__ winvt, double thr, int *stat) {
    __builtin_amdgcn_s_setprio(3);
    const int id = blockIdx.x;
    const int kf = right ? kb + 1 : kb;  // the block factored here
    const int k1 = 64 * kf;
    if (id < npot) {
        const int c = cblks[id];
        panel_upd_tile<T>(S, PanelItem{c, k1, k1, 0}, kb, right, coef, coef, coef, 0);
        __syncthreads();  // the tile's stores reach the whole workgroup
        __shared__ PotrfLds<T> lds;
        potrf64_block<T>(S, c, kf, coef, winv, winvt, thr, stat, lds);
        return;
    }
    if (id - npot >= nitems) return;
    const PanelItem it = items[id - npot];
    if (it.r0 == k1 && it.c0 == k1) return;
    panel_upd_tile<T>(S, it, kb, right, coef, coef, coef, 0);
}

// Inverses of the solve nodes' diagonal blocks, by recursive doubling.  The
// diagonal-block kernels leave the inverse of every 64-column sub-block of a
// cblk in place (W, and WT = W^T); a merge round joins adjacent column runs
// [a, m) and [m, b) of a node whose inverses X11 and X22 are known:
//   X21 = -X22 M21 X11
// (M21: rows [m, b), columns [a, m) of the node's lower-triangular diagonal
// block), on 64x64 tiles of X21 in two phases: T = M21 X11, stored
// transposed, then X21 = -X22 T, written to W (column-major, lower) and to
// WT.  M is read in the members' panels: node column j lies in member c
// (o = its first node column), node row i >= o is panel row i - o of c
// (a node's members are coupled by full blocks).  k_winv_leaves first copies
// the members' 64-column inverses into the W and WT of nodes of several
// cblks (a node of one cblk already holds them in place).

// D(64x64) += X(64xK) Y(64xK)^T as wg_gemm_lds, where xcol(k) points at
// X(0, k) (rows contiguous; called with increasing k, also for k >= K, whose
// rows are not read) and Y(r, k) = Y[k + r*ldy] (consecutive k contiguous)
template <typename T, int CH, class XCol>
__device__ __forceinline__ void wg_gemm_lds_xcol_yt(Quad<T> &q, XCol &xcol, int mx, const T *Y, int my, int64_t ldy,
                                                    int K, StageBuf<T, CH> &sb, int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int i = lane & 15, kq = lane >> 4;
    constexpr int E = Vec<T>::E, NP = CH / (4 * E), NY = CH / 4;
    static_assert(CH % (4 * E) == 0 && CH % 4 == 0, "chunk of whole passes");
    const int re = E * (tid % (64 / E)), ke = tid / (64 / E);
    // Y: thread (row yr, k group yk) fetches NY consecutive k of its row
    const int yr = tid & 63, yk = (tid >> 6) * NY;
    T gx[NP][E], gy[NY];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NP; u++) {
            const int k = k0 + ke + 4 * E * u;
            ld_rows(xcol(k) + re, re, k < K ? mx : 0, gx[u]);
        }
#pragma unroll
        for (int e = 0; e < NY; e++) {
            const int k = k0 + yk + e;
            gy[e] = (k < K && yr < my) ? Y[k + (int64_t)yr * ldy] : (T)0;
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int u = 0; u < NP; u++) st_rows(&sb.xs[ke + 4 * E * u][re], gx[u]);
#pragma unroll
        for (int e = 0; e < NY; e++) sb.ys[yk + e][yr] = gy[e];
    };
    const bool active = (wr < mx) && (wc < my);
    gload(0);
    sstore();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += CH) {
        const bool more = k0 + CH < K;
        if (more) gload(k0 + CH);
        if (active) {
#pragma unroll
            for (int s = 0; s < CH / 4; s++) {
                if (k0 + 4 * s < K) {
                    const int kr = 4 * s + kq;
                    const T a0 = sb.xs[kr][wr + i], a1 = sb.xs[kr][wr + 16 + i];
                    const T b0 = sb.ys[kr][wc + i], b1 = sb.ys[kr][wc + 16 + i];
                    q.c[0][0] = Mfma<T>::mma(a0, b0, q.c[0][0]);
                    q.c[0][1] = Mfma<T>::mma(a0, b1, q.c[0][1]);
                    q.c[1][0] = Mfma<T>::mma(a1, b0, q.c[1][0]);
                    q.c[1][1] = Mfma<T>::mma(a1, b1, q.c[1][1]);
                }
            }
        }
        __syncthreads();
        if (more) sstore();
        __syncthreads();
    }
}

// One workgroup per leaf (64-column block ib of member c of node nd, a node
// of several cblks): the leaf's inverse block into the node's W (and WT).
template <typename T>
__global__ __launch_bounds__(256) void k_winv_leaves(DevSymbol S, const WinvLeaf *__restrict__ leaves,
                                                     T *__restrict__ w, T *__restrict__ wt) {
    const WinvLeaf lf = leaves[blockIdx.x];
    const int tid = threadIdx.x;
    const int nd = lf.nd, c = lf.c;
    const int f0 = S.cblk_fcol[S.node_cblk[nd]];
    const int K = S.cblk_fcol[S.node_cblk[nd + 1]] - f0;
    const int o = S.cblk_fcol[c] - f0, wc = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int j0 = 64 * lf.ib, bw = min(64, wc - j0);
    const int64_t xo = S.node_winv[nd] + (o + j0) + (int64_t)(o + j0) * K;
    const int64_t mo = S.cblk_winv[c] + j0 + (int64_t)j0 * wc;
    // 16 entries per thread, loads first
    T v[16], vt[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int t = tid + 256 * u, i = t & 63, j = t >> 6;
        const bool ok = i < bw && j < bw;
        v[u] = ok ? w[mo + i + (int64_t)j * wc] : (T)0;
        vt[u] = (ok && wt) ? wt[mo + i + (int64_t)j * wc] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int t = tid + 256 * u, i = t & 63, j = t >> 6;
        if (i < bw && j < bw) {
            w[xo + i + (int64_t)j * K] = v[u];
            if (wt) wt[xo + i + (int64_t)j * K] = vt[u];
        }
    }
}

// 64-deep staging chunks: the late rounds have few tiles per launch, so each
// workgroup's chunk latency is not hidden by others
constexpr int WINV_CH = 16;

template <typename T, int PHASE>
__global__ __launch_bounds__(256) void k_winv_merge(DevSymbol S, const WinvTile *__restrict__ tiles,
                                                    const T *__restrict__ coef, T *__restrict__ w,
                                                    T *__restrict__ wt, T *__restrict__ tbuf) {
    __shared__ StageBuf<T, WINV_CH> sb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const WinvTile t = tiles[blockIdx.x];
    const int nd = t.nd, c0 = S.node_cblk[nd];
    const int f0 = S.cblk_fcol[c0];
    const int K = S.cblk_fcol[S.node_cblk[nd + 1]] - f0;
    const int h1 = t.m - t.a, h2 = t.b - t.m;
    const int mi = min(64, h2 - t.i0), mj = min(64, h1 - t.j0);
    T *X = w + S.node_winv[nd];
    Quad<T> q;
    q.zero();
    if constexpr (PHASE == 1) {
        // T[i0+i, j0+j] = sum_{k >= j0} M[m+i0+i, a+k] X[a+k, a+j0+j], M in
        // the members' panels (the member of a column followed as k grows)
        const int row0 = t.m + t.i0;
        int mc = c0, mend = S.cblk_fcol[c0 + 1] - f0, mo = 0;
        const T *mbase = coef + S.cblk_coef[c0] + row0;
        int64_t mld = S.cblk_stride[c0];
        auto xcol = [&](int k) -> const T * {
            const int gc = min(t.a + t.j0 + k, t.m - 1);
            while (gc >= mend) {
                mc++;
                mo = mend;
                mend = S.cblk_fcol[mc + 1] - f0;
                mbase = coef + S.cblk_coef[mc] + (row0 - mo);
                mld = S.cblk_stride[mc];
            }
            return mbase + (int64_t)(gc - mo) * mld;
        };
        wg_gemm_lds_xcol_yt<T, WINV_CH>(q, xcol, mi, X + (t.a + t.j0) + (int64_t)(t.a + t.j0) * K, mj, K, h1 - t.j0, sb,
                                   tid);
        T *Tt = tbuf + t.toff;
        QFOR(if (wr + qrow < mi && wc + qcol < mj)
                 Tt[(t.j0 + wc + qcol) + (int64_t)(t.i0 + wr + qrow) * h1] = q.c[m_][n_][r_];)
    } else {
        // X[m+i0+i, a+j0+j] = -sum_{k <= i0+i} X[m+i0+i, m+k] T[k, j0+j]
        wg_gemm_lds<T, WINV_CH>(q, X + (t.m + t.i0) + (int64_t)t.m * K, mi, K, tbuf + t.toff + t.j0, mj, h1,
                           min(t.i0 + 64, h2), sb, tid);
        T *XT = wt ? wt + S.node_winv[nd] : nullptr;
        QFOR(if (wr + qrow < mi && wc + qcol < mj) {
            const int gi = t.m + t.i0 + wr + qrow, gj = t.a + t.j0 + wc + qcol;
            const T v = -q.c[m_][n_][r_];
            X[gi + (int64_t)gj * K] = v;
            if (XT) XT[gj + (int64_t)gi * K] = v;
        })
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
// (upper), with explicit zeros in the other triangles of the sub-block; the
// inverses come from lower_inverse (recursive doubling, in place in LDS: U^T
// first, then L rebuilt from the registers).
template <typename T>
__global__ __launch_bounds__(256) void k_getrf64(DevSymbol S, const int32_t *__restrict__ cblks, int kb,
                                                 T *__restrict__ lcoef, T *__restrict__ ucoef, T *__restrict__ w1,
                                                 T *__restrict__ w2, T *__restrict__ wb, double thr, int *stat) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ T fac[1024];
    T(*colb)[4][64] = reinterpret_cast<T(*)[4][64]>(fac);
    T(*rowb)[4][64] = reinterpret_cast<T(*)[4][64]>(fac + 512);
    __shared__ T Ls[64][65];
    __shared__ T dinv[64];
    T *Ts = fac;  // inverse scratch once the factor is done
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
        if (k >= nb) break;  // uniform: the padding needs no elimination
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
    // U^T (lower, diagonal of U) in LDS for the first in-place inverse
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
        const int j = 4 * jj + w;
        Ls[j][i] = (j >= i) ? a[jj] : (T)0;
        if (i < nb && j < nb) {
            if (j <= i) PL[i + (int64_t)j * ld] = a[jj];
            else PU[j + (int64_t)i * ld] = a[jj];
        }
    }
    if (tid == 0 && npiv) atomicAdd(&stat[0], npiv);
    __syncthreads();
    if (tid < 64) dinv[tid] = recip(Ls[tid][tid]);
    __syncthreads();
    T *W1 = w1 + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    T *W2 = w2 + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    T *WB = wb + S.cblk_winv[c] + k0 + (int64_t)k0 * K;
    const int wu = __builtin_amdgcn_readfirstlane(w);
    lower_inverse<T, false, false, true>(Ls, Ls, Ts, dinv, nb, tid, wu);  // inv(U^T)
    for (int e = tid; e < 64 * 64; e += 256) {
        const int lo = e & 63, hi = e >> 6;
        if (lo < nb && hi < nb) {
            W2[lo + (int64_t)hi * K] = Ls[lo][hi];
            WB[lo + (int64_t)hi * K] = Ls[hi][lo];
        }
    }
    __syncthreads();
    // then the unit lower L (its diagonal is not read)
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
        const int j = 4 * jj + w;
        Ls[i][j] = (j < i) ? a[jj] : (T)0;
    }
    __syncthreads();
    lower_inverse<T, true, false, true>(Ls, Ls, Ts, dinv, nb, tid, wu);  // inv(L)
    for (int e = tid; e < 64 * 64; e += 256) {
        const int lo = e & 63, hi = e >> 6;
        if (lo < nb && hi < nb) W1[lo + (int64_t)hi * K] = Ls[lo][hi];
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

// ---------------------------------------------------------------------------
// Leaf region: one persistent launch whose tasks are driven by device-side
// dependency counters (reference kernels/cpucblk_zmpi_coeftab.c:482-562,
// cpucblk_zincoming_deps / cpucblk_zrelease_deps, with counters as
// blend/solver_backup.c:95-101 sets them, and the ready queue of the dynamic
// scheduler, sopalin/sequential_zpotrf.c:129-234).
//
// A task is one region cblk c (at most 64 columns): it pulls the updates of
// every region cblk facing it (its incoming tiles, entries in source order,
// owner computes: no other workgroup writes c), factors the diagonal block
// with its inverse, solves the rows below, then decrements the counter of
// every region cblk it faces; the decrement that reaches zero appends that
// cblk to the ready queue.  Workgroups take queue slots in order with one
// atomic; a slot's cblk is published with a release store once its counter
// reaches zero.  Every cblk enters the queue exactly once, and a taken slot
// whose cblk is not yet published waits only for tasks taken before it, all
// of them running on workgroups that wait for nothing: the queue drains
// whatever the number of resident workgroups.  Every wait is bounded: after
// `maxwait` ticks of the constant-rate wall clock without its slot being
// published a workgroup sets stat[2], every workgroup leaves, and the host
// reports the error.
//
// Memory ordering: a task's stores are released at agent scope by every wave
// (the writes of other compute units' L2 become visible to the consumer),
// the counter decrement a release, the last decrement's thread acquires the
// earlier ones before the release store of the queue slot, and the consumer
// polls the slot with relaxed loads, then takes one acquire fence; the other
// waves of its workgroup are ordered after that fence by the barrier.
// sync[0]: next queue slot to take, sync[1]: next slot to publish.
template <typename T> union LeafLds {
    UpdAccLds<T> u;
    PotrfLds<T> p;
};

__global__ __launch_bounds__(256) void k_leaf_init(LeafDev LP, int32_t *__restrict__ ctr,
                                                   int32_t *__restrict__ queue, int32_t *__restrict__ sync) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < LP.cblknbr) ctr[i] = LP.ctr0[i];
    if (i < LP.nregion) queue[i] = (i < LP.ninit) ? LP.init[i] : -1;
    if (i == 0) {
        sync[0] = 0;
        sync[1] = LP.ninit;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_leaf_tasks(DevSymbol S, LeafDev LP, const UpdEntry *__restrict__ ents,
                                                    const int32_t *__restrict__ shift, T *__restrict__ coef,
                                                    T *__restrict__ winv, T *__restrict__ winvt, double thr,
                                                    int *stat, int32_t *__restrict__ ctr,
                                                    int32_t *__restrict__ queue, int32_t *__restrict__ sync,
                                                    long long maxwait) {
    __shared__ LeafLds<T> lds;
    __shared__ int task;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            int c = -1;
            const int t = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t < LP.nregion) {
                // relaxed polls (an acquire load would invalidate the L2 on
                // every poll), one acquire fence once the slot is published;
                // the wait is bounded in time (constant-rate wall clock)
                const unsigned long long t0 = wall_clock64();
                for (;;) {
                    c = __hip_atomic_load(&queue[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (c >= 0) break;
                    if (__hip_atomic_load(&stat[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                        wall_clock64() - t0 > (unsigned long long)maxwait) {
                        __hip_atomic_store(&stat[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        c = -1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(8);
                }
                if (c >= 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            task = c;
        }
        __syncthreads();
        const int c = task;
        __syncthreads();  // every thread holds the task before thread 0 takes the next one
        if (c < 0 || c >= LP.cblknbr) break;
        // incoming updates (each tile ends with a barrier after its stores)
        for (int k = LP.in_ptr[c]; k < LP.in_ptr[c + 1]; k++)
            update_acc_tile<T, 0>(S, LP.in[k], ents, shift, coef, coef, coef, lds.u);
        // the diagonal block and its inverse, then the rows below it
        potrf64_block<T>(S, c, 0, coef, winv, winvt, thr, stat, lds.p);
        __syncthreads();
        const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
        const int stride = S.cblk_stride[c];
        for (int r0 = K; r0 < stride; r0 += 64) {
            panel_trsm_tile<T>(S, c, r0, 0, coef, winv, lds.u.sb);
            __syncthreads();
        }
        // release the panel and the inverse, then the counters it feeds
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            for (int k = LP.tgt_ptr[c]; k < LP.tgt_ptr[c + 1]; k++) {
                const int f = LP.tgt[k];
                // release: this task's writes (fenced above by every wave);
                // the last decrement also acquires the earlier ones before it
                // publishes f
                if (__hip_atomic_fetch_add(&ctr[f], -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == 1) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    const int pos = __hip_atomic_fetch_add(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (pos < LP.nregion) __hip_atomic_store(&queue[pos], f, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    else __hip_atomic_store(&stat[2], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
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
    hipLaunchKernelGGL(k_panel_trsm<T>, dim3(grid8(n)), dim3(256), 0, st, S, items, n, kb, coef, winv, coef, winv);
}
template <typename T>
void launch_panel_trsm2(const DevSymbol &S, const PanelItem *items, int n, int kb, T *coef, const T *winv, T *coef2,
                        const T *winv2, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_panel_trsm<T>, dim3(grid8(n), 2), dim3(256), 0, st, S, items, n, kb, coef, winv, coef2,
                       winv2);
}
template <typename T>
void launch_panel_upd(const DevSymbol &S, const PanelItem *items, int n, int kb, bool right, T *C, const T *X,
                      const T *Y, bool strict, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_panel_upd<T>, dim3(grid8(n)), dim3(256), 0, st, S, items, n, kb, right ? 1 : 0, C, X,
                       const_cast<T *>(Y), strict ? 1 : 0);
}
template <typename T>
void launch_panel_upd_lu(const DevSymbol &S, const PanelItem *items, int n, int kb, bool right, T *L, T *U,
                         hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_panel_upd<T>, dim3(grid8(n), 2), dim3(256), 0, st, S, items, n, kb, right ? 1 : 0, L, L, U,
                       0);
}
template <typename T>
void launch_panel_upd_potrf(const DevSymbol &S, const int32_t *cblks, int npot, const PanelItem *items, int n, int kb,
                            bool right, T *coef, T *winv, T *winvt, double thr, int *stat, hipStream_t st) {
    if (npot + n <= 0) return;
    hipLaunchKernelGGL(k_panel_upd_potrf<T>, dim3(npot + n), dim3(256), 0, st, S, cblks, npot, items, n, kb,
                       right ? 1 : 0, coef, winv, winvt, thr, stat);
}
template <typename T>
void launch_winv_nodes(const DevSymbol &S, const WinvLeaf *leaves, int nleaves, const WinvTile *tiles,
                       const int32_t *round_ptr, int nrounds, const T *coef, T *w, T *wt, T *tbuf, hipStream_t st) {
    if (nleaves > 0) hipLaunchKernelGGL(k_winv_leaves<T>, dim3(nleaves), dim3(256), 0, st, S, leaves, w, wt);
    for (int r = 0; r < nrounds; r++) {
        const int t0 = round_ptr[r], nt = round_ptr[r + 1] - t0;
        if (nt <= 0) continue;
        hipLaunchKernelGGL((k_winv_merge<T, 1>), dim3(nt), dim3(256), 0, st, S, tiles + t0, coef, w, wt, tbuf);
        hipLaunchKernelGGL((k_winv_merge<T, 2>), dim3(nt), dim3(256), 0, st, S, tiles + t0, coef, w, wt, tbuf);
    }
}

void launch_leaf_init(const LeafDev &LP, int32_t *ctr, int32_t *queue, int32_t *sync, hipStream_t st) {
    const int n = std::max(std::max(LP.cblknbr, LP.nregion), 1);
    hipLaunchKernelGGL(k_leaf_init, dim3((n + 255) / 256), dim3(256), 0, st, LP, ctr, queue, sync);
}
template <typename T>
void launch_leaf_tasks(const DevSymbol &S, const LeafDev &LP, const UpdEntry *ents, const int32_t *shift, T *coef,
                       T *winv, T *winvt, double thr, int *stat, int32_t *ctr, int32_t *queue, int32_t *sync,
                       int grid, long long maxwait, hipStream_t st) {
    if (LP.nregion <= 0) return;
    const int g = std::max(1, std::min(grid, LP.nregion));
    hipLaunchKernelGGL(k_leaf_tasks<T>, dim3(g), dim3(256), 0, st, S, LP, ents, shift, coef, winv, winvt, thr, stat,
                       ctr, queue, sync, maxwait);
}

#define INST(T)                                                                                                    \
    template void launch_leaf_tasks<T>(const DevSymbol &, const LeafDev &, const UpdEntry *, const int32_t *, T *, \
                                       T *, T *, double, int *, int32_t *, int32_t *, int32_t *, int, long long,   \
                                       hipStream_t);                                                               \
    template void launch_getrf64<T>(const DevSymbol &, const int32_t *, int, int, T *, T *, T *, T *, T *, double, \
                                    int *, hipStream_t);                                                           \
    template void launch_lu_square_layout<T>(const DevSymbol &, const int32_t *, int, T *, T *, hipStream_t);      \
    template void launch_lu_square_merge<T>(const DevSymbol &, const int32_t *, int, T *, T *, hipStream_t);       \
    template void launch_potrf64<T>(const DevSymbol &, const int32_t *, int, int, T *, T *, T *, double, int *,    \
                                    hipStream_t);                                                                  \
    template void launch_panel_trsm<T>(const DevSymbol &, const PanelItem *, int, int, T *, const T *, hipStream_t); \
    template void launch_panel_trsm2<T>(const DevSymbol &, const PanelItem *, int, int, T *, const T *, T *,         \
                                        const T *, hipStream_t);                                                 \
    template void launch_panel_upd<T>(const DevSymbol &, const PanelItem *, int, int, bool, T *, const T *,       \
                                      const T *, bool, hipStream_t);                                             \
    template void launch_panel_upd_lu<T>(const DevSymbol &, const PanelItem *, int, int, bool, T *, T *,          \
                                         hipStream_t);                                                           \
    template void launch_panel_upd_potrf<T>(const DevSymbol &, const int32_t *, int, const PanelItem *, int, int,  \
                                            bool, T *, T *, T *, double, int *, hipStream_t);                    \
    template void launch_winv_nodes<T>(const DevSymbol &, const WinvLeaf *, int, const WinvTile *,               \
                                       const int32_t *, int, const T *, T *, T *, T *, hipStream_t);
INST(double)
INST(float)

}  // namespace pxa
