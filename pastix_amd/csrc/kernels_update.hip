// Schur-complement update for CDNA4 (gfx950)
// (reference kernels/core_zgemmsp.c:396-465 and cpucblk_zgemmsp :1337-1409).
//
// One workgroup owns each 64x64 tile of a target cblk's panel and applies, in
// source order, every contribution of the level's sources that lands in it:
// C[f] -= A[s](p rows) * B[s](q rows)^T scattered through the row/column maps
// (no atomics, bitwise-reproducible).  Operands are staged through LDS.
//   k_update_direct : tiles fed by a single source: the product is kept in
//                     registers (computed transposed so that stores walk down
//                     target columns) and subtracted straight into C;
//   k_update_acc    : tiles fed by several sources: contributions are summed
//                     in an LDS tile, written once at the end.
// MODE 0 (LL^t and the L side of LU) keeps pairs with block(p) >= block(q),
// MODE 1 (U side of LU) keeps block(p) > block(q), as the reference's
// iterblok >= blok / iterblok > blok loops.
#include "dev_common.h"
#include "dev_types.h"
#include "kernels.h"

namespace pxa {

#define UQFOR(...)                                                       \
    for (int m_ = 0; m_ < 2; m_++)                                       \
        for (int n_ = 0; n_ < 2; n_++)                                   \
            for (int r_ = 0; r_ < 4; r_++) {                             \
                const int qrow = m_ * 16 + Mfma<T>::drow(lane, r_);       \
                const int qcol = n_ * 16 + (lane & 15);                   \
                __VA_ARGS__                                              \
            }

// Row/column maps of one entry into the target tile.
__device__ __forceinline__ void update_maps(const DevSymbol &S, const UpdEntry &en, const int32_t *shift, int r0,
                                            int fcol_c0, int *rmap, int *cmap, int *rblk, int *qblk, int tid) {
    const int64_t rb = S.cblk_rowbase[en.s];
    if (tid < 64) {
        const int p = en.p0 + tid;
        if (p < en.p1) {
            const int bk = S.panel_blk[rb + p];
            rmap[tid] = S.panel_grow[rb + p] + shift[en.sb + bk - en.ib0] - r0;
            rblk[tid] = bk;
        }
    } else if (tid < 128) {
        const int q = en.q0 + tid - 64;
        if (q < en.q1) {
            cmap[tid - 64] = S.panel_grow[rb + q] - fcol_c0;
            qblk[tid - 64] = S.panel_blk[rb + q];
        }
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update_direct(DevSymbol S, const UpdItem *__restrict__ items, int nitems,
                                                       const UpdEntry *__restrict__ ents,
                                                       const int32_t *__restrict__ shift, const T *__restrict__ Acoef,
                                                       const T *__restrict__ Bcoef, T *__restrict__ Ccoef) {
    __shared__ StageBuf<T> sb;
    __shared__ int rmap[64], cmap[64], rblk[64], qblk[64];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const UpdItem it = items[id];
    const int f = it.f, r0 = it.r0, c0 = it.c0;
    const int fcol = S.cblk_fcol[f];
    const UpdEntry en = ents[it.e0];
    update_maps(S, en, shift, r0, fcol + c0, rmap, cmap, rblk, qblk, tid);
    const int s = en.s;
    const int np = en.p1 - en.p0, nq = en.q1 - en.q0;
    const int K = S.cblk_fcol[s + 1] - S.cblk_fcol[s];
    const int64_t ld = S.cblk_stride[s];
    Quad<T> q;
    q.zero();
    // D(j, i) = B(q_j) . A(p_i): quadrant rows index columns of the target
    wg_gemm_lds<T>(q, Bcoef + S.cblk_coef[s] + en.q0, nq, ld, Acoef + S.cblk_coef[s] + en.p0, np, ld, K, sb, tid);
    const int64_t stf = S.cblk_stride[f];
    T *C = Ccoef + S.cblk_coef[f] + r0 + (int64_t)c0 * stf;
    if (wr < nq && wc < np) {
        UQFOR({
            const int qj = wr + qrow, pi = wc + qcol;
            if (qj < nq && pi < np) {
                const bool keep = (MODE == 0) ? (rblk[pi] >= qblk[qj]) : (rblk[pi] > qblk[qj]);
                if (keep) C[rmap[pi] + (int64_t)cmap[qj] * stf] -= q.c[m_][n_][r_];
            }
        })
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update_acc(DevSymbol S, const UpdItem *__restrict__ items, int nitems,
                                                    const UpdEntry *__restrict__ ents, const int32_t *__restrict__ shift,
                                                    const T *__restrict__ Acoef, const T *__restrict__ Bcoef,
                                                    T *__restrict__ Ccoef) {
    __shared__ T acc[64][65];
    __shared__ StageBuf<T> sb;
    __shared__ int rmap[64], cmap[64], rblk[64], qblk[64];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const UpdItem it = items[id];
    const int f = it.f, r0 = it.r0, c0 = it.c0;
    const int fcol = S.cblk_fcol[f];
    for (int e = tid; e < 64 * 65; e += 256) (&acc[0][0])[e] = (T)0;
    for (int e = it.e0; e < it.e1; e++) {
        const UpdEntry en = ents[e];
        update_maps(S, en, shift, r0, fcol + c0, rmap, cmap, rblk, qblk, tid);
        const int s = en.s;
        const int np = en.p1 - en.p0, nq = en.q1 - en.q0;
        const int K = S.cblk_fcol[s + 1] - S.cblk_fcol[s];
        const int64_t ld = S.cblk_stride[s];
        Quad<T> q;
        q.zero();
        // wg_gemm_lds starts with a barrier-protected LDS fill: the maps are
        // visible to every thread once it returns
        wg_gemm_lds<T>(q, Acoef + S.cblk_coef[s] + en.p0, np, ld, Bcoef + S.cblk_coef[s] + en.q0, nq, ld, K, sb,
                       tid);
        if (wr < np && wc < nq) {
            UQFOR({
                const int pi = wr + qrow, qj = wc + qcol;
                if (pi < np && qj < nq) {
                    const bool keep = (MODE == 0) ? (rblk[pi] >= qblk[qj]) : (rblk[pi] > qblk[qj]);
                    if (keep) acc[rmap[pi]][cmap[qj]] += q.c[m_][n_][r_];
                }
            })
        }
        __syncthreads();
    }
    const int Kf = S.cblk_fcol[f + 1] - fcol;
    const int stf = S.cblk_stride[f];
    T *C = Ccoef + S.cblk_coef[f];
    for (int e = tid; e < 64 * 64; e += 256) {
        const int i = e & 63, j = e >> 6;
        const int r = r0 + i, cc = c0 + j;
        if (r < stf && cc < Kf) {
            const T v = acc[i][j];
            if (v != (T)0) C[r + (int64_t)cc * stf] -= v;
        }
    }
}

static inline int grid8(int n) { return ((n + 7) / 8) * 8; }

template <typename T>
void launch_update(const DevSymbol &S, int mode, bool direct, const UpdItem *items, int n, const UpdEntry *ents,
                   const int32_t *shift, const T *A, const T *B, T *C, hipStream_t st) {
    if (n <= 0) return;
    const dim3 g(grid8(n)), b(256);
    if (direct) {
        if (mode == 0) hipLaunchKernelGGL((k_update_direct<T, 0>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
        else hipLaunchKernelGGL((k_update_direct<T, 1>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
    } else {
        if (mode == 0) hipLaunchKernelGGL((k_update_acc<T, 0>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
        else hipLaunchKernelGGL((k_update_acc<T, 1>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
    }
}

template void launch_update<double>(const DevSymbol &, int, bool, const UpdItem *, int, const UpdEntry *,
                                    const int32_t *, const double *, const double *, double *, hipStream_t);
template void launch_update<float>(const DevSymbol &, int, bool, const UpdItem *, int, const UpdEntry *,
                                   const int32_t *, const float *, const float *, float *, hipStream_t);

}  // namespace pxa
