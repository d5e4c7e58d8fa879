// Device pieces of the Schur-complement update shared by the level kernels
// (kernels_update.hip) and the leaf-region task kernel (kernels_panel.hip):
// the row/column maps of one update entry into a 64x64 target tile and the
// one-pass sum of a tile's entries (reference kernels/core_zgemmsp.c:396-465).
#pragma once
#include "dev_common.h"
#include "dev_types.h"

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

// LDS of one accumulated tile: the 64x64 sum, the operand staging and the
// maps of the current entry.
template <typename T> struct UpdAccLds {
    T acc[64][65];
    StageBuf<T, 16> sb;
    int rmap[64], cmap[64], rblk[64], qblk[64];
};

// Target tile `it` (rows [r0, r0+64) x columns [c0, c0+64) of cblk f) minus
// the sum of its entries' products, the entries taken in order and summed
// in LDS, the tile written once.  MODE 0 keeps pairs with block(p) >=
// block(q) (LL^t, L side of LU), MODE 1 block(p) > block(q) (U side).  Must
// be called by all 256 threads of the workgroup (contains barriers); ends
// with a barrier after its stores are issued.
template <typename T, int MODE>
__device__ __forceinline__ void update_acc_tile(const DevSymbol &S, const UpdItem it,
                                                const UpdEntry *__restrict__ ents, const int32_t *__restrict__ shift,
                                                const T *__restrict__ Acoef, const T *__restrict__ Bcoef,
                                                T *__restrict__ Ccoef, UpdAccLds<T> &lds) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int f = it.f, r0 = it.r0, c0 = it.c0;
    const int fcol = S.cblk_fcol[f];
    for (int e = tid; e < 64 * 65; e += 256) (&lds.acc[0][0])[e] = (T)0;
    for (int e = it.e0; e < it.e1; e++) {
        const UpdEntry en = ents[e];
        update_maps(S, en, shift, r0, fcol + c0, lds.rmap, lds.cmap, lds.rblk, lds.qblk, tid);
        const int s = en.s;
        const int np = en.p1 - en.p0, nq = en.q1 - en.q0;
        const int K = S.cblk_fcol[s + 1] - S.cblk_fcol[s];
        const int64_t ld = S.cblk_stride[s];
        Quad<T> q;
        q.zero();
        // wg_gemm_lds starts with a barrier-protected LDS fill: the maps are
        // visible to every thread once it returns
        wg_gemm_lds<T, 16>(q, Acoef + S.cblk_coef[s] + en.p0, np, ld, Bcoef + S.cblk_coef[s] + en.q0, nq, ld, K,
                           lds.sb, tid);
        if (wr < np && wc < nq) {
            UQFOR({
                const int pi = wr + qrow, qj = wc + qcol;
                if (pi < np && qj < nq) {
                    const bool keep = (MODE == 0) ? (lds.rblk[pi] >= lds.qblk[qj]) : (lds.rblk[pi] > lds.qblk[qj]);
                    if (keep) lds.acc[lds.rmap[pi]][lds.cmap[qj]] += q.c[m_][n_][r_];
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
            const T v = lds.acc[i][j];
            if (v != (T)0) C[r + (int64_t)cc * stf] -= v;
        }
    }
    __syncthreads();
}

}  // namespace pxa
