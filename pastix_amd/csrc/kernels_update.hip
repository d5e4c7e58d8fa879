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
//   k_update_seq    : tiles fed by several sources (FP64): each entry's
//                     product as in k_update_direct, subtracted into C in
//                     entry order;
//   k_update_acc    : tiles fed by several sources (FP32): contributions are
//                     summed in an LDS tile, written once at the end.
// MODE 0 (LL^t and the L side of LU) keeps pairs with block(p) >= block(q),
// MODE 1 (U side of LU) keeps block(p) > block(q), as the reference's
// iterblok >= blok / iterblok > blok loops.
#include "dev_common.h"
#include "dev_types.h"
#include "dev_update.h"
#include "kernels.h"

namespace pxa {

// K columns staged per LDS chunk in the single-tile GEMMs (16 measured
// faster than 32 on gfx950: the smaller LDS footprint keeps 6 workgroups per
// CU resident).
constexpr int UPD_CH = 16;

// register budget held to seven waves per SIMD (72 registers in FP64, no
// spills; the compiler's own choice, 80, allows six): 54.4 -> 53.9 ms FP64
// numfact on the 80^3 Laplacian
template <typename T, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_update_direct(DevSymbol S, const UpdItem *__restrict__ items, int nitems,
                                                       const UpdEntry *__restrict__ ents,
                                                       const int32_t *__restrict__ shift, const T *__restrict__ Acoef,
                                                       const T *__restrict__ Bcoef, T *__restrict__ Ccoef) {
    __shared__ StageBuf<T, UPD_CH> sb;
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
    wg_gemm_lds<T, UPD_CH>(q, Bcoef + S.cblk_coef[s] + en.q0, nq, ld, Acoef + S.cblk_coef[s] + en.p0, np, ld, K, sb, tid);
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

// Tiles fed by several sources: the entries in order, each one's product
// computed in registers as in k_update_direct and subtracted straight into
// C (C -= p1, then C -= p2, ..., the reference's source order), with a
// barrier between entries so that the next entry's read-modify-write of an
// element another wave wrote sees it.  The direct kernel's LDS footprint
// (no 64x64 accumulator), so six waves per SIMD instead of three.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update_seq(DevSymbol S, const UpdItem *__restrict__ items, int nitems,
                                                    const UpdEntry *__restrict__ ents, const int32_t *__restrict__ shift,
                                                    const T *__restrict__ Acoef, const T *__restrict__ Bcoef,
                                                    T *__restrict__ Ccoef) {
    __shared__ StageBuf<T, UPD_CH> sb;
    __shared__ int rmap[64], cmap[64], rblk[64], qblk[64];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const UpdItem it = items[id];
    const int f = it.f, r0 = it.r0, c0 = it.c0;
    const int fcol = S.cblk_fcol[f];
    const int64_t stf = S.cblk_stride[f];
    T *C = Ccoef + S.cblk_coef[f] + r0 + (int64_t)c0 * stf;
    for (int e = it.e0; e < it.e1; e++) {
        // the previous entry's stores and map reads are done by every wave
        if (e > it.e0) __syncthreads();
        const UpdEntry en = ents[e];
        update_maps(S, en, shift, r0, fcol + c0, rmap, cmap, rblk, qblk, tid);
        const int s = en.s;
        const int np = en.p1 - en.p0, nq = en.q1 - en.q0;
        const int K = S.cblk_fcol[s + 1] - S.cblk_fcol[s];
        const int64_t ld = S.cblk_stride[s];
        Quad<T> q;
        q.zero();
        // D(j, i) = B(q_j) . A(p_i): quadrant rows index columns of the target
        wg_gemm_lds<T, UPD_CH>(q, Bcoef + S.cblk_coef[s] + en.q0, nq, ld, Acoef + S.cblk_coef[s] + en.p0, np, ld, K,
                               sb, tid);
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
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update_acc(DevSymbol S, const UpdItem *__restrict__ items, int nitems,
                                                    const UpdEntry *__restrict__ ents, const int32_t *__restrict__ shift,
                                                    const T *__restrict__ Acoef, const T *__restrict__ Bcoef,
                                                    T *__restrict__ Ccoef) {
    __shared__ UpdAccLds<T> lds;
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    update_acc_tile<T, MODE>(S, items[id], ents, shift, Acoef, Bcoef, Ccoef, lds);
}

// Phase 1 of the multi-source tiles: dense 128 (p) x 64 (q) tile of a
// product group in the source's own row space (register budget held to four
// waves per SIMD: one more than the compiler picks unbounded, no spills), P(p, q) = A(p, :) . B(q, :)^T, stored to the
// level workspace.  Computed transposed (quadrant rows index q) so that each
// lane's stores walk down a workspace column.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_update_wgemm(DevSymbol S, const WTile *__restrict__ items, int nitems,
                                                      const T *__restrict__ Acoef, const T *__restrict__ Bcoef,
                                                      T *__restrict__ W) {
    __shared__ StageBuf2<T, UPD_CH> sb;
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;
    const WTile w = items[id];
    const int s = w.s;
    const int K = S.cblk_fcol[s + 1] - S.cblk_fcol[s];
    const int64_t ld = S.cblk_stride[s];
    Quad<T> q0, q1;
    q0.zero();
    q1.zero();
    // 64 q rows (B) x 128 p rows (A) of the group's product
    wg_gemm_lds_64x128<T, UPD_CH>(q0, q1, Bcoef + S.cblk_coef[s] + w.q0, w.mq, ld, Acoef + S.cblk_coef[s] + w.p0,
                                  w.mp, ld, K, sb, tid);
    T *Wt = W + w.wofs;
    if (wr < w.mq && wc < w.mp) {
        UQFOR({
            const int qj = wr + qrow, pi = wc + qcol;
            if (qj < w.mq && pi < w.mp) Wt[pi + (int64_t)qj * w.ldw] = q0.c[m_][n_][r_];
        })
    }
    if (wr < w.mq && wc + 32 < w.mp) {
        UQFOR({
            const int qj = wr + qrow, pi = wc + 32 + qcol;
            if (qj < w.mq && pi < w.mp) Wt[pi + (int64_t)qj * w.ldw] = q1.c[m_][n_][r_];
        })
    }
}

// Phase 2: every multi-source target tile sums, in entry order, the products
// of its entries read from the workspace and subtracts the sum once.
// Thread (r, ph) owns target row r and columns ph, ph+4, ..., ph+60 of the
// tile.  Entries are taken in batches of 8: first the source rows landing on
// each target row / column of every entry of the batch are published in LDS
// tables tagged with the entry index (64 threads per entry, independent
// loads for the whole batch), then the products are accumulated in entry
// order.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update_gather(DevSymbol S, const UpdItem *__restrict__ items, int nitems,
                                                       const UpdEntry *__restrict__ ents,
                                                       const int32_t *__restrict__ shift,
                                                       const WGroup *__restrict__ grp, const T *__restrict__ W,
                                                       T *__restrict__ Ccoef) {
    constexpr int NB = 8;
    __shared__ int rp[NB][64], rs[NB][64], rbk[NB][64];
    __shared__ int cq[NB][64], cs[NB][64], cbk[NB][64];
    __shared__ int64_t wb[NB];
    __shared__ int wl[NB];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, r = tid & 63, ph = tid >> 6;
    const UpdItem it = items[id];
    const int f = it.f, r0 = it.r0, c0 = it.c0;
    const int fcol = S.cblk_fcol[f];
    for (int e = tid; e < NB * 64; e += 256) {
        (&rs[0][0])[e] = -1;
        (&cs[0][0])[e] = -1;
    }
    // products: thread (g, cph) owns the E consecutive target rows E*g.. of
    // columns cph + CP*u; when the rows are one run of source rows of the
    // entry, a column's contribution is one 16-byte load
    constexpr int E = Vec<T>::E, RG = 64 / E, CP = 256 / RG, NC = 64 / CP;
    const int rr = E * (tid % RG), cph = tid / RG;
    T acc[NC][E];
#pragma unroll
    for (int u = 0; u < NC; u++)
#pragma unroll
        for (int i = 0; i < E; i++) acc[u][i] = (T)0;
    for (int eb = it.e0; eb < it.e1; eb += NB) {
        const int ne = min(NB, it.e1 - eb);
        __syncthreads();  // previous batch consumed (and the tables initialised)
#pragma unroll
        for (int pass = 0; pass < NB / 4; pass++) {
            const int slot = pass * 4 + ph;
            if (slot < ne) {
                const int e = eb + slot;
                const UpdEntry en = ents[e];
                const int64_t rb = S.cblk_rowbase[en.s];
                const int p = en.p0 + r;
                if (p < en.p1) {
                    const int bk = S.panel_blk[rb + p];
                    const int tr = S.panel_grow[rb + p] + shift[en.sb + bk - en.ib0] - r0;
                    rp[slot][tr] = p;
                    rbk[slot][tr] = bk;
                    rs[slot][tr] = e;
                }
                const int qq = en.q0 + r;
                if (qq < en.q1) {
                    const int tc = S.panel_grow[rb + qq] - fcol - c0;
                    cq[slot][tc] = qq;
                    cbk[slot][tc] = S.panel_blk[rb + qq];
                    cs[slot][tc] = e;
                }
                if (r == 0) {
                    const WGroup G = grp[en.grp];
                    const int base = S.blok_roff[en.ib0];
                    wb[slot] = G.wofs - base - (int64_t)base * G.ldw;
                    wl[slot] = G.ldw;
                }
            }
        }
        __syncthreads();
        for (int slot = 0; slot < ne; slot++) {
            const int e = eb + slot;
            bool cov[E], any = false, run = true;
            int pr[E], bkr[E];
            const int p0 = rp[slot][rr];
#pragma unroll
            for (int i = 0; i < E; i++) {
                cov[i] = rs[slot][rr + i] == e;
                pr[i] = rp[slot][rr + i];
                bkr[i] = rbk[slot][rr + i];
                any |= cov[i];
                run &= cov[i] && pr[i] == p0 + i;
            }
            if (!any) continue;
            const int64_t ldw = wl[slot];
            const T *Wb = W + wb[slot];
#pragma unroll
            for (int u = 0; u < NC; u++) {
                const int c = cph + CP * u;
                if (cs[slot][c] != e) continue;
                const T *Wc = Wb + (int64_t)cq[slot][c] * ldw;
                const int cb = cbk[slot][c];
                T v[E];
                if (run) {
                    const typename Vec<T>::u w = *reinterpret_cast<const typename Vec<T>::u *>(Wc + p0);
#pragma unroll
                    for (int i = 0; i < E; i++) v[i] = w[i];
                } else {
#pragma unroll
                    for (int i = 0; i < E; i++) v[i] = cov[i] ? Wc[pr[i]] : (T)0;
                }
#pragma unroll
                for (int i = 0; i < E; i++) {
                    const bool keep = cov[i] && ((MODE == 0) ? (bkr[i] >= cb) : (bkr[i] > cb));
                    if (keep) acc[u][i] += v[i];
                }
            }
        }
    }
    const int Kf = S.cblk_fcol[f + 1] - fcol;
    const int64_t stf = S.cblk_stride[f];
    T *C = Ccoef + S.cblk_coef[f];
#pragma unroll
    for (int u = 0; u < NC; u++) {
        const int cc = c0 + cph + CP * u;
        if (cc >= Kf) continue;
        bool nz = false;
#pragma unroll
        for (int i = 0; i < E; i++) nz |= acc[u][i] != (T)0;
        if (!nz) continue;
        T *Cp = C + r0 + rr + (int64_t)cc * stf;
        if (r0 + rr + E <= stf) {
            typename Vec<T>::u w = *reinterpret_cast<const typename Vec<T>::u *>(Cp);
#pragma unroll
            for (int i = 0; i < E; i++) w[i] -= acc[u][i];
            *reinterpret_cast<typename Vec<T>::u *>(Cp) = w;
        } else {
#pragma unroll
            for (int i = 0; i < E; i++)
                if (r0 + rr + i < stf && acc[u][i] != (T)0) Cp[i] -= acc[u][i];
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
        // FP64: entries subtracted in turn (six waves per SIMD; the LDS
        // accumulator's 52 KB allow three): 56.0 -> 54.3 ms numfact on the
        // 80^3 Laplacian.  FP32: the accumulator (26 KB, six waves already)
        // saves the extra read-modify-writes of C: 33.1 against 33.5 ms.
        if constexpr (sizeof(T) == 8) {
            if (mode == 0) hipLaunchKernelGGL((k_update_seq<T, 0>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
            else hipLaunchKernelGGL((k_update_seq<T, 1>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
        } else {
            if (mode == 0) hipLaunchKernelGGL((k_update_acc<T, 0>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
            else hipLaunchKernelGGL((k_update_acc<T, 1>), g, b, 0, st, S, items, n, ents, shift, A, B, C);
        }
    }
}

template <typename T>
void launch_update_wgemm(const DevSymbol &S, const WTile *items, int n, const T *A, const T *B, T *W, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_update_wgemm<T>), dim3(grid8(n)), dim3(256), 0, st, S, items, n, A, B, W);
}

template <typename T>
void launch_update_gather(const DevSymbol &S, int mode, const UpdItem *items, int n, const UpdEntry *ents,
                          const int32_t *shift, const WGroup *grp, const T *W, T *C, hipStream_t st) {
    if (n <= 0) return;
    const dim3 g(grid8(n)), b(256);
    if (mode == 0) hipLaunchKernelGGL((k_update_gather<T, 0>), g, b, 0, st, S, items, n, ents, shift, grp, W, C);
    else hipLaunchKernelGGL((k_update_gather<T, 1>), g, b, 0, st, S, items, n, ents, shift, grp, W, C);
}

#define PXA_UPD_INST(T)                                                                                           \
    template void launch_update<T>(const DevSymbol &, int, bool, const UpdItem *, int, const UpdEntry *,          \
                                   const int32_t *, const T *, const T *, T *, hipStream_t);                      \
    template void launch_update_wgemm<T>(const DevSymbol &, const WTile *, int, const T *, const T *, T *,        \
                                         hipStream_t);                                                            \
    template void launch_update_gather<T>(const DevSymbol &, int, const UpdItem *, int, const UpdEntry *,         \
                                          const int32_t *, const WGroup *, const T *, T *, hipStream_t);
PXA_UPD_INST(double)
PXA_UPD_INST(float)

}  // namespace pxa
