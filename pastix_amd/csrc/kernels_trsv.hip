// Supernodal forward / backward solve kernels for CDNA4 (gfx950)
// (reference kernels/solve_ztrsmsp.c:67-557, sopalin/sequential_ztrsm.c:34-103).
//
// The sweeps read b from x and leave the solution in x; the forward result
// z = inv(L) b lives in a separate buffer so that every workgroup of a cblk
// can read the cblk's input while others write their rows of the output.
// The sweeps go over solve nodes: a node is one cblk, or a run of cblks
// whose diagonal blocks and coupling blocks form one dense lower-triangular
// block, solved with that block's whole inverse.
// Forward sweep, level by level:
//   k_fwd_diag : workgroup (nd, ib) gathers, for rows [0, 64*ib+64) of node
//                nd, b minus the update-buffer entries of these rows (stored
//                contiguously per row, sources in increasing order), then
//                writes rows [64*ib, 64*ib+64) of z_nd = inv(L_nd) y;
//   k_fwd_prod : the products of the members' panel rows outside their node
//                on 64-row tiles, each product row stored at its entry of
//                the update buffer.
// Backward sweep, levels from the root down:
//   k_bwd_prod : partial[j] = sum over a strip of L[r, j] x[row(r)] (panel
//                rows outside the node);
//   k_bwd_diag : workgroup (nd, ib) forms y = z_nd - the strips' partials for
//                columns [64*ib, K) and writes rows [64*ib, 64*ib+64) of
//                x_nd = inv(L_nd)^T y.
// The diagonal inverses are stored whole (zeros in the other triangle), so
// the triangular products run over full 64-column blocks.  Every output
// element is produced in a fixed order by one workgroup: no atomics,
// bitwise-reproducible solves.  NR is the compile-time bound on the
// right-hand sides of one launch (1 or 8); NW the waves of a workgroup.
#include <algorithm>

#include "dev_common.h"
#include "dev_types.h"
#include "kernels.h"

namespace pxa {

static inline int grid8(int n) { return ((n + 7) / 8) * 8; }
// levels whose widest cblk has at most this many columns run the solve
// kernels on 4-wave workgroups,
constexpr int SOLVE_NARROW = 64;
// and one-wave workgroups at most this many (the leaf levels)
constexpr int SOLVE_TINY = 16;
// levels whose widest node has more columns gather once per 64-row block
// before the diagonal kernels (k_fwd_gather, k_bwd_reduce)
constexpr int SOLVE_SPLIT = 512;

// Loads and stores of the data one solve task hands to another (update
// buffer, forward result, solution rows, backward partials).  CO: device-
// coherent accesses for the single-launch solve, where the producer and the
// consumer may sit on different dies; plain accesses between level kernels.
template <bool CO, typename T> __device__ __forceinline__ T ld_co(const T *p) {
    if constexpr (CO) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool CO, typename T> __device__ __forceinline__ void st_co(T *p, T v) {
    if constexpr (CO) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// red holds NW x 64 x NR partial sums; the sum over waves is taken in wave
// order by the threads of wave 0.
template <bool CO, typename T, int NR, int NW>
__device__ __forceinline__ void wave_sum_store(T (&acc)[NR], T *red, int r, int q, bool ok, int nrhs, T *out,
                                               int64_t ldo, const T *sub, int64_t lds) {
#pragma unroll
    for (int u = 0; u < NR; u++) red[(q * 64 + r) * NR + u] = acc[u];
    __syncthreads();
    if (q == 0 && ok) {
#pragma unroll
        for (int u = 0; u < NR; u++) {
            if (u < nrhs) {
                T v = red[r * NR + u];
#pragma unroll
                for (int w = 1; w < NW; w++) v += red[(w * 64 + r) * NR + u];
                st_co<CO>(out + u * ldo, sub ? ld_co<CO>(sub + u * lds) - v : v);
            }
        }
    }
}

// The same for partial sums held as E consecutive rows per lane: lane l of
// wave q holds rows E*(l % (64/E)) + i of columns l / (64/E); the E lanes of a
// row group combine with lane shuffles, then the waves as above (row r =
// the calling lane).
template <bool CO, typename T, int NR, int NW, int E>
__device__ __forceinline__ void wave_sum_store_rows(T (&acc)[E][NR], T *red, int r, int q, bool ok, int nrhs,
                                                    T *out, int64_t ldo) {
    constexpr int G = 64 / E;
#pragma unroll
    for (int off = G; off < 64; off *= 2)
#pragma unroll
        for (int i = 0; i < E; i++)
#pragma unroll
            for (int u = 0; u < NR; u++) acc[i][u] += __shfl_xor(acc[i][u], off);
    if (r < G) {
#pragma unroll
        for (int i = 0; i < E; i++)
#pragma unroll
            for (int u = 0; u < NR; u++) red[(q * 64 + E * r + i) * NR + u] = acc[i][u];
    }
    __syncthreads();
    if (q == 0 && ok) {
#pragma unroll
        for (int u = 0; u < NR; u++) {
            if (u < nrhs) {
                T v = red[r * NR + u];
#pragma unroll
                for (int w = 1; w < NW; w++) v += red[(w * 64 + r) * NR + u];
                st_co<CO>(out + u * ldo, v);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Task bodies, each run by one whole workgroup of NW waves (the level
// kernels below and the dependency-driven solve share them).
// ---------------------------------------------------------------------------
template <bool CO, typename T, int NR, int NW, bool PRE = false>
__device__ __forceinline__ void fwd_diag_body(const DevSymbol &S, const DiagItem it,
                                              const YSlices ys, const T *__restrict__ ybuf,
                                              int64_t ldy, const T *__restrict__ b, int64_t ldb, T *__restrict__ z,
                                              int64_t ldz, int nrhs, const T *__restrict__ w1,
                                              const T *__restrict__ w1t, T *y, T *red) {
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int nd = it.nd, r0 = it.ib * 64;
    const int fc = S.cblk_fcol[S.node_cblk[nd]];
    const int Kc = S.cblk_fcol[S.node_cblk[nd + 1]] - fc;
    const int nrow = min(Kc, r0 + 64);
    // y = b minus the pending products of rows [0, nrow): gl consecutive
    // lanes per row, as many as fill the workgroup in one pass (at most 64,
    // at most the node's first row's entry count), lane j summing the
    // entries j, j + gl, ... of its row, UL loads issued together, in order;
    // the lanes' sums combined by xor shuffles in a fixed order.  With one
    // lane per row, consecutive lanes read consecutive rows' entries, which
    // the sliced update buffer keeps side by side.
    const int tasks = nrow * nrhs;
    const int ne = ys.cnt[fc];
    int gl = 1;
    while (gl < 64 && gl * 2 <= ne && tasks * gl * 2 <= NW * 64) gl *= 2;
    // rows of the inverse: E consecutive per lane (16-byte loads); the first
    // PF column groups are fetched before the gather, so that their latency
    // overlaps it
    constexpr int E = Vec<T>::E, G = 64 / E, PF = 4;
    const int rr = E * (r % G), kc = r / G;
    const bool mrow = r0 + rr < Kc;
    const T *M = w1 + S.node_winv[nd] + r0 + rr;
    // one right-hand side with W^T at hand (LL^t): row i of W is read as
    // column i of W^T (contiguous), NW lanes per row (row r0 + jt, lane lt)
    const bool tw = NR == 1 && w1t != nullptr;
    const int jt = tid / NW, lt = tid % NW, it_row = r0 + jt;
    const T *MT = tw ? w1t + S.node_winv[nd] + (int64_t)it_row * Kc : nullptr;
    T mpre[PF][E];
#pragma unroll
    for (int f = 0; f < PF; f++) {
        if (tw) {
            const int k = E * lt + f * NW * E;
            if (it_row < Kc && k <= it_row) ld_rows(MT + k, k, it_row + 1, mpre[f]);
        } else {
            const int k = q * E + kc + f * NW * E;
            if (mrow && k < nrow) ld_rows(M + (int64_t)k * Kc, rr, Kc - r0, mpre[f]);
        }
    }
    constexpr int UL = 16;
    if constexpr (PRE) {
        // the rows' pending products were subtracted in b already (k_fwd_gather)
        for (int t = tid; t < tasks; t += NW * 64) {
            const int i = t % nrow, u = t / nrow;
            y[i + u * Kc] = b[fc + i + u * ldb];
        }
    } else
    for (int base = 0; base < tasks * gl; base += NW * 64) {
        const int t = base + tid, task = t / gl, j = t % gl;
        const int i = task % nrow, u = task / nrow;
        T v = 0;
        if (task < tasks) {
            const int g = fc + i, ne_i = ys.cnt[g];
            const T *yb = ybuf + u * ldy + ys.base[g >> 6] + (g & 63);
            for (int e = j; e < ne_i; e += UL * gl) {
                T a[UL];
#pragma unroll
                for (int k = 0; k < UL; k++) a[k] = (e + k * gl < ne_i) ? ld_co<CO>(yb + (int64_t)(e + k * gl) * 64) : (T)0;
#pragma unroll
                for (int k = 0; k < UL; k++) v += a[k];
            }
        }
        for (int o = gl >> 1; o >= 1; o >>= 1) v += __shfl_xor(v, o);
        if (task < tasks && j == 0) y[i + u * Kc] = b[fc + i + u * ldb] - v;
    }
    __syncthreads();
    if (tw) {
        // W[i, k] = W^T[k, i] is nonzero for k <= i; y holds rows [0, nrow)
        T a2[E];
#pragma unroll
        for (int e = 0; e < E; e++) a2[e] = 0;
        if (it_row < Kc) {
#pragma unroll
            for (int f = 0; f < PF; f++) {
                const int k = E * lt + f * NW * E;
                if (k <= it_row) {
#pragma unroll
                    for (int e = 0; e < E; e++) a2[e] += mpre[f][e] * y[min(k + e, it_row)];
                }
            }
#pragma unroll 4
            for (int k = E * lt + PF * NW * E; k <= it_row; k += NW * E) {
                T m[E];
                ld_rows(MT + k, k, it_row + 1, m);
#pragma unroll
                for (int e = 0; e < E; e++) a2[e] += m[e] * y[min(k + e, it_row)];
            }
        }
        T v = 0;
#pragma unroll
        for (int e = 0; e < E; e++) v += a2[e];
#pragma unroll
        for (int o = NW / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, NW);
        if (lt == 0 && it_row < Kc) st_co<CO>(z + fc + it_row, v);
        return;  // uniform: tw is the same for the whole workgroup
    }
    const int i = r0 + r;
    T acc[E][NR];
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
        for (int u = 0; u < NR; u++) acc[e][u] = 0;
    if (mrow) {
#pragma unroll
        for (int f = 0; f < PF; f++) {
            const int k = q * E + kc + f * NW * E;
            if (k < nrow) {
#pragma unroll
                for (int u = 0; u < NR; u++) {
                    if (u < nrhs) {
                        const T yv = y[k + u * Kc];
#pragma unroll
                        for (int e = 0; e < E; e++) acc[e][u] += mpre[f][e] * yv;
                    }
                }
            }
        }
#pragma unroll 4
        for (int k = q * E + kc + PF * NW * E; k < nrow; k += NW * E) {
            T m[E];
            ld_rows(M + (int64_t)k * Kc, rr, Kc - r0, m);
#pragma unroll
            for (int u = 0; u < NR; u++) {
                if (u < nrhs) {
                    const T yv = y[k + u * Kc];
#pragma unroll
                    for (int e = 0; e < E; e++) acc[e][u] += m[e] * yv;
                }
            }
        }
    }
    wave_sum_store_rows<CO, T, NR, NW, E>(acc, red, r, q, i < Kc, nrhs, z + fc + i, ldz);
}

template <bool CO, typename T, int NR, int NW>
__device__ __forceinline__ void fwd_prod_body(const DevSymbol &S, const FwdItem it, const T *__restrict__ coef,
                                              const T *__restrict__ z, int64_t ldz, int nrhs,
                                              const int32_t *__restrict__ ydst, T *__restrict__ ybuf, int64_t ldy,
                                              T *xs, T *red) {
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int c = it.c;
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    const int64_t ld = S.cblk_stride[c];
    for (int t = tid; t < Kc * nrhs; t += NW * 64) xs[t] = ld_co<CO>(z + fc + (t % Kc) + (int64_t)(t / Kc) * ldz);
    __syncthreads();
    const int nr = it.r1 - it.r0;
    // E consecutive rows per lane (one 16-byte load), E columns per wave
    // instruction
    constexpr int E = Vec<T>::E, G = 64 / E;
    const int rr = E * (r % G), kc = r / G;
    T acc[E][NR];
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
        for (int u = 0; u < NR; u++) acc[i][u] = 0;
    if (rr < nr) {
        const T *rows = coef + S.cblk_coef[c] + it.r0 + rr;
#pragma unroll 2
        for (int k = q * E + kc; k < Kc; k += NW * E) {
            T a[E];
            ld_rows(rows + (int64_t)k * ld, rr, nr, a);
#pragma unroll
            for (int u = 0; u < NR; u++) {
                if (u < nrhs) {
                    const T xv = xs[k + u * Kc];
#pragma unroll
                    for (int i = 0; i < E; i++) acc[i][u] += a[i] * xv;
                }
            }
        }
    }
    const int dst = (r < nr) ? ydst[it.slot + r] : 0;
    wave_sum_store_rows<CO, T, NR, NW, E>(acc, red, r, q, r < nr, nrhs, ybuf + dst, ldy);
}

// one strip (panel rows [r0, r1), any multiple of 64 of them, column block
// k0/64) of the backward products, 4 waves: partial[j] = sum over the strip's
// rows of L[r, k0+j] x[row(r)], 64 rows at a time, the next chunk's loads in
// flight while the current one is summed.  nrhs right-hand sides starting at
// column u0 of the partials (laid out for nrhs_tot columns); x points at
// column u0 already
template <bool CO, typename T, int NR>
__device__ __forceinline__ void bwd_prod_body(const DevSymbol &S, const BwdItem it, int k0,
                                              const T *__restrict__ coef, const T *__restrict__ x, int64_t ldx,
                                              int nrhs, int nrhs_tot, int u0, T *__restrict__ part, T (*tile)[65],
                                              T *xr, T *red) {
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    const int c = it.c;
    const int Kc = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    const int64_t rb = S.cblk_rowbase[c];
    // E consecutive rows per lane (one 16-byte load), 4E columns per pass
    constexpr int E = Vec<T>::E, G = 64 / E, NP = 64 / (4 * E);
    const int rr = E * (lane % G), kc = lane / G;
    const T *P = coef + S.cblk_coef[c];
    T a[NP][E], xv[NR];
    auto gload = [&](int r0) {
        const int nr = min(64, it.r1 - r0);
#pragma unroll
        for (int u = 0; u < NP; u++) {
            const int cc = q * E + kc + 4 * E * u;
            ld_rows(P + r0 + rr + (int64_t)(k0 + cc) * ld, rr, (k0 + cc < Kc) ? nr : 0, a[u]);
        }
        if (tid < 64) {
#pragma unroll
            for (int v = 0; v < NR; v++)
                xv[v] = (tid < nr && v < nrhs) ? ld_co<CO>(x + S.panel_grow[rb + r0 + tid] + v * ldx) : (T)0;
        }
    };
    T acc[NR];
#pragma unroll
    for (int u = 0; u < NR; u++) acc[u] = 0;
    gload(it.r0);
    for (int r0 = it.r0; r0 < it.r1; r0 += 64) {
        // the chunk in registers goes to LDS, the next one is fetched
#pragma unroll
        for (int u = 0; u < NP; u++) {
            const int cc = q * E + kc + 4 * E * u;
#pragma unroll
            for (int i = 0; i < E; i++) tile[rr + i][cc] = a[u][i];
        }
        if (tid < 64) {
#pragma unroll
            for (int v = 0; v < NR; v++) xr[tid + v * 64] = xv[v];
        }
        __syncthreads();
        if (r0 + 64 < it.r1) gload(r0 + 64);
#pragma unroll 4
        for (int row = q * 16; row < q * 16 + 16; row++) {
            const T t = tile[row][lane];
#pragma unroll
            for (int u = 0; u < NR; u++) acc[u] += t * xr[row + u * 64];
        }
        __syncthreads();
    }
    T *out = part + (int64_t)it.slot * nrhs_tot + (int64_t)u0 * Kc + k0 + lane;
    wave_sum_store<CO, T, NR, 4>(acc, red, lane, q, k0 + lane < Kc, nrhs, out, Kc, (const T *)nullptr, 0);
}

// nrhs right-hand sides starting at column u0 of the partials, which are laid
// out for nrhs_tot columns (z and x point at column u0 already)
template <bool CO, typename T, int NR, int NW, bool PRE = false>
__device__ __forceinline__ void bwd_diag_body(const DevSymbol &S, const DiagItem it,
                                              const BwdItem *__restrict__ bitems, const T *__restrict__ z,
                                              int64_t ldz, T *__restrict__ x, int64_t ldx, int nrhs, int nrhs_tot,
                                              int u0, const T *__restrict__ part, const T *__restrict__ wb, T *y,
                                              T *red) {
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int nd = it.nd, r0 = it.ib * 64;
    const int cb0 = S.node_cblk[nd], cb1 = S.node_cblk[nd + 1];
    const int fc = S.cblk_fcol[cb0];
    const int Kc = S.cblk_fcol[cb1] - fc;
    // Column k of the node belongs to member cblk m (column k - o of it, o
    // = fcol[m] - fc); m's strips [tb, te) own consecutive partial slots, w
    // apart, the first at bitems[tb].slot.  pcol returns the address of the
    // partial of right-hand side 0 in m's first strip; the strips follow at
    // stride w * nrhs_tot.
    auto pcol = [&](int k, int &tb, int &te, int64_t &stride) -> const T * {
        int lo = cb0, hi = cb1 - 1;  // the last member whose first column is <= k
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S.cblk_fcol[mid] - fc <= k) lo = mid;
            else hi = mid - 1;
        }
        const int o = S.cblk_fcol[lo] - fc, w = S.cblk_fcol[lo + 1] - S.cblk_fcol[lo];
        tb = S.cblk_btile[2 * lo];
        te = S.cblk_btile[2 * lo + 1];
        stride = (int64_t)w * nrhs_tot;
        const int64_t s0 = (te > tb) ? (int64_t)bitems[tb].slot : 0;
        return part + s0 * nrhs_tot + (int64_t)u0 * w + (k - o);
    };
    if constexpr (PRE) {
        // the partials were subtracted from z already (k_bwd_reduce)
        for (int t = tid; t < (Kc - r0) * nrhs; t += NW * 64) {
            const int k = r0 + t % (Kc - r0), u = t / (Kc - r0);
            y[k + u * Kc] = z[fc + k + u * ldz];
        }
        __syncthreads();
    }
    // y = z_c - the tiles' partials for columns [r0, Kc).  Several
    // right-hand sides: 64 columns at a time, the waves splitting the tiles.
    if constexpr (NR > 1 && !PRE) {
        for (int kc = r0; kc < Kc; kc += 64) {
            const int k = kc + r;
            T acc[NR];
#pragma unroll
            for (int u = 0; u < NR; u++) acc[u] = 0;
            if (k < Kc) {
                int tb, te;
                int64_t stride;
                const T *p0 = pcol(k, tb, te, stride);
                const int w = (int)(stride / nrhs_tot);
#pragma unroll 4
                for (int j = tb + q; j < te; j += NW) {
                    const T *pp = p0 + (int64_t)(j - tb) * stride;
#pragma unroll
                    for (int u = 0; u < NR; u++)
                        if (u < nrhs) acc[u] += ld_co<CO>(pp + u * w);
                }
            }
            // y lives in LDS: a plain store; z comes from other tasks
            wave_sum_store<CO, T, NR, NW>(acc, red, r, q, k < Kc, nrhs, y + k, Kc, z + fc + k, ldz);
            __syncthreads();
        }
    }
    // One right-hand side: all columns at once (cpp per pass), each summed
    // over every L-th tile by L threads, whose sums are then added in thread
    // order; one barrier pair per pass instead of one per 64 columns.
    const int ncol = (NR > 1 || PRE) ? 0 : Kc - r0;
    const int cpp = max(1, min(ncol, NW * 64)), L = (NW * 64) / cpp;
    for (int c0 = 0; c0 < ncol; c0 += cpp) {
        const int cc = tid % cpp, h = tid / cpp;
        const int k = r0 + c0 + cc;
        const bool live = h < L && c0 + cc < ncol;
        T acc[NR];
#pragma unroll
        for (int u = 0; u < NR; u++) acc[u] = 0;
        if (live) {
            int tb, te;
            int64_t stride;
            const T *p0 = pcol(k, tb, te, stride);
            const int w = (int)(stride / nrhs_tot);
#pragma unroll 4
            for (int j = tb + h; j < te; j += L) {
                const T *pp = p0 + (int64_t)(j - tb) * stride;
#pragma unroll
                for (int u = 0; u < NR; u++)
                    if (u < nrhs) acc[u] += ld_co<CO>(pp + u * w);
            }
        }
        if (h < L) {
#pragma unroll
            for (int u = 0; u < NR; u++) red[(h * cpp + cc) * NR + u] = acc[u];
        }
        __syncthreads();
        if (h == 0 && c0 + cc < ncol) {
#pragma unroll
            for (int u = 0; u < NR; u++) {
                if (u < nrhs) {
                    T v = red[cc * NR + u];
                    for (int w = 1; w < L; w++) v += red[(w * cpp + cc) * NR + u];
                    // y lives in LDS: a plain store; z comes from other tasks
                    y[k + u * Kc] = ld_co<CO>(z + fc + k + u * ldz) - v;
                }
            }
        }
        __syncthreads();
    }
    const int i = r0 + r;
    // rows of the inverse: E consecutive per lane (16-byte loads)
    constexpr int E = Vec<T>::E, G = 64 / E;
    const int rr = E * (r % G), kc = r / G;
    T acc2[E][NR];
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
        for (int u = 0; u < NR; u++) acc2[e][u] = 0;
    if (r0 + rr < Kc) {
        // wb holds inv(L_cc)^T (or inv(U_cc)) column-major: upper triangular
        const T *M = wb + S.node_winv[nd] + r0 + rr;
#pragma unroll 4
        for (int k = r0 + q * E + kc; k < Kc; k += NW * E) {
            T m[E];
            ld_rows(M + (int64_t)k * Kc, rr, Kc - r0, m);
#pragma unroll
            for (int u = 0; u < NR; u++) {
                if (u < nrhs) {
                    const T yv = y[k + u * Kc];
#pragma unroll
                    for (int e = 0; e < E; e++) acc2[e][u] += m[e] * yv;
                }
            }
        }
    }
    wave_sum_store_rows<CO, T, NR, NW, E>(acc2, red, r, q, i < Kc, nrhs, x + fc + i, ldx);
}

// ------------------------------ level kernels -------------------------------
template <typename T, int NR, int NW, bool PRE = false>
__global__ __launch_bounds__(NW * 64) void k_fwd_diag(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                      const YSlices ys, const T *__restrict__ ybuf, int64_t ldy,
                                                      const T *__restrict__ b, int64_t ldb, T *__restrict__ z,
                                                      int64_t ldz, int nrhs, const T *__restrict__ w1,
                                                      const T *__restrict__ w1t) {
    extern __shared__ char smem_raw[];
    __shared__ T red[NW * 64 * NR];
    // the blocks with the most rows to gather and columns to read (the last
    // ones of a node) take the lowest workgroup ids, so that they start first
    const int id = (int)gridDim.x - 1 - xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    // right-hand sides [NR*blockIdx.y, NR*blockIdx.y + NR)
    const int u0 = NR * blockIdx.y;
    fwd_diag_body<false, T, NR, NW, PRE>(S, items[id], ys, ybuf + u0 * ldy, ldy, b + u0 * ldb, ldb, z + u0 * ldz,
                                    ldz, min(NR, nrhs - u0), w1, w1t, reinterpret_cast<T *>(smem_raw), red);
}

template <typename T, int NR, int NW>
__global__ __launch_bounds__(NW * 64) void k_fwd_prod(DevSymbol S, const FwdItem *__restrict__ items, int nitems,
                                                      const T *__restrict__ coef, const T *__restrict__ z,
                                                      int64_t ldz, int nrhs, const int32_t *__restrict__ ydst,
                                                      T *__restrict__ ybuf, int64_t ldy) {
    extern __shared__ char smem_raw[];
    __shared__ T red[NW * 64 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    // right-hand sides [NR*blockIdx.y, NR*blockIdx.y + NR)
    const int u0 = NR * blockIdx.y;
    fwd_prod_body<false, T, NR, NW>(S, items[id], coef, z + u0 * ldz, ldz, min(NR, nrhs - u0), ydst, ybuf + u0 * ldy,
                                    ldy, reinterpret_cast<T *>(smem_raw), red);
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_bwd_prod(DevSymbol S, const BwdItem *__restrict__ items, int nitems,
                                                  const T *__restrict__ coef, const T *__restrict__ x, int64_t ldx,
                                                  int nrhs, T *__restrict__ part) {
    __shared__ T tile[64][65];
    __shared__ T xr[64 * NR];
    __shared__ T red[256 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const BwdItem it = items[id];
    const int k0 = blockIdx.y * 64;
    if (k0 >= S.cblk_fcol[it.c + 1] - S.cblk_fcol[it.c]) return;
    // right-hand sides [NR*blockIdx.z, NR*blockIdx.z + NR)
    const int u0 = NR * blockIdx.z;
    bwd_prod_body<false, T, NR>(S, it, k0, coef, x + u0 * ldx, ldx, min(NR, nrhs - u0), nrhs, u0, part, tile, xr,
                                red);
}

// Backward strips of narrow cblks (at most C columns), one right-hand side:
// four strips per wave, 16 lanes each; lane l sums rows l, l + 16, ... of
// every column in registers, the 16 lanes combine by xor shuffles, lane j
// stores column j.
template <typename T, int C>
__global__ __launch_bounds__(64) void k_bwd_prod_narrow(DevSymbol S, const BwdItem *__restrict__ items, int nitems,
                                                        const T *__restrict__ coef, const T *__restrict__ x,
                                                        T *__restrict__ part) {
    const int lane = threadIdx.x, grp = lane >> 4, l = lane & 15;
    const int id = xcd_remap(blockIdx.x, gridDim.x) * 4 + grp;
    BwdItem it{0, 0, 0, 0};
    int Kc = 0;
    int64_t ld = 0, rb = 0;
    const T *P = coef;
    if (id < nitems) {
        it = items[id];
        Kc = S.cblk_fcol[it.c + 1] - S.cblk_fcol[it.c];
        ld = S.cblk_stride[it.c];
        rb = S.cblk_rowbase[it.c];
        P = coef + S.cblk_coef[it.c];
    }
    T acc[C];
#pragma unroll
    for (int k = 0; k < C; k++) acc[k] = 0;
    for (int r = it.r0 + l; r < it.r1; r += 16) {
        const T xr = x[S.panel_grow[rb + r]];
#pragma unroll
        for (int k = 0; k < C; k++)
            if (k < Kc) acc[k] += P[r + (int64_t)k * ld] * xr;
    }
#pragma unroll
    for (int off = 1; off < 16; off *= 2)
#pragma unroll
        for (int k = 0; k < C; k++) acc[k] += __shfl_xor(acc[k], off);
    // lane j of the group stores columns j, j + 16, ...
#pragma unroll
    for (int k = 0; k < C; k++)
        if ((k & 15) == l && k < Kc) part[(int64_t)it.slot + k] = acc[k];
}

template <typename T, int NR, int NW, bool PRE = false>
__global__ __launch_bounds__(NW * 64) void k_bwd_diag(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                      const BwdItem *__restrict__ bitems, const T *__restrict__ z,
                                                      int64_t ldz, T *__restrict__ x, int64_t ldx, int nrhs,
                                                      const T *__restrict__ part, const T *__restrict__ wb) {
    extern __shared__ char smem_raw[];
    __shared__ T red[NW * 64 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    // right-hand sides [NR*blockIdx.y, NR*blockIdx.y + NR)
    const int u0 = NR * blockIdx.y;
    bwd_diag_body<false, T, NR, NW, PRE>(S, items[id], bitems, z + u0 * ldz, ldz, x + u0 * ldx, ldx, min(NR, nrhs - u0),
                                    nrhs, u0, part, wb, reinterpret_cast<T *>(smem_raw), red);
}


// Levels of wide nodes (more than SOLVE_SPLIT columns): the gathers run once
// per 64-row block in a launch of their own instead of in every diagonal
// workgroup that needs the rows (workgroup ib needs rows [0, 64 ib + 64)).
// Forward: b minus the rows' update-buffer entries, in place in b; wave q
// sums the entries q, q + 4, ... of the block's rows (lane = row, so the
// sliced buffer's loads are contiguous), the waves combined in order.  One
// right-hand side per grid row.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_gather(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                    const YSlices ys, const T *__restrict__ ybuf, int64_t ldy,
                                                    T *__restrict__ b, int64_t ldb) {
    __shared__ T red[4][64];
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int u = blockIdx.y;
    const DiagItem it = items[id];
    const int fc = S.cblk_fcol[S.node_cblk[it.nd]];
    const int K = S.cblk_fcol[S.node_cblk[it.nd + 1]] - fc;
    const int i = it.ib * 64 + lane;
    constexpr int UL = 8;
    T v = 0;
    if (i < K) {
        const int g = fc + i, ne = ys.cnt[g];
        const T *yb = ybuf + u * ldy + ys.base[g >> 6] + (g & 63);
        for (int e = q; e < ne; e += 4 * UL) {
            T a[UL];
#pragma unroll
            for (int k = 0; k < UL; k++) a[k] = (e + 4 * k < ne) ? yb[(int64_t)(e + 4 * k) * 64] : (T)0;
#pragma unroll
            for (int k = 0; k < UL; k++) v += a[k];
        }
    }
    red[q][lane] = v;
    __syncthreads();
    if (q == 0 && i < K) b[fc + i + u * ldb] -= ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Backward: z minus the strips' partials of the block's columns, in place in
// z; wave q sums the strips q, q + 4, ... of each column's member (lane =
// column), the waves combined in order.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_reduce(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                    const BwdItem *__restrict__ bitems, T *__restrict__ z,
                                                    int64_t ldz, const T *__restrict__ part, int nrhs_tot) {
    __shared__ T red[4][64];
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int u = blockIdx.y;
    const DiagItem it = items[id];
    const int cb0 = S.node_cblk[it.nd], cb1 = S.node_cblk[it.nd + 1];
    const int fc = S.cblk_fcol[cb0];
    const int K = S.cblk_fcol[cb1] - fc;
    const int k = it.ib * 64 + lane;
    T v = 0;
    if (k < K) {
        int lo = cb0, hi = cb1 - 1;  // the last member whose first column is <= k
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S.cblk_fcol[mid] - fc <= k) lo = mid;
            else hi = mid - 1;
        }
        const int o = S.cblk_fcol[lo] - fc, w = S.cblk_fcol[lo + 1] - S.cblk_fcol[lo];
        const int tb = S.cblk_btile[2 * lo], te = S.cblk_btile[2 * lo + 1];
        if (te > tb) {
            const int64_t stride = (int64_t)w * nrhs_tot;
            const T *p0 = part + (int64_t)bitems[tb].slot * nrhs_tot + (int64_t)u * w + (k - o);
            for (int j = q; j < te - tb; j += 4) v += p0[(int64_t)j * stride];
        }
    }
    red[q][lane] = v;
    __syncthreads();
    if (q == 0 && k < K) z[fc + k + u * ldz] -= ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Wide levels, one right-hand side, after the gathers: the diagonal GEMV on
// 16-row blocks (four workgroups per 64-row item), so that a node's inverse
// is read by four times as many compute units.  Forward (UPPER false): rows
// [r, r+16) of z = W y, y = b's node rows (already gathered), columns [0,
// r+16); backward (UPPER true): rows of x = WB y, y = z's node rows,
// columns [r, K).  Thread t keeps E consecutive rows (16-byte loads) and the
// columns c, c + 64/... of its column phase; the phases combine through LDS
// in a fixed order.
template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void k_diag_gemv16(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                     const T *__restrict__ src, T *__restrict__ dst,
                                                     const T *__restrict__ w) {
    constexpr int E = Vec<T>::E, RL = 16 / E, NC = 256 / RL;  // lanes per column, column phases
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);
    __shared__ T red[NC][17];
    const int tid = threadIdx.x;
    // forward: the blocks with the most columns (the last rows of a node)
    // take the lowest workgroup ids, so that they start first
    const int g0 = xcd_remap(blockIdx.x, gridDim.x);
    const int gid = UPPER ? g0 : (int)gridDim.x - 1 - g0;
    const int id = gid >> 2, sub = gid & 3;
    if (id >= nitems) return;
    const DiagItem it = items[id];
    const int fc = S.cblk_fcol[S.node_cblk[it.nd]];
    const int K = S.cblk_fcol[S.node_cblk[it.nd + 1]] - fc;
    const int r = it.ib * 64 + sub * 16;
    if (r >= K) return;  // uniform: the whole workgroup leaves
    const int nr = min(16, K - r);
    const int k0 = UPPER ? r : 0, k1 = UPPER ? K : min(K, r + 16);
    for (int k = k0 + tid; k < k1; k += 256) y[k - k0] = src[fc + k];
    __syncthreads();
    const int rr = E * (tid % RL), c = tid / RL;
    const T *M = w + S.node_winv[it.nd] + r + rr;
    T acc[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = 0;
#pragma unroll 4
    for (int k = k0 + c; k < k1; k += NC) {
        T m[E];
        ld_rows(M + (int64_t)k * K, rr, nr, m);
        const T yv = y[k - k0];
#pragma unroll
        for (int e = 0; e < E; e++) acc[e] += m[e] * yv;
    }
#pragma unroll
    for (int e = 0; e < E; e++) red[c][rr + e] = acc[e];
    __syncthreads();
    if (tid < nr) {
        T v = red[0][tid];
        for (int q = 1; q < NC; q++) v += red[q][tid];
        dst[fc + r + tid] = v;
    }
}

// Forward, LL^t, one right-hand side, after the gathers: rows [r, r+16) of
// z = W y as dot products down columns r..r+15 of WT = W^T (the backward
// inverse): row i of W is column i of WT, one contiguous run, so a block
// reads 16 adjacent columns instead of 16 rows strided across the whole
// inverse (FP64 80^3: 305 -> 175 us per solve).  Sixteen lanes per output
// row, E consecutive k per lane (16-byte loads); the lanes' sums combine by
// xor shuffles in a fixed order.  (The same form for the backward GEMV,
// down the columns of the forward inverse, was slower: 178 against 150 us.)
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_gemv16t(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                     const T *__restrict__ src, T *__restrict__ dst,
                                                     const T *__restrict__ wt) {
    constexpr int E = Vec<T>::E;
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);
    const int tid = threadIdx.x;
    // the blocks with the most columns (the last rows of a node) take the
    // lowest workgroup ids, so that they start first
    const int gid = (int)gridDim.x - 1 - xcd_remap(blockIdx.x, gridDim.x);
    const int id = gid >> 2, sub = gid & 3;
    if (id >= nitems) return;
    const DiagItem it = items[id];
    const int fc = S.cblk_fcol[S.node_cblk[it.nd]];
    const int K = S.cblk_fcol[S.node_cblk[it.nd + 1]] - fc;
    const int r = it.ib * 64 + sub * 16;
    if (r >= K) return;  // uniform: the whole workgroup leaves
    const int k1 = min(K, r + 16);
    for (int k = tid; k < k1; k += 256) y[k] = src[fc + k];
    __syncthreads();
    const int j = tid >> 4, l = tid & 15;
    const int i = r + j;  // output row; W[i, k] = WT[k, i] is nonzero for k <= i
    T acc = 0;
    if (i < K) {
        const T *col = wt + S.node_winv[it.nd] + (int64_t)i * K;
        T a2[E];
#pragma unroll
        for (int e = 0; e < E; e++) a2[e] = 0;
#pragma unroll 4
        for (int k = E * l; k <= i; k += 16 * E) {
            T m[E];
            ld_rows(col + k, k, i + 1, m);
#pragma unroll
            for (int e = 0; e < E; e++) a2[e] += m[e] * y[min(k + e, i)];
        }
#pragma unroll
        for (int e = 0; e < E; e++) acc += a2[e];
    }
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 16);
    if (l == 0 && i < K) dst[fc + i] = acc;
}

// Levels of narrow nodes (at most SOLVE_TINY columns), one right-hand side:
// four nodes per wave, 16 lanes each, lane i owning row (forward) or column
// (backward) i of its node; the node's inverse is read into registers
// before the gathers, the product takes y by lane shuffles.
template <typename T>
__global__ __launch_bounds__(64) void k_fwd_diag_tiny(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                      const YSlices ys, const T *__restrict__ ybuf,
                                                      const T *__restrict__ b, T *__restrict__ z,
                                                      const T *__restrict__ w1) {
    const int lane = threadIdx.x, grp = lane >> 4, i = lane & 15;
    const int id = xcd_remap(blockIdx.x, gridDim.x) * 4 + grp;
    int fc = 0, K = 0;
    int64_t wo = 0;
    if (id < nitems) {
        const int nd = items[id].nd;
        fc = S.cblk_fcol[S.node_cblk[nd]];
        K = S.cblk_fcol[S.node_cblk[nd + 1]] - fc;
        wo = S.node_winv[nd];
    }
    const bool row = i < K;
    T wv[16];
#pragma unroll
    for (int k = 0; k < 16; k++) wv[k] = (row && k < K) ? w1[wo + i + (int64_t)k * K] : (T)0;
    T y = 0;
    if (row) {
        const int g = fc + i, ne = ys.cnt[g];
        const T *yb = ybuf + ys.base[g >> 6] + (g & 63);
        T v = 0;
        for (int e = 0; e < ne; e += 8) {
            T a[8];
#pragma unroll
            for (int k = 0; k < 8; k++) a[k] = (e + k < ne) ? yb[(int64_t)(e + k) * 64] : (T)0;
#pragma unroll
            for (int k = 0; k < 8; k++) v += a[k];
        }
        y = b[g] - v;
    }
    T acc = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) acc += wv[k] * __shfl(y, (grp << 4) + k);
    if (row) z[fc + i] = acc;
}

template <typename T>
__global__ __launch_bounds__(64) void k_bwd_diag_tiny(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                      const BwdItem *__restrict__ bitems, const T *__restrict__ z,
                                                      T *__restrict__ x, const T *__restrict__ part,
                                                      const T *__restrict__ wb) {
    const int lane = threadIdx.x, grp = lane >> 4, k = lane & 15;
    const int id = xcd_remap(blockIdx.x, gridDim.x) * 4 + grp;
    int fc = 0, K = 0, cb0 = 0, cb1 = 0;
    int64_t wo = 0;
    if (id < nitems) {
        const int nd = items[id].nd;
        cb0 = S.node_cblk[nd];
        cb1 = S.node_cblk[nd + 1];
        fc = S.cblk_fcol[cb0];
        K = S.cblk_fcol[cb1] - fc;
        wo = S.node_winv[nd];
    }
    const bool col = k < K;
    // row k of inv(L_nd)^T: WB[k + j*K], j >= k
    T wv[16];
#pragma unroll
    for (int j = 0; j < 16; j++) wv[j] = (col && j < K) ? wb[wo + k + (int64_t)j * K] : (T)0;
    T y = 0;
    if (col) {
        // member m of column k, its strips' partials (consecutive slots, w apart)
        int m = cb0;
        while (m + 1 < cb1 && S.cblk_fcol[m + 1] - fc <= k) m++;
        const int o = S.cblk_fcol[m] - fc, w = S.cblk_fcol[m + 1] - S.cblk_fcol[m];
        const int tb = S.cblk_btile[2 * m], te = S.cblk_btile[2 * m + 1];
        T v = 0;
        if (te > tb) {
            const T *pp = part + (int64_t)bitems[tb].slot + (k - o);
            for (int j = 0; j < te - tb; j++) v += pp[(int64_t)j * w];
        }
        y = z[fc + k] - v;
    }
    T acc = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) acc += wv[j] * __shfl(y, (grp << 4) + j);
    if (col) x[fc + k] = acc;
}

// ------------------------------- launchers ---------------------------------
template <typename T>
void launch_fwd_diag(const DevSymbol &S, const DiagItem *items, int n, const YSlices &ys, const T *ybuf,
                     int64_t ldy, T *b, int64_t ldb, T *z, int64_t ldz, int nrhs, const T *w1, const T *w1t,
                     int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    if (maxwidth > SOLVE_SPLIT) {
        hipLaunchKernelGGL(k_fwd_gather<T>, dim3(grid8(n), nrhs), dim3(256), 0, st, S, items, n, ys, ybuf, ldy, b, ldb);
        if (nrhs == 1 && w1t)
            hipLaunchKernelGGL(k_fwd_gemv16t<T>, dim3(grid8(4 * n)), dim3(256), (size_t)maxwidth * sizeof(T), st, S,
                               items, n, (const T *)b, z, w1t);
        else if (nrhs == 1)
            hipLaunchKernelGGL((k_diag_gemv16<T, false>), dim3(grid8(4 * n)), dim3(256), (size_t)maxwidth * sizeof(T),
                               st, S, items, n, (const T *)b, z, w1);
        else
            hipLaunchKernelGGL((k_fwd_diag<T, 2, 16, true>), dim3(grid8(n), (nrhs + 1) / 2), dim3(1024),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, ys, ybuf, ldy, b, ldb, z, ldz,
                               nrhs, w1, (const T *)nullptr);
        return;
    }
    // several right-hand sides: pairs of columns on the grid's y dimension,
    // so that the few diagonal blocks of the top levels use more of the chip;
    // levels of narrow cblks (maxwidth = the level's widest) take 4-wave
    // workgroups
    const bool narrow = maxwidth <= SOLVE_NARROW;
    if (nrhs == 1) {
        if (maxwidth <= SOLVE_TINY)
            hipLaunchKernelGGL(k_fwd_diag_tiny<T>, dim3(grid8((n + 3) / 4)), dim3(64), 0, st, S, items, n, ys, ybuf, b,
                               z, w1);
        else if (narrow)
            hipLaunchKernelGGL((k_fwd_diag<T, 1, 4>), dim3(grid8(n)), dim3(256), (size_t)maxwidth * sizeof(T), st, S,
                               items, n, ys, ybuf, ldy, b, ldb, z, ldz, nrhs, w1, w1t);
        else
            hipLaunchKernelGGL((k_fwd_diag<T, 1, 16>), dim3(grid8(n)), dim3(1024), (size_t)maxwidth * sizeof(T), st,
                               S, items, n, ys, ybuf, ldy, b, ldb, z, ldz, nrhs, w1, w1t);
    } else {
        if (narrow)
            hipLaunchKernelGGL((k_fwd_diag<T, 2, 4>), dim3(grid8(n), (nrhs + 1) / 2), dim3(256),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, ys, ybuf, ldy, b, ldb, z, ldz,
                               nrhs, w1, (const T *)nullptr);
        else
            hipLaunchKernelGGL((k_fwd_diag<T, 2, 16>), dim3(grid8(n), (nrhs + 1) / 2), dim3(1024),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, ys, ybuf, ldy, b, ldb, z, ldz,
                               nrhs, w1, (const T *)nullptr);
    }
}
template <typename T>
void launch_fwd_prod(const DevSymbol &S, const FwdItem *items, int n, const T *coef, const T *z, int64_t ldz,
                     int nrhs, const int32_t *ydst, T *ybuf, int64_t ldy, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    if (nrhs == 1 && maxwidth <= SOLVE_TINY)
        hipLaunchKernelGGL((k_fwd_prod<T, 1, 1>), dim3(grid8(n)), dim3(64), (size_t)maxwidth * sizeof(T), st, S,
                           items, n, coef, z, ldz, nrhs, ydst, ybuf, ldy);
    else if (nrhs == 1 && maxwidth <= SOLVE_NARROW)
        hipLaunchKernelGGL((k_fwd_prod<T, 1, 4>), dim3(grid8(n)), dim3(256), (size_t)maxwidth * sizeof(T), st, S,
                           items, n, coef, z, ldz, nrhs, ydst, ybuf, ldy);
    else if (nrhs == 1)
        hipLaunchKernelGGL((k_fwd_prod<T, 1, 16>), dim3(grid8(n)), dim3(1024), (size_t)maxwidth * sizeof(T), st, S,
                           items, n, coef, z, ldz, nrhs, ydst, ybuf, ldy);
    else
        hipLaunchKernelGGL((k_fwd_prod<T, 8, 8>), dim3(grid8(n), (nrhs + 7) / 8), dim3(512),
                           (size_t)maxwidth * 8 * sizeof(T), st, S, items, n, coef, z, ldz, nrhs, ydst, ybuf, ldy);
}
template <typename T>
void launch_bwd_prod(const DevSymbol &S, const BwdItem *items, int n, const T *coef, const T *x, int64_t ldx,
                     int nrhs, T *part, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    if (nrhs == 1 && maxwidth <= SOLVE_TINY) {
        hipLaunchKernelGGL((k_bwd_prod_narrow<T, SOLVE_TINY>), dim3(grid8((n + 3) / 4)), dim3(64), 0, st, S, items, n,
                           coef, x, part);
        return;
    }
    const dim3 g(grid8(n), (maxwidth + 63) / 64, nrhs == 1 ? 1 : (nrhs + 7) / 8);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_bwd_prod<T, 1>), g, dim3(256), 0, st, S, items, n, coef, x, ldx, nrhs, part);
    else
        hipLaunchKernelGGL((k_bwd_prod<T, 8>), g, dim3(256), 0, st, S, items, n, coef, x, ldx, nrhs, part);
}
template <typename T>
void launch_bwd_diag(const DevSymbol &S, const DiagItem *items, int n, const BwdItem *bitems, T *z,
                     int64_t ldz, T *x, int64_t ldx, int nrhs, const T *part, const T *wb, int maxwidth,
                     hipStream_t st) {
    if (n <= 0) return;
    if (maxwidth > SOLVE_SPLIT) {
        hipLaunchKernelGGL(k_bwd_reduce<T>, dim3(grid8(n), nrhs), dim3(256), 0, st, S, items, n, bitems, z, ldz, part,
                           nrhs);
        if (nrhs == 1)
            hipLaunchKernelGGL((k_diag_gemv16<T, true>), dim3(grid8(4 * n)), dim3(256), (size_t)maxwidth * sizeof(T),
                               st, S, items, n, (const T *)z, x, wb);
        else
            hipLaunchKernelGGL((k_bwd_diag<T, 2, 16, true>), dim3(grid8(n), (nrhs + 1) / 2), dim3(1024),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, bitems, z, ldz, x, ldx, nrhs, part,
                               wb);
        return;
    }
    const bool narrow = maxwidth <= SOLVE_NARROW;
    if (nrhs == 1) {
        if (maxwidth <= SOLVE_TINY)
            hipLaunchKernelGGL(k_bwd_diag_tiny<T>, dim3(grid8((n + 3) / 4)), dim3(64), 0, st, S, items, n, bitems, z, x,
                               part, wb);
        else if (narrow)
            hipLaunchKernelGGL((k_bwd_diag<T, 1, 4>), dim3(grid8(n)), dim3(256), (size_t)maxwidth * sizeof(T), st, S,
                               items, n, bitems, z, ldz, x, ldx, nrhs, part, wb);
        else
            hipLaunchKernelGGL((k_bwd_diag<T, 1, 16>), dim3(grid8(n)), dim3(1024), (size_t)maxwidth * sizeof(T), st,
                               S, items, n, bitems, z, ldz, x, ldx, nrhs, part, wb);
    } else {
        if (narrow)
            hipLaunchKernelGGL((k_bwd_diag<T, 2, 4>), dim3(grid8(n), (nrhs + 1) / 2), dim3(256),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, bitems, z, ldz, x, ldx, nrhs, part,
                               wb);
        else
            hipLaunchKernelGGL((k_bwd_diag<T, 2, 16>), dim3(grid8(n), (nrhs + 1) / 2), dim3(1024),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, bitems, z, ldz, x, ldx, nrhs, part,
                               wb);
    }
}


#define INST(T)                                                                                                   \
    template void launch_fwd_diag<T>(const DevSymbol &, const DiagItem *, int, const YSlices &, const T *,       \
                                     int64_t, T *, int64_t, T *, int64_t, int, const T *, const T *, int,        \
                                     hipStream_t);                                                               \
    template void launch_fwd_prod<T>(const DevSymbol &, const FwdItem *, int, const T *, const T *, int64_t, int, \
                                     const int32_t *, T *, int64_t, int, hipStream_t);                           \
    template void launch_bwd_prod<T>(const DevSymbol &, const BwdItem *, int, const T *, const T *, int64_t, int, \
                                     T *, int, hipStream_t);                                                     \
    template void launch_bwd_diag<T>(const DevSymbol &, const DiagItem *, int, const BwdItem *, T *, int64_t,     \
                                     T *, int64_t, int, const T *, const T *, int, hipStream_t);
INST(double)
INST(float)


}  // namespace pxa
