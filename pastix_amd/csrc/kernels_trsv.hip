// Supernodal forward / backward solve kernels for CDNA4 (gfx950)
// (reference kernels/solve_ztrsmsp.c:67-557, sopalin/sequential_ztrsm.c:34-103).
//
// The sweeps read b from x and leave the solution in x; the forward result
// z = inv(L) b lives in a separate buffer so that every workgroup of a cblk
// can read the cblk's input while others write their rows of the output.
// Forward sweep, level by level:
//   k_fwd_diag : workgroup (c, ib) gathers, for rows [0, 64*ib+64) of cblk c,
//                b minus the update-buffer entries of these rows (stored
//                contiguously per row, sources in increasing order), then
//                writes rows [64*ib, 64*ib+64) of z_c = inv(L_cc) y;
//   k_fwd_prod : L_off z_c on 64-row tiles, each product row stored at its
//                entry of the update buffer.
// Backward sweep, levels from the root down:
//   k_bwd_prod : partial[j] = sum over a 64x64 tile of L[r, j] x[row(r)];
//   k_bwd_diag : workgroup (c, ib) forms y = z_c - the tiles' partials for
//                columns [64*ib, K) and writes rows [64*ib, 64*ib+64) of
//                x_c = inv(L_cc)^T y.
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
template <bool CO, typename T, int NR, int NW>
__device__ __forceinline__ void fwd_diag_body(const DevSymbol &S, const DiagItem it,
                                              const int64_t *__restrict__ yrow_ptr, const T *__restrict__ ybuf,
                                              int64_t ldy, const T *__restrict__ b, int64_t ldb, T *__restrict__ z,
                                              int64_t ldz, int nrhs, const T *__restrict__ w1, T *y, T *red) {
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int c = it.c, r0 = it.ib * 64;
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    const int nrow = min(Kc, r0 + 64);
    for (int t = tid; t < nrow * nrhs; t += NW * 64) {
        const int i = t % nrow, u = t / nrow;
        const int g = fc + i;
        const T *yb = ybuf + u * ldy;
        T v = b[g + u * ldb];
        int64_t p = yrow_ptr[g];
        const int64_t pe = yrow_ptr[g + 1];
#pragma unroll 4
        for (; p + 4 <= pe; p += 4)
            v -= ((ld_co<CO>(yb + p) + ld_co<CO>(yb + p + 1)) + (ld_co<CO>(yb + p + 2) + ld_co<CO>(yb + p + 3)));
        for (; p < pe; p++) v -= ld_co<CO>(yb + p);
        y[i + u * Kc] = v;
    }
    __syncthreads();
    const int i = r0 + r;
    // rows of the inverse: E consecutive per lane (16-byte loads)
    constexpr int E = Vec<T>::E, G = 64 / E;
    const int rr = E * (r % G), kc = r / G;
    T acc[E][NR];
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
        for (int u = 0; u < NR; u++) acc[e][u] = 0;
    if (r0 + rr < Kc) {
        const T *M = w1 + S.cblk_winv[c] + r0 + rr;
#pragma unroll 4
        for (int k = q * E + kc; k < nrow; k += NW * E) {
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

// one 64x64 tile (column block k0/64) of the backward products, 4 waves
// nrhs right-hand sides starting at column u0 of the partials (laid out for
// nrhs_tot columns); x points at column u0 already
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
    const int nr = it.r1 - it.r0;
    for (int t = tid; t < 64 * NR; t += 256) {
        const int row = t & 63, u = t >> 6;
        xr[t] = (row < nr && u < nrhs) ? ld_co<CO>(x + S.panel_grow[rb + it.r0 + row] + u * ldx) : (T)0;
    }
    const T *P = coef + S.cblk_coef[c] + it.r0;
    {
        // E consecutive rows per lane (one 16-byte load), 4E columns per pass
        constexpr int E = Vec<T>::E, G = 64 / E;
        const int rr = E * (lane % G), kc = lane / G;
#pragma unroll
        for (int cc = q * E + kc; cc < 64; cc += 4 * E) {
            T a[E];
            ld_rows(P + rr + (int64_t)(k0 + cc) * ld, rr, (k0 + cc < Kc) ? nr : 0, a);
#pragma unroll
            for (int i = 0; i < E; i++) tile[rr + i][cc] = a[i];
        }
    }
    __syncthreads();
    T acc[NR];
#pragma unroll
    for (int u = 0; u < NR; u++) acc[u] = 0;
#pragma unroll 4
    for (int row = q * 16; row < q * 16 + 16; row++) {
        const T a = tile[row][lane];
#pragma unroll
        for (int u = 0; u < NR; u++) acc[u] += a * xr[row + u * 64];
    }
    T *out = part + (int64_t)it.slot * nrhs_tot + (int64_t)u0 * Kc + k0 + lane;
    wave_sum_store<CO, T, NR, 4>(acc, red, lane, q, k0 + lane < Kc, nrhs, out, Kc, (const T *)nullptr, 0);
}

// nrhs right-hand sides starting at column u0 of the partials, which are laid
// out for nrhs_tot columns (z and x point at column u0 already)
template <bool CO, typename T, int NR, int NW>
__device__ __forceinline__ void bwd_diag_body(const DevSymbol &S, const DiagItem it,
                                              const BwdItem *__restrict__ bitems, const T *__restrict__ z,
                                              int64_t ldz, T *__restrict__ x, int64_t ldx, int nrhs, int nrhs_tot,
                                              int u0, const T *__restrict__ part, const T *__restrict__ wb, T *y,
                                              T *red) {
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int c = it.c, r0 = it.ib * 64;
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    // the tiles of one cblk own consecutive partial slots, Kc apart
    const int64_t s0 = (it.b1 > it.b0) ? (int64_t)bitems[it.b0].slot : 0;
    // y = z_c - the tiles' partials for columns [r0, Kc).  Several
    // right-hand sides: 64 columns at a time, the waves splitting the tiles.
    if constexpr (NR > 1) {
        for (int kc = r0; kc < Kc; kc += 64) {
            const int k = kc + r;
            T acc[NR];
#pragma unroll
            for (int u = 0; u < NR; u++) acc[u] = 0;
            if (k < Kc) {
#pragma unroll 4
                for (int j = it.b0 + q; j < it.b1; j += NW) {
                    const T *pp = part + (s0 + (int64_t)(j - it.b0) * Kc) * nrhs_tot + (int64_t)u0 * Kc + k;
#pragma unroll
                    for (int u = 0; u < NR; u++)
                        if (u < nrhs) acc[u] += ld_co<CO>(pp + u * Kc);
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
    const int ncol = (NR > 1) ? 0 : Kc - r0;
    const int cpp = max(1, min(ncol, NW * 64)), L = (NW * 64) / cpp;
    for (int c0 = 0; c0 < ncol; c0 += cpp) {
        const int cc = tid % cpp, h = tid / cpp;
        const int k = r0 + c0 + cc;
        const bool live = h < L && c0 + cc < ncol;
        T acc[NR];
#pragma unroll
        for (int u = 0; u < NR; u++) acc[u] = 0;
        if (live) {
#pragma unroll 4
            for (int j = it.b0 + h; j < it.b1; j += L) {
                const T *pp = part + (s0 + (int64_t)(j - it.b0) * Kc) * nrhs_tot + (int64_t)u0 * Kc + k;
#pragma unroll
                for (int u = 0; u < NR; u++)
                    if (u < nrhs) acc[u] += ld_co<CO>(pp + u * Kc);
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
        const T *M = wb + S.cblk_winv[c] + r0 + rr;
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
template <typename T, int NR, int NW>
__global__ __launch_bounds__(NW * 64) void k_fwd_diag(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                      const int64_t *__restrict__ yrow_ptr,
                                                      const T *__restrict__ ybuf, int64_t ldy,
                                                      const T *__restrict__ b, int64_t ldb, T *__restrict__ z,
                                                      int64_t ldz, int nrhs, const T *__restrict__ w1) {
    extern __shared__ char smem_raw[];
    __shared__ T red[NW * 64 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    // right-hand sides [NR*blockIdx.y, NR*blockIdx.y + NR)
    const int u0 = NR * blockIdx.y;
    fwd_diag_body<false, T, NR, NW>(S, items[id], yrow_ptr, ybuf + u0 * ldy, ldy, b + u0 * ldb, ldb, z + u0 * ldz,
                                    ldz, min(NR, nrhs - u0), w1, reinterpret_cast<T *>(smem_raw), red);
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

template <typename T, int NR, int NW>
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
    bwd_diag_body<false, T, NR, NW>(S, items[id], bitems, z + u0 * ldz, ldz, x + u0 * ldx, ldx, min(NR, nrhs - u0),
                                    nrhs, u0, part, wb, reinterpret_cast<T *>(smem_raw), red);
}

// --------------------- dependency-driven solve (one launch) -----------------
// Workgroups take tasks from a queue in the order of the task list (forward
// levels upward, then backward levels downward) and wait, spinning on
// device counters, for the tasks theirs depends on:
//   0 forward diagonal block (c, ib): every forward product tile facing c;
//   1 forward product tile of c: every forward diagonal block of c;
//   2 backward product tile (item, column block): every backward diagonal
//     block of the cblks its rows belong to;
//   3 backward diagonal block of c: every backward product tile of c and
//     every forward diagonal block of c.
// A task only waits for tasks placed before it in the list, which running
// workgroups have already taken, so the queue always drains.  A wait longer
// than SOLVE_SPIN_LIMIT polls raises the error flag, and every workgroup then
// leaves: the launch always ends.  The data handed from task to task is
// written and read with device-coherent accesses, and a task signals only
// after its stores are complete, so a consumer that sees the counter reads
// the producer's values.  Counters: cnt[0..cb) forward products
// received, [cb..2cb) forward diagonal blocks done, [2cb..3cb) backward
// products done, [3cb..4cb) backward diagonal blocks done, cnt[4cb] queue
// head, cnt[4cb+1] error flag; need[0..cb) products facing c, [cb..2cb)
// diagonal blocks of c, [2cb..3cb) backward product tiles of c.
constexpr int SOLVE_SPIN_LIMIT = 1 << 20;

__device__ __forceinline__ bool solve_wait(const int *c, int need, int *err) {
    if (need <= 0) return true;
    int spins = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if (++spins > SOLVE_SPIN_LIMIT) {
            __hip_atomic_fetch_add(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_solve_tasks(DevSymbol S, const SolveTask *__restrict__ tasks,
                                                     const int32_t *__restrict__ chunks, int nchunks,
                                                     const DiagItem *__restrict__ diag,
                                                     const FwdItem *__restrict__ fwd, const BwdItem *__restrict__ bwd,
                                                     const int32_t *__restrict__ tgt_ptr,
                                                     const int32_t *__restrict__ tgt, const int32_t *__restrict__ need,
                                                     int32_t *cnt, int cb, const int64_t *__restrict__ yrow_ptr,
                                                     T *ybuf, int64_t ldy, T *x, int64_t ldx, T *z, int64_t ldz,
                                                     int nrhs, const T *__restrict__ L, const T *__restrict__ U,
                                                     const T *__restrict__ w1, const T *__restrict__ wb,
                                                     const int32_t *__restrict__ ydst, T *part) {
    // dynamic LDS: the y / z staging of the diagonal and forward tasks or the
    // 64x65 tile of a backward product task
    extern __shared__ char smem_raw[];
    T *ysh = reinterpret_cast<T *>(smem_raw);
    T(*tile)[65] = reinterpret_cast<T(*)[65]>(smem_raw);
    __shared__ T red[4 * 64 * NR];
    __shared__ T xr[64 * NR];
    __shared__ int sh_ok, sh_c;
    int *fwd_in = cnt, *fdiag = cnt + cb, *bprod = cnt + 2 * cb, *bdiag = cnt + 3 * cb;
    int *head = cnt + 4 * cb, *err = head + 1;
    const int tid = threadIdx.x;
    int t = 0, tend = 0;  // tasks [t, tend) of the chunk in hand
    for (;;) {
        if (t == tend) {
            // next chunk of the queue
            if (tid == 0) sh_c = __hip_atomic_fetch_add(head, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            const int ck = sh_c;
            if (ck >= nchunks) return;
            t = chunks[ck];
            tend = chunks[ck + 1];
        }
        if (tid == 0) {
            bool ok = true;
            {
                const SolveTask tk = tasks[t];
                const int type = tk.code & 3;
                if (type == 0) {
                    const int c = diag[tk.idx].c;
                    ok = solve_wait(fwd_in + c, need[c], err);
                } else if (type == 1) {
                    const int c = fwd[tk.idx].c;
                    ok = solve_wait(fdiag + c, need[cb + c], err);
                } else if (type == 2) {
                    for (int k = tgt_ptr[tk.idx]; ok && k < tgt_ptr[tk.idx + 1]; k++) {
                        const int f = tgt[k];
                        ok = solve_wait(bdiag + f, need[cb + f], err);
                    }
                } else {
                    const int c = diag[tk.idx].c;
                    ok = solve_wait(fdiag + c, need[cb + c], err) && solve_wait(bprod + c, need[2 * cb + c], err);
                }
            }
            sh_ok = ok ? 1 : 0;
        }
        __syncthreads();
        if (!sh_ok) return;
        const SolveTask tk = tasks[t];
        const int type = tk.code & 3;
        int sig = -1;
        if (type == 0) {
            const DiagItem it = diag[tk.idx];
            fwd_diag_body<true, T, NR, 4>(S, it, yrow_ptr, ybuf, ldy, x, ldx, z, ldz, nrhs, w1, ysh, red);
            sig = cb + it.c;
        } else if (type == 1) {
            const FwdItem it = fwd[tk.idx];
            fwd_prod_body<true, T, NR, 4>(S, it, L, z, ldz, nrhs, ydst, ybuf, ldy, ysh, red);
        } else if (type == 2) {
            const BwdItem it = bwd[tk.idx];
            bwd_prod_body<true, T, NR>(S, it, (tk.code >> 2) * 64, U, x, ldx, nrhs, nrhs, 0, part, tile, xr, red);
            sig = 2 * cb + it.c;
        } else {
            const DiagItem it = diag[tk.idx];
            bwd_diag_body<true, T, NR, 4>(S, it, bwd, z, ldz, x, ldx, nrhs, nrhs, 0, part, wb, ysh, red);
            sig = 3 * cb + it.c;
        }
        // the task's hand-over data went out with device-coherent stores:
        // once every wave has them complete, one thread signals
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            if (type == 1) {
                for (int k = tgt_ptr[tk.idx]; k < tgt_ptr[tk.idx + 1]; k++)
                    __hip_atomic_fetch_add(fwd_in + tgt[k], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_fetch_add(cnt + sig, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        t++;
    }
}

// ------------------------------- launchers ---------------------------------
template <typename T>
void launch_fwd_diag(const DevSymbol &S, const DiagItem *items, int n, const int64_t *yrow_ptr, const T *ybuf,
                     int64_t ldy, const T *b, int64_t ldb, T *z, int64_t ldz, int nrhs, const T *w1, int maxwidth,
                     hipStream_t st) {
    if (n <= 0) return;
    // several right-hand sides: pairs of columns on the grid's y dimension,
    // so that the few diagonal blocks of the top levels use more of the chip;
    // levels of narrow cblks (maxwidth = the level's widest) take 4-wave
    // workgroups
    const bool narrow = maxwidth <= SOLVE_NARROW;
    if (nrhs == 1) {
        if (maxwidth <= SOLVE_TINY)
            hipLaunchKernelGGL((k_fwd_diag<T, 1, 1>), dim3(grid8(n)), dim3(64), (size_t)maxwidth * sizeof(T), st, S,
                               items, n, yrow_ptr, ybuf, ldy, b, ldb, z, ldz, nrhs, w1);
        else if (narrow)
            hipLaunchKernelGGL((k_fwd_diag<T, 1, 4>), dim3(grid8(n)), dim3(256), (size_t)maxwidth * sizeof(T), st, S,
                               items, n, yrow_ptr, ybuf, ldy, b, ldb, z, ldz, nrhs, w1);
        else
            hipLaunchKernelGGL((k_fwd_diag<T, 1, 16>), dim3(grid8(n)), dim3(1024), (size_t)maxwidth * sizeof(T), st,
                               S, items, n, yrow_ptr, ybuf, ldy, b, ldb, z, ldz, nrhs, w1);
    } else {
        if (narrow)
            hipLaunchKernelGGL((k_fwd_diag<T, 2, 4>), dim3(grid8(n), (nrhs + 1) / 2), dim3(256),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, yrow_ptr, ybuf, ldy, b, ldb, z, ldz,
                               nrhs, w1);
        else
            hipLaunchKernelGGL((k_fwd_diag<T, 2, 16>), dim3(grid8(n), (nrhs + 1) / 2), dim3(1024),
                               (size_t)maxwidth * 2 * sizeof(T), st, S, items, n, yrow_ptr, ybuf, ldy, b, ldb, z, ldz,
                               nrhs, w1);
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
    const dim3 g(grid8(n), (maxwidth + 63) / 64, nrhs == 1 ? 1 : (nrhs + 7) / 8);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_bwd_prod<T, 1>), g, dim3(256), 0, st, S, items, n, coef, x, ldx, nrhs, part);
    else
        hipLaunchKernelGGL((k_bwd_prod<T, 8>), g, dim3(256), 0, st, S, items, n, coef, x, ldx, nrhs, part);
}
template <typename T>
void launch_bwd_diag(const DevSymbol &S, const DiagItem *items, int n, const BwdItem *bitems, const T *z,
                     int64_t ldz, T *x, int64_t ldx, int nrhs, const T *part, const T *wb, int maxwidth,
                     hipStream_t st) {
    if (n <= 0) return;
    const bool narrow = maxwidth <= SOLVE_NARROW;
    if (nrhs == 1) {
        if (maxwidth <= SOLVE_TINY)
            hipLaunchKernelGGL((k_bwd_diag<T, 1, 1>), dim3(grid8(n)), dim3(64), (size_t)maxwidth * sizeof(T), st, S,
                               items, n, bitems, z, ldz, x, ldx, nrhs, part, wb);
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

template <typename T>
void launch_solve_tasks(const DevSymbol &S, const SolveTask *tasks, const int32_t *chunks, int nchunks,
                        const DiagItem *diag, const FwdItem *fwd, const BwdItem *bwd, const int32_t *tgt_ptr,
                        const int32_t *tgt, const int32_t *need, int32_t *cnt, int cb, const int64_t *yrow_ptr,
                        T *ybuf, int64_t ldy, T *x, int64_t ldx, T *z, int64_t ldz, int nrhs, const T *L, const T *U,
                        const T *w1, const T *wb, const int32_t *ydst, T *part, int maxwidth, int nwg,
                        hipStream_t st) {
    if (nchunks <= 0) return;
    const int g = std::max(1, std::min(nwg, nchunks));
    // the right-hand sides of one launch times the widest cblk stay within
    // the level kernels' 24 KiB bound, so this stays below 33 KiB
    const size_t sh = (size_t)std::max(maxwidth * nrhs, 64 * 65) * sizeof(T);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_solve_tasks<T, 1>), dim3(g), dim3(256), sh, st, S, tasks, chunks, nchunks, diag, fwd, bwd, tgt_ptr,
                           tgt, need, cnt, cb, yrow_ptr, ybuf, ldy, x, ldx, z, ldz, nrhs, L, U, w1, wb, ydst, part);
    else
        hipLaunchKernelGGL((k_solve_tasks<T, 8>), dim3(g), dim3(256), sh, st, S, tasks, chunks, nchunks, diag, fwd, bwd, tgt_ptr,
                           tgt, need, cnt, cb, yrow_ptr, ybuf, ldy, x, ldx, z, ldz, nrhs, L, U, w1, wb, ydst, part);
}

#define INST(T)                                                                                                   \
    template void launch_fwd_diag<T>(const DevSymbol &, const DiagItem *, int, const int64_t *, const T *,       \
                                     int64_t, const T *, int64_t, T *, int64_t, int, const T *, int,             \
                                     hipStream_t);                                                               \
    template void launch_fwd_prod<T>(const DevSymbol &, const FwdItem *, int, const T *, const T *, int64_t, int, \
                                     const int32_t *, T *, int64_t, int, hipStream_t);                           \
    template void launch_bwd_prod<T>(const DevSymbol &, const BwdItem *, int, const T *, const T *, int64_t, int, \
                                     T *, int, hipStream_t);                                                     \
    template void launch_bwd_diag<T>(const DevSymbol &, const DiagItem *, int, const BwdItem *, const T *,        \
                                     int64_t, T *, int64_t, int, const T *, const T *, int, hipStream_t);
INST(double)
INST(float)

#define INST_TASKS(T)                                                                                             \
    template void launch_solve_tasks<T>(const DevSymbol &, const SolveTask *, const int32_t *, int,              \
                                        const DiagItem *,                                                        \
                                        const FwdItem *, const BwdItem *, const int32_t *, const int32_t *,      \
                                        const int32_t *, int32_t *, int, const int64_t *, T *, int64_t, T *,     \
                                        int64_t, T *, int64_t, int, const T *, const T *, const T *, const T *,  \
                                        const int32_t *, T *, int, int, hipStream_t);
INST_TASKS(double)
INST_TASKS(float)

}  // namespace pxa
