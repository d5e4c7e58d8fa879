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
#include "dev_common.h"
#include "dev_types.h"
#include "kernels.h"

namespace pxa {

static inline int grid8(int n) { return ((n + 7) / 8) * 8; }

// red holds NW x 64 x NR partial sums; the sum over waves is taken in wave
// order by the threads of wave 0.
template <typename T, int NR, int NW>
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
                out[u * ldo] = sub ? sub[u * lds] - v : v;
            }
        }
    }
}

template <typename T, int NR, int NW>
__global__ __launch_bounds__(NW * 64) void k_fwd_diag(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                      const int64_t *__restrict__ yrow_ptr,
                                                      const T *__restrict__ ybuf, int64_t ldy,
                                                      const T *__restrict__ b, int64_t ldb, T *__restrict__ z,
                                                      int64_t ldz, int nrhs, const T *__restrict__ w1) {
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);  // maxwidth * NR
    __shared__ T red[NW * 64 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const DiagItem it = items[id];
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
        for (; p + 4 <= pe; p += 4) v -= ((yb[p] + yb[p + 1]) + (yb[p + 2] + yb[p + 3]));
        for (; p < pe; p++) v -= yb[p];
        y[i + u * Kc] = v;
    }
    __syncthreads();
    const int i = r0 + r;
    T acc[NR];
#pragma unroll
    for (int u = 0; u < NR; u++) acc[u] = 0;
    if (i < Kc) {
        const T *M = w1 + S.cblk_winv[c] + i;
#pragma unroll 8
        for (int k = q; k < nrow; k += NW) {
            const T m = M[(int64_t)k * Kc];
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nrhs) acc[u] += m * y[k + u * Kc];
        }
    }
    wave_sum_store<T, NR, NW>(acc, red, r, q, i < Kc, nrhs, z + fc + i, ldz, (const T *)nullptr, 0);
}

template <typename T, int NR, int NW>
__global__ __launch_bounds__(NW * 64) void k_fwd_prod(DevSymbol S, const FwdItem *__restrict__ items, int nitems,
                                                      const T *__restrict__ coef, const T *__restrict__ z,
                                                      int64_t ldz, int nrhs, const int32_t *__restrict__ ydst,
                                                      T *__restrict__ ybuf, int64_t ldy) {
    extern __shared__ char smem_raw[];
    T *xs = reinterpret_cast<T *>(smem_raw);  // Kc * NR
    __shared__ T red[NW * 64 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const FwdItem it = items[id];
    const int c = it.c;
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    const int64_t ld = S.cblk_stride[c];
    for (int t = tid; t < Kc * nrhs; t += NW * 64) xs[t] = z[fc + (t % Kc) + (int64_t)(t / Kc) * ldz];
    __syncthreads();
    const int nr = it.r1 - it.r0;
    T acc[NR];
#pragma unroll
    for (int u = 0; u < NR; u++) acc[u] = 0;
    if (r < nr) {
        const T *row = coef + S.cblk_coef[c] + it.r0 + r;
#pragma unroll 4
        for (int k = q; k < Kc; k += NW) {
            const T a = row[(int64_t)k * ld];
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nrhs) acc[u] += a * xs[k + u * Kc];
        }
    }
    const int dst = (r < nr) ? ydst[it.slot + r] : 0;
    wave_sum_store<T, NR, NW>(acc, red, r, q, r < nr, nrhs, ybuf + dst, ldy, (const T *)nullptr, 0);
}

// One 64x64 tile (item, column block blockIdx.y) of the backward products.
template <typename T, int NR>
__global__ __launch_bounds__(256) void k_bwd_prod(DevSymbol S, const BwdItem *__restrict__ items, int nitems,
                                                  const T *__restrict__ coef, const T *__restrict__ x, int64_t ldx,
                                                  int nrhs, T *__restrict__ part) {
    __shared__ T tile[64][65];
    __shared__ T xr[64 * NR];
    __shared__ T red[256 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    const BwdItem it = items[id];
    const int c = it.c;
    const int Kc = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int k0 = blockIdx.y * 64;
    if (k0 >= Kc) return;
    const int64_t ld = S.cblk_stride[c];
    const int64_t rb = S.cblk_rowbase[c];
    const int nr = it.r1 - it.r0;
    for (int t = tid; t < 64 * NR; t += 256) {
        const int row = t & 63, u = t >> 6;
        xr[t] = (row < nr && u < nrhs) ? x[S.panel_grow[rb + it.r0 + row] + u * ldx] : (T)0;
    }
    const T *P = coef + S.cblk_coef[c] + it.r0;
    for (int cc = q; cc < 64; cc += 4)
        tile[lane][cc] = (lane < nr && k0 + cc < Kc) ? P[lane + (int64_t)(k0 + cc) * ld] : (T)0;
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
    T *out = part + (int64_t)it.slot * nrhs + k0 + lane;
    wave_sum_store<T, NR, 4>(acc, red, lane, q, k0 + lane < Kc, nrhs, out, Kc, (const T *)nullptr, 0);
}

template <typename T, int NR, int NW>
__global__ __launch_bounds__(NW * 64) void k_bwd_diag(DevSymbol S, const DiagItem *__restrict__ items, int nitems,
                                                      const BwdItem *__restrict__ bitems, const T *__restrict__ z,
                                                      int64_t ldz, T *__restrict__ x, int64_t ldx, int nrhs,
                                                      const T *__restrict__ part, const T *__restrict__ wb) {
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);  // maxwidth * NR
    __shared__ T red[NW * 64 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const DiagItem it = items[id];
    const int c = it.c, r0 = it.ib * 64;
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    // the tiles of one cblk own consecutive partial slots, Kc apart
    const int64_t s0 = (it.b1 > it.b0) ? (int64_t)bitems[it.b0].slot : 0;
    for (int kc = r0; kc < Kc; kc += 64) {
        const int k = kc + r;
        T acc[NR];
#pragma unroll
        for (int u = 0; u < NR; u++) acc[u] = 0;
        if (k < Kc) {
#pragma unroll 4
            for (int j = it.b0 + q; j < it.b1; j += NW) {
                const T *pp = part + (s0 + (int64_t)(j - it.b0) * Kc) * nrhs + k;
#pragma unroll
                for (int u = 0; u < NR; u++)
                    if (u < nrhs) acc[u] += pp[u * Kc];
            }
        }
        wave_sum_store<T, NR, NW>(acc, red, r, q, k < Kc, nrhs, y + k, Kc, z + fc + k, ldz);
        __syncthreads();
    }
    const int i = r0 + r;
    T acc[NR];
#pragma unroll
    for (int u = 0; u < NR; u++) acc[u] = 0;
    if (i < Kc) {
        // wb holds inv(L_cc)^T (or inv(U_cc)) column-major: upper triangular
        const T *M = wb + S.cblk_winv[c] + i;
#pragma unroll 8
        for (int k = r0 + q; k < Kc; k += NW) {
            const T m = M[(int64_t)k * Kc];
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nrhs) acc[u] += m * y[k + u * Kc];
        }
    }
    wave_sum_store<T, NR, NW>(acc, red, r, q, i < Kc, nrhs, x + fc + i, ldx, (const T *)nullptr, 0);
}

// ------------------------------- launchers ---------------------------------
template <typename T>
void launch_fwd_diag(const DevSymbol &S, const DiagItem *items, int n, const int64_t *yrow_ptr, const T *ybuf,
                     int64_t ldy, const T *b, int64_t ldb, T *z, int64_t ldz, int nrhs, const T *w1, int maxwidth,
                     hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_fwd_diag<T, 1, 16>), dim3(grid8(n)), dim3(1024), sh, st, S, items, n, yrow_ptr, ybuf,
                           ldy, b, ldb, z, ldz, nrhs, w1);
    else
        hipLaunchKernelGGL((k_fwd_diag<T, 8, 8>), dim3(grid8(n)), dim3(512), sh, st, S, items, n, yrow_ptr, ybuf, ldy,
                           b, ldb, z, ldz, nrhs, w1);
}
template <typename T>
void launch_fwd_prod(const DevSymbol &S, const FwdItem *items, int n, const T *coef, const T *z, int64_t ldz,
                     int nrhs, const int32_t *ydst, T *ybuf, int64_t ldy, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_fwd_prod<T, 1, 16>), dim3(grid8(n)), dim3(1024), sh, st, S, items, n, coef, z, ldz,
                           nrhs, ydst, ybuf, ldy);
    else
        hipLaunchKernelGGL((k_fwd_prod<T, 8, 8>), dim3(grid8(n)), dim3(512), sh, st, S, items, n, coef, z, ldz, nrhs,
                           ydst, ybuf, ldy);
}
template <typename T>
void launch_bwd_prod(const DevSymbol &S, const BwdItem *items, int n, const T *coef, const T *x, int64_t ldx,
                     int nrhs, T *part, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    const dim3 g(grid8(n), (maxwidth + 63) / 64);
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
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_bwd_diag<T, 1, 16>), dim3(grid8(n)), dim3(1024), sh, st, S, items, n, bitems, z, ldz, x,
                           ldx, nrhs, part, wb);
    else
        hipLaunchKernelGGL((k_bwd_diag<T, 8, 8>), dim3(grid8(n)), dim3(512), sh, st, S, items, n, bitems, z, ldz, x,
                           ldx, nrhs, part, wb);
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

}  // namespace pxa
