// Supernodal forward / backward solve kernels for CDNA4 (gfx950)
// (reference kernels/solve_ztrsmsp.c:67-557, sopalin/sequential_ztrsm.c:34-103).
//
// Forward sweep, level by level:
//   k_fwd_diag : for every cblk c of the level, x_c -= the update-buffer rows
//                that face c (a per-row CSR built at analysis, sources in
//                increasing order), then x_c = inv(L_cc) x_c;
//   k_fwd_prod : for the same cblks, ybuf[off-diagonal rows] = L_off * x_c on
//                64-row tiles, the column range split over the 4 waves.
// Backward sweep, levels from the root down:
//   k_bwd_prod : partial[j] = sum over a 64-row tile r of L[r, j] x[row(r)],
//                the tile staged through LDS so panel reads coalesce;
//   k_bwd_diag : x_c = inv(L_cc)^T (x_c - sum of the tiles' partials), with the
//                transposed inverse stored column-major.
// Every output element is produced in a fixed order by one workgroup: no
// atomics, bitwise-reproducible solves.  NR is the compile-time bound on the
// right-hand sides of one launch (1 or 8).
#include "dev_common.h"
#include "dev_types.h"
#include "kernels.h"

namespace pxa {

static inline int grid8(int n) { return ((n + 7) / 8) * 8; }

// out[i] = sum_k M[i, k] y[k] over the triangle of an n x n column-major M:
// k <= i (LOWER) or k >= i (upper); rows in blocks of 64, k split in 4.
template <typename T, int NR, bool LOWER>
__device__ void tri_apply(const T *__restrict__ M, int n, const T *y, int nrhs, T *red, T *out, int64_t ldo) {
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    for (int ib = 0; ib < n; ib += 64) {
        const int i = ib + r;
        T acc[NR];
#pragma unroll
        for (int u = 0; u < NR; u++) acc[u] = 0;
        if (i < n) {
            const int klo = LOWER ? 0 : i, khi = LOWER ? i + 1 : n;
            const int len = khi - klo, part = (len + 3) / 4;
            const int ka = klo + q * part, kz = min(khi, ka + part);
#pragma unroll 4
            for (int k = ka; k < kz; k++) {
                const T m = M[i + (int64_t)k * n];
#pragma unroll
                for (int u = 0; u < NR; u++)
                    if (u < nrhs) acc[u] += m * y[k + u * n];
            }
        }
#pragma unroll
        for (int u = 0; u < NR; u++) red[(q * 64 + r) * NR + u] = acc[u];
        __syncthreads();
        if (q == 0 && i < n) {
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nrhs)
                    out[i + u * ldo] = red[(0 * 64 + r) * NR + u] + red[(1 * 64 + r) * NR + u] +
                                       red[(2 * 64 + r) * NR + u] + red[(3 * 64 + r) * NR + u];
        }
        __syncthreads();
    }
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_fwd_diag(DevSymbol S, const int32_t *__restrict__ cblks,
                                                  const int64_t *__restrict__ yrow_ptr,
                                                  const int32_t *__restrict__ yrow_pos, const T *__restrict__ ybuf,
                                                  int64_t ldy, T *__restrict__ x, int64_t ldx, int nrhs,
                                                  const T *__restrict__ w1) {
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);  // maxwidth * NR
    __shared__ T red[256 * NR];
    const int tid = threadIdx.x;
    const int c = cblks[blockIdx.x];
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    for (int t = tid; t < Kc * nrhs; t += 256) {
        const int i = t % Kc, rhs = t / Kc;
        const int g = fc + i;
        const T *yb = ybuf + rhs * ldy;
        T v = x[g + rhs * ldx];
        int64_t p = yrow_ptr[g];
        const int64_t pe = yrow_ptr[g + 1];
        for (; p + 4 <= pe; p += 4) {
            const int a = yrow_pos[p], b = yrow_pos[p + 1], cc = yrow_pos[p + 2], d = yrow_pos[p + 3];
            v -= ((yb[a] + yb[b]) + (yb[cc] + yb[d]));
        }
        for (; p < pe; p++) v -= yb[yrow_pos[p]];
        y[t] = v;
    }
    __syncthreads();
    tri_apply<T, NR, true>(w1 + S.cblk_winv[c], Kc, y, nrhs, red, x + fc, ldx);
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_fwd_prod(DevSymbol S, const FwdItem *__restrict__ items, int nitems,
                                                  const T *__restrict__ coef, const T *__restrict__ x, int64_t ldx,
                                                  int nrhs, T *__restrict__ ybuf, int64_t ldy) {
    extern __shared__ char smem_raw[];
    T *xs = reinterpret_cast<T *>(smem_raw);  // Kc * NR
    __shared__ T red[256 * NR];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const FwdItem it = items[id];
    const int c = it.c;
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    const int64_t ld = S.cblk_stride[c];
    for (int t = tid; t < Kc * nrhs; t += 256) xs[t] = x[fc + (t % Kc) + (int64_t)(t / Kc) * ldx];
    __syncthreads();
    const int nr = it.r1 - it.r0;
    T acc[NR];
#pragma unroll
    for (int u = 0; u < NR; u++) acc[u] = 0;
    if (r < nr) {
        const T *row = coef + S.cblk_coef[c] + it.r0 + r;
        const int part = (Kc + 3) / 4;
        const int ka = q * part, kz = min(Kc, ka + part);
#pragma unroll 4
        for (int k = ka; k < kz; k++) {
            const T a = row[(int64_t)k * ld];
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nrhs) acc[u] += a * xs[k + u * Kc];
        }
    }
#pragma unroll
    for (int u = 0; u < NR; u++) red[(q * 64 + r) * NR + u] = acc[u];
    __syncthreads();
    if (q == 0 && r < nr) {
#pragma unroll
        for (int u = 0; u < NR; u++)
            if (u < nrhs)
                ybuf[it.slot + r + u * ldy] = red[(0 * 64 + r) * NR + u] + red[(1 * 64 + r) * NR + u] +
                                              red[(2 * 64 + r) * NR + u] + red[(3 * 64 + r) * NR + u];
    }
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
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    const BwdItem it = items[id];
    const int c = it.c;
    const int Kc = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    const int64_t rb = S.cblk_rowbase[c];
    const int nr = it.r1 - it.r0;
    for (int t = tid; t < 64 * NR; t += 256) {
        const int row = t & 63, u = t >> 6;
        xr[t] = (row < nr && u < nrhs) ? x[S.panel_grow[rb + it.r0 + row] + u * ldx] : (T)0;
    }
    const T *P = coef + S.cblk_coef[c] + it.r0;
    T *out = part + (int64_t)it.slot * nrhs;
    for (int k0 = 0; k0 < Kc; k0 += 64) {
        __syncthreads();
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
#pragma unroll
        for (int u = 0; u < NR; u++) red[(q * 64 + lane) * NR + u] = acc[u];
        __syncthreads();
        if (q == 0 && k0 + lane < Kc) {
#pragma unroll
            for (int u = 0; u < NR; u++)
                if (u < nrhs)
                    out[k0 + lane + u * Kc] = red[(0 * 64 + lane) * NR + u] + red[(1 * 64 + lane) * NR + u] +
                                              red[(2 * 64 + lane) * NR + u] + red[(3 * 64 + lane) * NR + u];
        }
    }
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_bwd_diag(DevSymbol S, const int32_t *__restrict__ cblks,
                                                  const int32_t *__restrict__ itemptr,
                                                  const BwdItem *__restrict__ items, T *__restrict__ x, int64_t ldx,
                                                  int nrhs, const T *__restrict__ part, const T *__restrict__ wb) {
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);
    __shared__ T red[256 * NR];
    const int tid = threadIdx.x;
    const int c = cblks[blockIdx.x];
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    const int i0 = itemptr[blockIdx.x], i1 = itemptr[blockIdx.x + 1];
    for (int t = tid; t < Kc * nrhs; t += 256) {
        const int i = t % Kc, rhs = t / Kc;
        T v = x[fc + i + rhs * ldx];
        for (int k = i0; k < i1; k++) v -= part[(int64_t)items[k].slot * nrhs + t];
        y[t] = v;
    }
    __syncthreads();
    // wb holds inv(L_cc)^T (or inv(U_cc)) column-major: upper triangular
    tri_apply<T, NR, false>(wb + S.cblk_winv[c], Kc, y, nrhs, red, x + fc, ldx);
}

// Transposed copy of the diagonal inverses of a level: WT[c] = W[c]^T.
template <typename T>
__global__ __launch_bounds__(256) void k_winv_transpose(DevSymbol S, const int32_t *__restrict__ cblks,
                                                        const T *__restrict__ w, T *__restrict__ wt) {
    __shared__ T tile[32][33];
    const int c = cblks[blockIdx.x];
    const int K = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const T *W = w + S.cblk_winv[c];
    T *WT = wt + S.cblk_winv[c];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int bi = 0; bi < K; bi += 32)
        for (int bj = 0; bj < K; bj += 32) {
            for (int r = ty; r < 32; r += 8) {
                const int i = bi + tx, j = bj + r;
                tile[r][tx] = (i < K && j < K) ? W[i + (int64_t)j * K] : (T)0;
            }
            __syncthreads();
            for (int r = ty; r < 32; r += 8) {
                const int i = bj + tx, j = bi + r;  // WT(i, j) = W(j, i)
                if (i < K && j < K) WT[i + (int64_t)j * K] = tile[tx][r];
            }
            __syncthreads();
        }
}

// ------------------------------- launchers ---------------------------------
template <typename T>
void launch_fwd_diag(const DevSymbol &S, const int32_t *cblks, int n, const int64_t *yrow_ptr,
                     const int32_t *yrow_pos, const T *ybuf, int64_t ldy, T *x, int64_t ldx, int nrhs, const T *w1,
                     int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_fwd_diag<T, 1>), dim3(n), dim3(256), sh, st, S, cblks, yrow_ptr, yrow_pos, ybuf, ldy, x,
                           ldx, nrhs, w1);
    else
        hipLaunchKernelGGL((k_fwd_diag<T, 8>), dim3(n), dim3(256), sh, st, S, cblks, yrow_ptr, yrow_pos, ybuf, ldy, x,
                           ldx, nrhs, w1);
}
template <typename T>
void launch_fwd_prod(const DevSymbol &S, const FwdItem *items, int n, const T *coef, const T *x, int64_t ldx,
                     int nrhs, T *ybuf, int64_t ldy, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_fwd_prod<T, 1>), dim3(grid8(n)), dim3(256), sh, st, S, items, n, coef, x, ldx, nrhs, ybuf,
                           ldy);
    else
        hipLaunchKernelGGL((k_fwd_prod<T, 8>), dim3(grid8(n)), dim3(256), sh, st, S, items, n, coef, x, ldx, nrhs, ybuf,
                           ldy);
}
template <typename T>
void launch_bwd_prod(const DevSymbol &S, const BwdItem *items, int n, const T *coef, const T *x, int64_t ldx,
                     int nrhs, T *part, hipStream_t st) {
    if (n <= 0) return;
    if (nrhs == 1)
        hipLaunchKernelGGL((k_bwd_prod<T, 1>), dim3(grid8(n)), dim3(256), 0, st, S, items, n, coef, x, ldx, nrhs, part);
    else
        hipLaunchKernelGGL((k_bwd_prod<T, 8>), dim3(grid8(n)), dim3(256), 0, st, S, items, n, coef, x, ldx, nrhs, part);
}
template <typename T>
void launch_bwd_diag(const DevSymbol &S, const int32_t *cblks, int n, const int32_t *itemptr, const BwdItem *items,
                     T *x, int64_t ldx, int nrhs, const T *part, const T *wb, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    if (nrhs == 1)
        hipLaunchKernelGGL((k_bwd_diag<T, 1>), dim3(n), dim3(256), sh, st, S, cblks, itemptr, items, x, ldx, nrhs, part,
                           wb);
    else
        hipLaunchKernelGGL((k_bwd_diag<T, 8>), dim3(n), dim3(256), sh, st, S, cblks, itemptr, items, x, ldx, nrhs, part,
                           wb);
}
template <typename T>
void launch_winv_transpose(const DevSymbol &S, const int32_t *cblks, int n, const T *w, T *wt, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_winv_transpose<T>, dim3(n), dim3(256), 0, st, S, cblks, w, wt);
}

#define INST(T)                                                                                                   \
    template void launch_fwd_diag<T>(const DevSymbol &, const int32_t *, int, const int64_t *, const int32_t *,  \
                                     const T *, int64_t, T *, int64_t, int, const T *, int, hipStream_t);        \
    template void launch_fwd_prod<T>(const DevSymbol &, const FwdItem *, int, const T *, const T *, int64_t, int, \
                                     T *, int64_t, int, hipStream_t);                                            \
    template void launch_bwd_prod<T>(const DevSymbol &, const BwdItem *, int, const T *, const T *, int64_t, int, \
                                     T *, hipStream_t);                                                          \
    template void launch_bwd_diag<T>(const DevSymbol &, const int32_t *, int, const int32_t *, const BwdItem *,   \
                                     T *, int64_t, int, const T *, const T *, int, hipStream_t);                 \
    template void launch_winv_transpose<T>(const DevSymbol &, const int32_t *, int, const T *, T *, hipStream_t);
INST(double)
INST(float)

}  // namespace pxa
