// Solve, SpMV and vector kernels for CDNA4 (gfx950).
//
// Forward sweep (reference kernels/solve_ztrsmsp.c:332-410): per level,
// k_fwd_gather sums the contributions L[rows facing c] * x[source] of the
// blocks in browtab(c) into per-item partial vectors, then k_fwd_diag
// subtracts them in item order and applies inv(L_cc).  Backward sweep
// (:459-557): k_bwd_gather computes L[off-diagonal rows of c]^T * x[rows] on
// row chunks, k_bwd_diag subtracts them and applies inv(L_cc)^T (inv(U_cc) for
// LU).  Every output is written by one workgroup with a fixed summation order.
#include "dev_common.h"
#include "dev_types.h"
#include "kernels.h"

namespace pxa {

static inline int grid8(int n) { return ((n + 7) / 8) * 8; }

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_gather(DevSymbol S, const FwdItem *__restrict__ items, int nitems,
                                                    const T *__restrict__ coef, const T *__restrict__ x, int64_t ldx,
                                                    int nrhs, T *__restrict__ part) {
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x;
    const FwdItem it = items[id];
    const int c = it.c;
    const int fc = S.cblk_fcol[c];
    const int Kc = S.cblk_fcol[c + 1] - fc;
    for (int e = tid; e < Kc * nrhs; e += 256) y[e] = (T)0;
    __syncthreads();
    for (int e = it.e0; e < it.e1; e++) {
        const int b = S.browtab[e];
        const int s = S.blok_lcblk[b];
        const int nr = S.blok_lrow[b] - S.blok_frow[b] + 1;
        const int off = S.blok_frow[b] - fc;
        const int fs = S.cblk_fcol[s];
        const int Ks = S.cblk_fcol[s + 1] - fs;
        const int64_t lds = S.cblk_stride[s];
        const T *A = coef + S.cblk_coef[s] + S.blok_roff[b];
        for (int t = tid; t < nr * nrhs; t += 256) {
            const int row = t % nr, rhs = t / nr;
            const T *xs = x + fs + rhs * ldx;
            T acc = 0;
            for (int k = 0; k < Ks; k++) acc += A[row + k * lds] * xs[k];
            y[off + row + rhs * Kc] += acc;
        }
        __syncthreads();
    }
    T *out = part + (int64_t)it.slot * nrhs;
    for (int e = tid; e < Kc * nrhs; e += 256) out[e] = y[e];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_diag(DevSymbol S, const int32_t *__restrict__ cblks,
                                                  const int32_t *__restrict__ itemptr,
                                                  const FwdItem *__restrict__ items, T *__restrict__ x, int64_t ldx,
                                                  int nrhs, const T *__restrict__ part, const T *__restrict__ w1) {
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);
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
    const T *W = w1 + S.cblk_winv[c];
    for (int t = tid; t < Kc * nrhs; t += 256) {
        const int i = t % Kc, rhs = t / Kc;
        const T *yy = y + rhs * Kc;
        T s = 0;
        for (int k = 0; k <= i; k++) s += W[i + (int64_t)k * Kc] * yy[k];
        x[fc + i + rhs * ldx] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_gather(DevSymbol S, const BwdItem *__restrict__ items, int nitems,
                                                    const T *__restrict__ coef, const T *__restrict__ x, int64_t ldx,
                                                    int nrhs, T *__restrict__ part) {
    extern __shared__ char smem_raw[];
    T *xr = reinterpret_cast<T *>(smem_raw);
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    if (id >= nitems) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const BwdItem it = items[id];
    const int c = it.c;
    const int Kc = S.cblk_fcol[c + 1] - S.cblk_fcol[c];
    const int64_t ld = S.cblk_stride[c];
    const int64_t rb = S.cblk_rowbase[c];
    const int nr = it.r1 - it.r0;
    const T *P = coef + S.cblk_coef[c] + it.r0;
    for (int t = tid; t < nr * nrhs; t += 256) {
        const int row = t % nr, rhs = t / nr;
        xr[t] = x[S.panel_grow[rb + it.r0 + row] + rhs * ldx];
    }
    __syncthreads();
    for (int j = wave; j < Kc; j += 4) {
        for (int rhs = 0; rhs < nrhs; rhs++) {
            T acc = 0;
            for (int t = lane; t < nr; t += 64) acc += P[t + (int64_t)j * ld] * xr[t + rhs * nr];
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (lane == 0) part[(int64_t)it.slot * nrhs + j + rhs * Kc] = acc;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_diag(DevSymbol S, const int32_t *__restrict__ cblks,
                                                  const int32_t *__restrict__ itemptr,
                                                  const BwdItem *__restrict__ items, T *__restrict__ x, int64_t ldx,
                                                  int nrhs, const T *__restrict__ part, const T *__restrict__ w2) {
    extern __shared__ char smem_raw[];
    T *y = reinterpret_cast<T *>(smem_raw);
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
    const T *W = w2 + S.cblk_winv[c];
    for (int t = tid; t < Kc * nrhs; t += 256) {
        const int i = t % Kc, rhs = t / Kc;
        const T *yy = y + rhs * Kc;
        const T *Wc = W + (int64_t)i * Kc;
        T s = 0;
        for (int k = i; k < Kc; k++) s += Wc[k] * yy[k];
        x[fc + i + rhs * ldx] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, const int64_t *__restrict__ colptr,
                                              const int32_t *__restrict__ rows, const T *__restrict__ vals, T alpha,
                                              const T *__restrict__ x, T beta, T *__restrict__ y) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    T s = 0;
    for (int64_t k = colptr[j]; k < colptr[j + 1]; k++) s += vals[k] * x[rows[k]];
    y[j] = (beta == (T)0) ? alpha * s : alpha * s + beta * y[j];
}

// --------------------------- vector kernels --------------------------------
static constexpr int RED_BLOCKS = 512;

__device__ __forceinline__ double block_sum(double v, double *sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0;
    if (tid == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(256) void k_dot_part(int64_t n, const double *__restrict__ x,
                                                  const double *__restrict__ y, double *__restrict__ partials) {
    __shared__ double sh[4];
    double v = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) v += x[i] * y[i];
    double t = block_sum(v, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_sum_final(int nparts, int m, const double *__restrict__ partials,
                                                   double *__restrict__ out) {
    // out[j] = sum_b partials[b*m + j], fixed order
    const int j = threadIdx.x;
    if (j >= m) return;
    double t = 0;
    for (int b = 0; b < nparts; b++) t += partials[(int64_t)b * m + j];
    out[j] = t;
}

__global__ __launch_bounds__(256) void k_axpy(int64_t n, double alpha, const double *__restrict__ x,
                                              double *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] += alpha * x[i];
}
__global__ __launch_bounds__(256) void k_axpy_dev(int64_t n, const double *__restrict__ alpha, double sign,
                                                  const double *__restrict__ x, double *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] += sign * alpha[0] * x[i];
}
__global__ __launch_bounds__(256) void k_scal(int64_t n, double alpha, double *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] *= alpha;
}
__global__ __launch_bounds__(256) void k_copy(int64_t n, const double *__restrict__ x, double *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[i];
}
__global__ __launch_bounds__(256) void k_d2s(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}
__global__ __launch_bounds__(256) void k_s2d(int64_t n, const float *__restrict__ x, double *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = (double)x[i];
}

// partials[b*m + j] = sum over this block's rows of V[i, j] * x[i]
__global__ __launch_bounds__(256) void k_gemv_t_part(int64_t n, int m, const double *__restrict__ V, int64_t ldv,
                                                     const double *__restrict__ x, double *__restrict__ partials) {
    __shared__ double sh[4];
    for (int j = 0; j < m; j++) {
        double v = 0;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
            v += V[i + j * ldv] * x[i];
        double t = block_sum(v, sh);
        if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * m + j] = t;
    }
}
__global__ __launch_bounds__(256) void k_gemv_n(int64_t n, int m, const double *__restrict__ V, int64_t ldv,
                                                const double *__restrict__ y, double alpha, double *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int j = 0; j < m; j++) s += V[i + j * ldv] * y[j];
    x[i] += alpha * s;
}

// Normwise backward error pieces: out = [max|b - A x|, max|x|, max|b|] for a
// matrix stored by rows (colptr/rows of the transpose = CSR of A).
__global__ __launch_bounds__(256) void k_berr_part(int64_t n, const int64_t *__restrict__ colptr,
                                                   const int32_t *__restrict__ rows, const double *__restrict__ vals,
                                                   const double *__restrict__ x, const double *__restrict__ b,
                                                   double *__restrict__ partials) {
    __shared__ double sh[3][4];
    double r = 0, xm = 0, bm = 0;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
        double s = 0;
        for (int64_t k = colptr[j]; k < colptr[j + 1]; k++) s += vals[k] * x[rows[k]];
        r = fmax(r, fabs(b[j] - s));
        xm = fmax(xm, fabs(x[j]));
        bm = fmax(bm, fabs(b[j]));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        r = fmax(r, __shfl_xor(r, o, 64));
        xm = fmax(xm, __shfl_xor(xm, o, 64));
        bm = fmax(bm, __shfl_xor(bm, o, 64));
    }
    if (lane == 0) {
        sh[0][wave] = r;
        sh[1][wave] = xm;
        sh[2][wave] = bm;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0;
        for (int w = 0; w < 4; w++) t = fmax(t, sh[threadIdx.x][w]);
        partials[(int64_t)blockIdx.x * 3 + threadIdx.x] = t;
    }
}
__global__ void k_max_final(int nparts, int m, const double *__restrict__ partials, double *__restrict__ out) {
    const int j = threadIdx.x;
    if (j >= m) return;
    double t = 0;
    for (int b = 0; b < nparts; b++) t = fmax(t, partials[(int64_t)b * m + j]);
    out[j] = t;
}

template <typename T>
__global__ __launch_bounds__(256) void k_permute(int64_t m, int nrhs, const int32_t *__restrict__ perm, int scatter,
                                                 const T *__restrict__ x, int64_t ldx, T *__restrict__ y,
                                                 int64_t ldy) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int64_t p = perm[i];
    for (int r = 0; r < nrhs; r++) {
        if (scatter) y[p + r * ldy] = x[i + r * ldx];
        else y[i + r * ldy] = x[p + r * ldx];
    }
}

// ------------------------------- launchers ---------------------------------
static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

template <typename T>
void launch_fwd_gather(const DevSymbol &S, const FwdItem *items, int n, const T *coef, const T *x, int64_t ldx,
                       int nrhs, T *part, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    hipLaunchKernelGGL(k_fwd_gather<T>, dim3(grid8(n)), dim3(256), sh, st, S, items, n, coef, x, ldx, nrhs, part);
}
template <typename T>
void launch_fwd_diag(const DevSymbol &S, const int32_t *cblks, int n, const int32_t *itemptr, const FwdItem *items,
                     T *x, int64_t ldx, int nrhs, const T *part, const T *w1, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    hipLaunchKernelGGL(k_fwd_diag<T>, dim3(n), dim3(256), sh, st, S, cblks, itemptr, items, x, ldx, nrhs, part, w1);
}
template <typename T>
void launch_bwd_gather(const DevSymbol &S, const BwdItem *items, int n, const T *coef, const T *x, int64_t ldx,
                       int nrhs, T *part, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)256 * nrhs * sizeof(T);
    hipLaunchKernelGGL(k_bwd_gather<T>, dim3(grid8(n)), dim3(256), sh, st, S, items, n, coef, x, ldx, nrhs, part);
}
template <typename T>
void launch_bwd_diag(const DevSymbol &S, const int32_t *cblks, int n, const int32_t *itemptr, const BwdItem *items,
                     T *x, int64_t ldx, int nrhs, const T *part, const T *w2, int maxwidth, hipStream_t st) {
    if (n <= 0) return;
    size_t sh = (size_t)maxwidth * nrhs * sizeof(T);
    hipLaunchKernelGGL(k_bwd_diag<T>, dim3(n), dim3(256), sh, st, S, cblks, itemptr, items, x, ldx, nrhs, part, w2);
}
template <typename T>
void launch_spmv(int64_t n, const int64_t *colptr, const int32_t *rows, const T *vals, T alpha, const T *x, T beta,
                 T *y, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spmv<T>, dim3(nblk(n)), dim3(256), 0, st, n, colptr, rows, vals, alpha, x, beta, y);
}
template <typename T>
void launch_permute(int64_t m, int nrhs, const int32_t *perm, bool scatter, const T *x, int64_t ldx, T *y,
                    int64_t ldy, hipStream_t st) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_permute<T>, dim3(nblk(m)), dim3(256), 0, st, m, nrhs, perm, scatter ? 1 : 0, x, ldx, y, ldy);
}

void launch_dot(int64_t n, const double *x, const double *y, double *partials, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(256), 0, st, n, x, y, partials);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, st, RED_BLOCKS, 1, partials, out);
}
void launch_nrm2sq(int64_t n, const double *x, double *partials, double *out, hipStream_t st) {
    launch_dot(n, x, x, partials, out, st);
}
void launch_axpy(int64_t n, double alpha, const double *x, double *y, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_axpy, dim3(nblk(n)), dim3(256), 0, st, n, alpha, x, y);
}
void launch_axpy_dev(int64_t n, const double *alpha, double sign, const double *x, double *y, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_axpy_dev, dim3(nblk(n)), dim3(256), 0, st, n, alpha, sign, x, y);
}
void launch_scal(int64_t n, double alpha, double *x, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_scal, dim3(nblk(n)), dim3(256), 0, st, n, alpha, x);
}
void launch_copy(int64_t n, const double *x, double *y, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(256), 0, st, n, x, y);
}
void launch_d2s(int64_t n, const double *x, float *y, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_d2s, dim3(nblk(n)), dim3(256), 0, st, n, x, y);
}
void launch_s2d(int64_t n, const float *x, double *y, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_s2d, dim3(nblk(n)), dim3(256), 0, st, n, x, y);
}
void launch_gemv_t(int64_t n, int m, const double *V, int64_t ldv, const double *x, double *partials, double *out,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_gemv_t_part, dim3(RED_BLOCKS), dim3(256), 0, st, n, m, V, ldv, x, partials);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, RED_BLOCKS, m, partials, out);
}
void launch_gemv_n(int64_t n, int m, const double *V, int64_t ldv, const double *y, double alpha, double *x,
                   hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_gemv_n, dim3(nblk(n)), dim3(256), 0, st, n, m, V, ldv, y, alpha, x);
}
void launch_berr_parts(int64_t n, const int64_t *colptr, const int32_t *rows, const double *vals, const double *x,
                       const double *b, double *partials, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_berr_part, dim3(RED_BLOCKS), dim3(256), 0, st, n, colptr, rows, vals, x, b, partials);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(64), 0, st, RED_BLOCKS, 3, partials, out);
}

#define INST(T)                                                                                               \
    template void launch_fwd_gather<T>(const DevSymbol &, const FwdItem *, int, const T *, const T *, int64_t, \
                                       int, T *, int, hipStream_t);                                           \
    template void launch_fwd_diag<T>(const DevSymbol &, const int32_t *, int, const int32_t *, const FwdItem *, \
                                     T *, int64_t, int, const T *, const T *, int, hipStream_t);              \
    template void launch_bwd_gather<T>(const DevSymbol &, const BwdItem *, int, const T *, const T *, int64_t, \
                                       int, T *, hipStream_t);                                                \
    template void launch_bwd_diag<T>(const DevSymbol &, const int32_t *, int, const int32_t *, const BwdItem *, \
                                     T *, int64_t, int, const T *, const T *, int, hipStream_t);              \
    template void launch_spmv<T>(int64_t, const int64_t *, const int32_t *, const T *, T, const T *, T, T *,   \
                                 hipStream_t);                                                                \
    template void launch_permute<T>(int64_t, int, const int32_t *, bool, const T *, int64_t, T *, int64_t,     \
                                    hipStream_t);
INST(double)
INST(float)

}  // namespace pxa
