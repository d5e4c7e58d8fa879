// SpMV, vector (BLAS-1) and permutation kernels for CDNA4 (gfx950).
// Reductions use a fixed grid and a fixed combination order, so results are
// bitwise reproducible from run to run.
#include "dev_common.h"
#include "dev_types.h"
#include <mutex>

#include "kernels.h"

namespace pxa {


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

// out[j] = sum_b partials[b*m + j] by one wave per j (block j): lane l adds
// b = l, l+64, ..., then a fixed butterfly over the lanes.
__global__ __launch_bounds__(64) void k_sum_final(int nparts, int m, const double *__restrict__ partials,
                                                  double *__restrict__ out) {
    const int j = blockIdx.x, lane = threadIdx.x;
    double t = 0;
    for (int b = lane; b < nparts; b += 64) t += partials[(int64_t)b * m + j];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) out[j] = t;
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
__global__ __launch_bounds__(256) void k_zero_int(int n, int *__restrict__ p) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
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

// ---- reference-style BLAS-1 / BLAS-2 (bcsc/bvec_zcompute.c) -------------------
// Euclidean norm by scaled sums of squares (frobenius_update, the LAPACK
// dlassq rule used by bvec_znrm2_seq :67-120): a pair (scale, ssq) stands for
// scale^2 * ssq; pairs merge by rescaling to the larger scale.  Each thread
// walks its strided entries in order, lanes and waves merge in a fixed
// butterfly and index order, then the block pairs in block order.
__device__ __forceinline__ void ssq_add(double &scale, double &ssq, double v) {
    const double a = fabs(v);
    if (a == 0.0) return;
    if (scale < a) {
        const double r = scale / a;
        ssq = 1.0 + ssq * r * r;
        scale = a;
    } else {
        const double r = a / scale;
        ssq += r * r;
    }
}
__device__ __forceinline__ void ssq_merge(double &scale, double &ssq, double s2, double q2) {
    if (s2 == 0.0) return;
    if (scale < s2) {
        const double r = scale / s2;
        ssq = q2 + ssq * r * r;
        scale = s2;
    } else {
        const double r = s2 / scale;
        ssq += q2 * r * r;
    }
}
__global__ __launch_bounds__(256) void k_nrm2_part(int64_t n, const double *__restrict__ x,
                                                   double *__restrict__ partials) {
    __shared__ double sh[2][4];
    double scale = 0.0, ssq = 1.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        ssq_add(scale, ssq, x[i]);
    for (int o = 32; o > 0; o >>= 1) {
        const double s2 = __shfl_xor(scale, o, 64), q2 = __shfl_xor(ssq, o, 64);
        ssq_merge(scale, ssq, s2, q2);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wave] = scale;
        sh[1][wave] = ssq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, q = 1.0;
        for (int w = 0; w < 4; w++) ssq_merge(s, q, sh[0][w], sh[1][w]);
        partials[2 * blockIdx.x] = s;
        partials[2 * blockIdx.x + 1] = q;
    }
}
__global__ __launch_bounds__(64) void k_nrm2_final(int nparts, const double *__restrict__ partials,
                                                   double *__restrict__ out) {
    if (threadIdx.x != 0) return;
    double s = 0.0, q = 1.0;
    for (int b = 0; b < nparts; b++) ssq_merge(s, q, partials[2 * b], partials[2 * b + 1]);
    out[0] = s * sqrt(q);
}
// y = alpha A x + beta y, A m x n column-major (bvec_zgemv_seq :1168-1210);
// y is not read when beta = 0
__global__ __launch_bounds__(256) void k_gemv(int64_t m, int64_t n, double alpha, const double *__restrict__ A,
                                              int64_t lda, const double *__restrict__ x, double beta,
                                              double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double s = 0;
    for (int64_t j = 0; j < n; j++) s += A[i + j * lda] * x[j];
    y[i] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[i];
}
// y = alpha A^T x + beta y through the transposed pattern of the CSR: row j of
// A^T holds the entries vals[tpos[k]], k in [trowptr[j], trowptr[j+1]), of
// column j of A, at rows tcols[k] (increasing)
__global__ __launch_bounds__(256) void k_spmv_t(int64_t n, const int64_t *__restrict__ trowptr,
                                                const int32_t *__restrict__ tcols, const int64_t *__restrict__ tpos,
                                                const double *__restrict__ vals, double alpha,
                                                const double *__restrict__ x, double beta, double *__restrict__ y) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double s = 0;
    for (int64_t k = trowptr[j]; k < trowptr[j + 1]; k++) s += vals[tpos[k]] * x[tcols[k]];
    y[j] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[j];
}

// Rows [r0, r1) of the Schur complement in the solve (reference
// kernels/solve_ztrsmsp.c:383-400,513-546): 0 z = x, 1 z = x minus the
// forward products pending for the row (update-buffer entries, in source
// order), 2 x = 0, 3 x = z.
template <typename T>
__global__ __launch_bounds__(256) void k_schur_rows(int64_t r0, int64_t r1, int nrhs, int mode,
                                                    const YSlices ys, const T *__restrict__ ybuf,
                                                    int64_t ldy, T *__restrict__ x, int64_t ldx, T *__restrict__ z,
                                                    int64_t ldz) {
    const int64_t i = r0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= r1) return;
    for (int u = 0; u < nrhs; u++) {
        if (mode == 0) {
            z[i + u * ldz] = x[i + u * ldx];
        } else if (mode == 1) {
            T v = x[i + u * ldx];
            const T *yb = ybuf + u * ldy + ys.base[i >> 6] + (i & 63);
            for (int j = 0; j < ys.cnt[i]; j++) v -= yb[(int64_t)j * 64];
            z[i + u * ldz] = v;
        } else if (mode == 2) {
            x[i + u * ldx] = (T)0;
        } else {
            x[i + u * ldx] = z[i + u * ldz];
        }
    }
}

// ------------------------------- launchers ---------------------------------
static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

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
// ---- segmented vector kernels for batches of independent systems ----------
// Segment s owns rows [off[s], off[s+1]); column c of an n x R block is at
// x + c*ld.  out[s + c*nseg] = sum over the segment of x*y, one workgroup per
// (segment, column), fixed-order tree reduction.
// Segmented dot products in two passes of fixed order: pass 1 reduces
// chunks of SEG_CHUNK entries (chunk k lies in segment cseg[k]; the chunks of
// segment s are [coff[s], coff[s+1])), pass 2 adds the chunk sums of each
// segment in chunk order.
static constexpr int SEG_CHUNK = 4096;
__global__ __launch_bounds__(256) void k_seg_dot_part(const int64_t *__restrict__ off,
                                                      const int32_t *__restrict__ coff,
                                                      const int32_t *__restrict__ cseg, int nchunk, int64_t ld,
                                                      const double *__restrict__ x, const double *__restrict__ y,
                                                      double *__restrict__ part) {
    __shared__ double red[256];
    const int k = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
    const int s = cseg[k];
    const int64_t i0 = off[s] + (int64_t)(k - coff[s]) * SEG_CHUNK;
    const int64_t i1 = min(i0 + SEG_CHUNK, off[s + 1]);
    const double *xc = x + c * ld, *yc = y + c * ld;
    double v = 0;
    for (int64_t i = i0 + tid; i < i1; i += 256) v += xc[i] * yc[i];
    red[tid] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) part[k + (int64_t)c * nchunk] = red[0];
}

__global__ __launch_bounds__(64) void k_seg_dot_final(const int32_t *__restrict__ coff, int nseg, int nchunk,
                                                      const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double red[64];
    const int s = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
    const double *pc = part + (int64_t)c * nchunk;
    double v = 0;
    for (int k = coff[s] + tid; k < coff[s + 1]; k += 64) v += pc[k];
    red[tid] = v;
    __syncthreads();
    for (int o = 32; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) out[s + (int64_t)c * nseg] = red[0];
}

// y[:, c] += sign * a[seg(i) + c*nseg] * x[:, c]
__global__ __launch_bounds__(256) void k_seg_axpy(int64_t n, int R, int64_t ld, const int32_t *__restrict__ seg,
                                                  int nseg, const double *__restrict__ a, double sign,
                                                  const double *__restrict__ x, double *__restrict__ y) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * R) return;
    const int64_t i = t % n, c = t / n;
    y[i + c * ld] += sign * a[seg[i] + c * nseg] * x[i + c * ld];
}

// p[:, c] = z[:, c] + b[seg(i) + c*nseg] * p[:, c]
__global__ __launch_bounds__(256) void k_seg_xpby(int64_t n, int R, int64_t ld, const int32_t *__restrict__ seg,
                                                  int nseg, const double *__restrict__ b,
                                                  const double *__restrict__ z, double *__restrict__ p) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * R) return;
    const int64_t i = t % n, c = t / n;
    p[i + c * ld] = z[i + c * ld] + b[seg[i] + c * nseg] * p[i + c * ld];
}

int seg_dot_chunk() { return SEG_CHUNK; }
void launch_seg_dot(const int64_t *off, const int32_t *coff, const int32_t *cseg, int nseg, int nchunk, int R,
                    int64_t ld, const double *x, const double *y, double *part, double *out, hipStream_t st) {
    if (nseg <= 0 || nchunk <= 0 || R <= 0) return;
    hipLaunchKernelGGL(k_seg_dot_part, dim3(nchunk, R), dim3(256), 0, st, off, coff, cseg, nchunk, ld, x, y, part);
    hipLaunchKernelGGL(k_seg_dot_final, dim3(nseg, R), dim3(64), 0, st, coff, nseg, nchunk, part, out);
}
void launch_seg_axpy(int64_t n, int R, int64_t ld, const int32_t *seg, int nseg, const double *a, double sign,
                     const double *x, double *y, hipStream_t st) {
    const int64_t m = n * R;
    if (m > 0)
        hipLaunchKernelGGL(k_seg_axpy, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, n, R, ld, seg, nseg, a,
                           sign, x, y);
}
void launch_seg_xpby(int64_t n, int R, int64_t ld, const int32_t *seg, int nseg, const double *b, const double *z,
                     double *p, hipStream_t st) {
    const int64_t m = n * R;
    if (m > 0)
        hipLaunchKernelGGL(k_seg_xpby, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, n, R, ld, seg, nseg, b,
                           z, p);
}

template <typename T>
__global__ __launch_bounds__(256) void k_add_diag(int64_t n, const int64_t *__restrict__ pos, double delta,
                                                  T *__restrict__ coef) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) coef[pos[i]] += (T)delta;
}
__global__ __launch_bounds__(256) void k_set_csr_diag(int64_t n, const int64_t *__restrict__ pos,
                                                      const double *__restrict__ d0, double delta,
                                                      double *__restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && pos[i] >= 0) vals[pos[i]] = d0[i] + delta;
}
template <typename T>
void launch_add_diag(int64_t n, const int64_t *pos, double delta, T *coef, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_add_diag<T>, dim3(nblk(n)), dim3(256), 0, st, n, pos, delta, coef);
}
template void launch_add_diag<double>(int64_t, const int64_t *, double, double *, hipStream_t);
template void launch_add_diag<float>(int64_t, const int64_t *, double, float *, hipStream_t);
void launch_set_csr_diag(int64_t n, const int64_t *pos, const double *d0, double delta, double *vals, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_set_csr_diag, dim3(nblk(n)), dim3(256), 0, st, n, pos, d0, delta, vals);
}

// out[i] = coef[pos[i]] (the factor's diagonal, widened to FP64)
template <typename T>
__global__ __launch_bounds__(256) void k_gather_diag(int64_t n, const int64_t *__restrict__ pos,
                                                     const T *__restrict__ coef, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (double)coef[pos[i]];
}
template <typename T> void launch_gather_diag(int64_t n, const int64_t *pos, const T *coef, double *out, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_gather_diag<T>, dim3(nblk(n)), dim3(256), 0, st, n, pos, coef, out);
}
template void launch_gather_diag<double>(int64_t, const int64_t *, const double *, double *, hipStream_t);
template void launch_gather_diag<float>(int64_t, const int64_t *, const float *, double *, hipStream_t);

// Coefficient-table fill (reference kernels/cpucblk_zinit.c:229-309) from
// the FP64 CSR values of the permuted matrix: entry e goes to L[cmap[e]]
// when cmap[e] >= 0, to U[-2 - cmap[e]] when cmap[e] <= -2 (U given), and
// nowhere at -1 (outside the symbol).  The tables are zero beforehand; every
// position receives at most one entry.
template <typename T>
__global__ __launch_bounds__(256) void k_fill_coef(int64_t nnz, const int64_t *__restrict__ cmap,
                                                   const double *__restrict__ vals, T *__restrict__ L,
                                                   T *__restrict__ U) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nnz; e += stride) {
        const int64_t p = cmap[e];
        if (p >= 0) L[p] = (T)vals[e];
        else if (p <= -2 && U) U[-2 - p] = (T)vals[e];
    }
}
template <typename T>
void launch_fill_coef(int64_t nnz, const int64_t *cmap, const double *vals, T *L, T *U, hipStream_t st) {
    if (nnz > 0)
        hipLaunchKernelGGL(k_fill_coef<T>, dim3(nnz > 16384 * 256 ? 16384u : nblk(nnz)), dim3(256), 0, st, nnz,
                           cmap, vals, L, U);
}
template void launch_fill_coef<double>(int64_t, const int64_t *, const double *, double *, double *, hipStream_t);
template void launch_fill_coef<float>(int64_t, const int64_t *, const double *, float *, float *, hipStream_t);

// columns [0, nr) of an n-row block: y = x (different leading dimensions)
template <typename T>
__global__ __launch_bounds__(256) void k_copy2d(int64_t n, const T *__restrict__ x, int64_t ldx,
                                                T *__restrict__ y, int64_t ldy) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t u = blockIdx.y;
    if (i < n) y[i + u * ldy] = x[i + u * ldx];
}
template <typename T> void launch_copy2d(int64_t n, int nr, const T *x, int64_t ldx, T *y, int64_t ldy, hipStream_t st) {
    if (n > 0 && nr > 0) hipLaunchKernelGGL(k_copy2d<T>, dim3(nblk(n), nr), dim3(256), 0, st, n, x, ldx, y, ldy);
}
template void launch_copy2d<double>(int64_t, int, const double *, int64_t, double *, int64_t, hipStream_t);
template void launch_copy2d<float>(int64_t, int, const float *, int64_t, float *, int64_t, hipStream_t);

// x_i *= d_i (divide = 0) or x_i /= d_i (divide = 1), d_i = coef[pos[i]]: the
// factor's diagonal (reference solve_cblk_zdiag, kernels/solve_ztrsmsp.c:584-633)
template <typename T>
__global__ __launch_bounds__(256) void k_diag_scale(int64_t n, const int64_t *__restrict__ pos,
                                                    const T *__restrict__ coef, int divide, T *__restrict__ x,
                                                    int64_t ldx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t u = blockIdx.y;
    if (i < n) {
        const T d = coef[pos[i]];
        x[i + u * ldx] = divide ? x[i + u * ldx] / d : x[i + u * ldx] * d;
    }
}
template <typename T>
void launch_diag_scale(int64_t n, int nr, const int64_t *pos, const T *coef, bool divide, T *x, int64_t ldx,
                       hipStream_t st) {
    if (n > 0 && nr > 0)
        hipLaunchKernelGGL(k_diag_scale<T>, dim3(nblk(n), nr), dim3(256), 0, st, n, pos, coef, divide ? 1 : 0, x, ldx);
}
template void launch_diag_scale<double>(int64_t, int, const int64_t *, const double *, bool, double *, int64_t,
                                        hipStream_t);
template void launch_diag_scale<float>(int64_t, int, const int64_t *, const float *, bool, float *, int64_t,
                                       hipStream_t);

void launch_zero_int(int n, int *p, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, n, p);
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
    hipLaunchKernelGGL(k_sum_final, dim3(m), dim3(64), 0, st, RED_BLOCKS, m, partials, out);
}
void launch_gemv_n(int64_t n, int m, const double *V, int64_t ldv, const double *y, double alpha, double *x,
                   hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_gemv_n, dim3(nblk(n)), dim3(256), 0, st, n, m, V, ldv, y, alpha, x);
}
template <typename T>
void launch_schur_rows(int64_t r0, int64_t r1, int nrhs, int mode, const YSlices &ys, const T *ybuf, int64_t ldy,
                       T *x, int64_t ldx, T *z, int64_t ldz, hipStream_t st) {
    if (r1 > r0)
        hipLaunchKernelGGL(k_schur_rows<T>, dim3(nblk(r1 - r0)), dim3(256), 0, st, r0, r1, nrhs, mode, ys, ybuf,
                           ldy, x, ldx, z, ldz);
}
template void launch_schur_rows<double>(int64_t, int64_t, int, int, const YSlices &, const double *, int64_t, double *,
                                        int64_t, double *, int64_t, hipStream_t);
template void launch_schur_rows<float>(int64_t, int64_t, int, int, const YSlices &, const float *, int64_t, float *,
                                       int64_t, float *, int64_t, hipStream_t);
void launch_nrm2(int64_t n, const double *x, double *partials, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_nrm2_part, dim3(RED_BLOCKS), dim3(256), 0, st, n, x, partials);
    hipLaunchKernelGGL(k_nrm2_final, dim3(1), dim3(64), 0, st, RED_BLOCKS, partials, out);
}
void launch_gemv(int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x, double beta,
                 double *y, hipStream_t st) {
    if (m > 0) hipLaunchKernelGGL(k_gemv, dim3(nblk(m)), dim3(256), 0, st, m, n, alpha, A, lda, x, beta, y);
}
void launch_spmv_t(int64_t n, const int64_t *trowptr, const int32_t *tcols, const int64_t *tpos, const double *vals,
                   double alpha, const double *x, double beta, double *y, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(k_spmv_t, dim3(nblk(n)), dim3(256), 0, st, n, trowptr, tcols, tpos, vals, alpha, x, beta,
                           y);
}
void launch_berr_parts(int64_t n, const int64_t *colptr, const int32_t *rows, const double *vals, const double *x,
                       const double *b, double *partials, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_berr_part, dim3(RED_BLOCKS), dim3(256), 0, st, n, colptr, rows, vals, x, b, partials);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(64), 0, st, RED_BLOCKS, 3, partials, out);
}

// ---- backward error without allocation --------------------------------------
// Library-static scratch of pastix_amd_backward_error_device_noalloc: the
// per-block partial maxima (|b - A x|, |x|, |b|) of up to NOALLOC_MAX_RHS
// right-hand sides.  Every GPU holds its own copy for the life of the
// process.  Host code never passes the array's name to a kernel (on the host
// it is only a stand-in for the symbol): it uses the address that
// hipGetSymbolAddress returned for the device, resolved once per device by
// noalloc_scratch_init.  Every block writes its partials on every call, so
// nothing needs clearing between calls; calls that overlap on one GPU race.
static_assert(NOALLOC_SCRATCH_DOUBLES == (size_t)RED_BLOCKS * 3 * NOALLOC_MAX_RHS, "one partial triple per block and column");
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
static std::mutex scratch_mu;
static double *scratch_dev[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> g(scratch_mu);
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return -1;
    if (scratch_dev[device]) return 0;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    scratch_dev[device] = static_cast<double *>(p);
    return 0;
}

double *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> g(scratch_mu);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? scratch_dev[device] : nullptr;
}

// k_berr_part on right-hand sides and solutions in the user's ordering: row
// j of the permuted matrix is user row peri[j], and its column m reads x at
// peri[m], so the products and their order are those of k_berr_part on the
// permuted vectors.  Column c of the block (blockIdx.y) writes its partials
// at partials[(c * gridDim.x + blockIdx.x) * 3 + t].
__global__ __launch_bounds__(256) void k_berr_part_user(int64_t n, const int64_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ cols,
                                                        const double *__restrict__ vals,
                                                        const int32_t *__restrict__ peri, const double *__restrict__ x,
                                                        int64_t ldx, const double *__restrict__ b, int64_t ldb,
                                                        double *__restrict__ partials) {
    __shared__ double sh[3][4];
    const int c = blockIdx.y;
    x += (int64_t)c * ldx;
    b += (int64_t)c * ldb;
    double r = 0, xm = 0, bm = 0;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
        double s = 0;
        for (int64_t k = rowptr[j]; k < rowptr[j + 1]; k++) s += vals[k] * x[peri[cols[k]]];
        const int64_t u = peri[j];
        r = fmax(r, fabs(b[u] - s));
        xm = fmax(xm, fabs(x[u]));
        bm = fmax(bm, fabs(b[u]));
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
        partials[((int64_t)c * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = t;
    }
}

// berr[c] = max|b - A x| / (normInf max|x| + max|b|) of column c (thread c),
// rounded step by step as the host computes it (no contraction)
__global__ void k_berr_final_user(int nparts, int nrhs, const double *__restrict__ partials, double normInf,
                                  double *__restrict__ berr) {
    const int c = threadIdx.x;
    if (c >= nrhs) return;
    double t[3] = {0, 0, 0};
    for (int b = 0; b < nparts; b++)
        for (int i = 0; i < 3; i++) t[i] = fmax(t[i], partials[((int64_t)c * nparts + b) * 3 + i]);
    const double den = __dadd_rn(__dmul_rn(normInf, t[1]), t[2]);
    berr[c] = den > 0 ? __ddiv_rn(t[0], den) : t[0];
}

void launch_berr_user(int64_t n, const int64_t *rowptr, const int32_t *cols, const double *vals,
                      const int32_t *peri, const double *x, int64_t ldx, const double *b, int64_t ldb, int nrhs,
                      double normInf, double *partials, double *berr, hipStream_t st) {
    if (nrhs < 1) return;
    hipLaunchKernelGGL(k_berr_part_user, dim3(RED_BLOCKS, nrhs), dim3(256), 0, st, n, rowptr, cols, vals, peri, x,
                       ldx, b, ldb, partials);
    hipLaunchKernelGGL(k_berr_final_user, dim3(1), dim3(64), 0, st, RED_BLOCKS, nrhs, partials, normInf, berr);
}

#define INST(T)                                                                                               \
    template void launch_spmv<T>(int64_t, const int64_t *, const int32_t *, const T *, T, const T *, T, T *,   \
                                 hipStream_t);                                                                \
    template void launch_permute<T>(int64_t, int, const int32_t *, bool, const T *, int64_t, T *, int64_t,     \
                                    hipStream_t);
INST(double)
INST(float)

}  // namespace pxa
