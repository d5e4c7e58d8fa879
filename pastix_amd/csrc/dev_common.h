// Device helpers shared by the CDNA4 kernels: MFMA wrappers and the
// wave-level 32x32 tile product used by every GEMM-shaped kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pxa {

// v_mfma_{f64,f32}_16x16x4: A is 16x4, B is 4x16, D is 16x16.  Lane l holds
// A[l%16][l/16] and B[l/16][l%16] for both types; the accumulator layouts
// differ (measured on gfx950):
//   f32: D[4*(l/16) + r][l%16] in register r
//   f64: D[4*r + l/16][l%16]   in register r
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Strided read-only view: element (r, k) at p[r*rs + k*cs].
template <typename T> struct View {
    const T *p;
    int64_t rs, cs;
    __device__ __forceinline__ T at(int r, int k) const { return p[(int64_t)r * rs + (int64_t)k * cs]; }
};

// 32x32 accumulator of one wave, as 2x2 MFMA blocks of 16x16.
template <typename T> struct Quad {
    typename Mfma<T>::v4 c[2][2];
    __device__ __forceinline__ void zero() {
        for (int m = 0; m < 2; m++)
            for (int n = 0; n < 2; n++)
                for (int r = 0; r < 4; r++) c[m][n][r] = (T)0;
    }
    // Element of block (m,n), register r, owned by `lane`:
    //   row = m*16 + Mfma<T>::drow(lane, r), col = n*16 + (lane&15)
};

// acc += A[0:32, 0:K] * B[0:32, 0:K]^T where rows >= ma of A and rows >= nb
// of B read as zero.  K is consumed in groups of QG k-steps (4*QG columns):
// all operand loads of group g+1 are issued before the MFMAs of group g, so
// 16*QG independent loads per lane are in flight while the matrix core works.
template <typename T>
__device__ __forceinline__ void quad_load_group(View<T> A, View<T> B, int k0, int K, int i, int kq, bool ra0,
                                                bool ra1, bool rb0, bool rb1, T (&a0)[4], T (&a1)[4], T (&b0)[4],
                                                T (&b1)[4]) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int k = k0 + 4 * u + kq;
        const bool ok = k < K;
        a0[u] = (ok && ra0) ? A.at(i, k) : (T)0;
        a1[u] = (ok && ra1) ? A.at(i + 16, k) : (T)0;
        b0[u] = (ok && rb0) ? B.at(i, k) : (T)0;
        b1[u] = (ok && rb1) ? B.at(i + 16, k) : (T)0;
    }
}

template <typename T>
__device__ __forceinline__ void quad_gemm_nt(Quad<T> &q, View<T> A, int ma, View<T> B, int nb, int K, int lane) {
    const int i = lane & 15, kq = lane >> 4;
    const bool ra0 = i < ma, ra1 = i + 16 < ma, rb0 = i < nb, rb1 = i + 16 < nb;
    T a0[4], a1[4], b0[4], b1[4];
    quad_load_group(A, B, 0, K, i, kq, ra0, ra1, rb0, rb1, a0, a1, b0, b1);
    for (int k = 0; k < K; k += 16) {
        T na0[4], na1[4], nb0[4], nb1[4];
        if (k + 16 < K) quad_load_group(A, B, k + 16, K, i, kq, ra0, ra1, rb0, rb1, na0, na1, nb0, nb1);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k + 4 * u < K) {  // uniform across the wave
                q.c[0][0] = Mfma<T>::mma(a0[u], b0[u], q.c[0][0]);
                q.c[0][1] = Mfma<T>::mma(a0[u], b1[u], q.c[0][1]);
                q.c[1][0] = Mfma<T>::mma(a1[u], b0[u], q.c[1][0]);
                q.c[1][1] = Mfma<T>::mma(a1[u], b1[u], q.c[1][1]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0[u] = na0[u]; a1[u] = na1[u]; b0[u] = nb0[u]; b1[u] = nb1[u];
        }
    }
}

// Workgroup-level D(64x64) = X(64xK) * Y(64xK)^T with operands staged through
// LDS in chunks of CH columns (16 or 32): the 256 threads fetch each chunk
// once with coalesced column segments (rows contiguous in memory), the 4
// waves read their MFMA fragments from LDS.  Global loads of chunk c+1 are in
// flight while chunk c is multiplied, so a chunk's MFMA time (CH/4 * 4
// MFMAs per wave) is the latency the prefetch can hide.  Rows >= mx of X /
// >= my of Y read as zero; X(r, k) = X[r + k*ldx].  Wave w owns the 32x32
// quadrant (w>>1, w&1).  Must be called by all 256 threads (contains
// barriers).
// two elements: V2<T>::u with the alignment of one panel element (panels
// start and lead at any element), V2<T>::a fully aligned (LDS staging rows)
// Operand staging moves 16 bytes per lane: E = 16 / sizeof(T) consecutive
// rows of one column (4 floats or 2 doubles).  Vec<T>::u has the alignment
// of one panel element (panels start and lead at any element), Vec<T>::a is
// fully aligned (LDS staging rows, whose pitch is a multiple of 16 bytes).
template <typename T> struct Vec;
template <> struct Vec<float> {
    static constexpr int E = 4;
    typedef float u __attribute__((ext_vector_type(4), aligned(4)));
    typedef float a __attribute__((ext_vector_type(4)));
};
template <> struct Vec<double> {
    static constexpr int E = 2;
    typedef double u __attribute__((ext_vector_type(2), aligned(8)));
    typedef double a __attribute__((ext_vector_type(2)));
};
// E rows [r, r+E) of column p (rows >= m read as zero) into o
template <typename T> __device__ __forceinline__ void ld_rows(const T *p, int r, int m, T *o) {
    constexpr int E = Vec<T>::E;
    if (r + E <= m) {
        const typename Vec<T>::u v = *reinterpret_cast<const typename Vec<T>::u *>(p);
#pragma unroll
        for (int e = 0; e < E; e++) o[e] = v[e];
    } else {
#pragma unroll
        for (int e = 0; e < E; e++) o[e] = (r + e < m) ? p[e] : (T)0;
    }
}
template <typename T> __device__ __forceinline__ void st_rows(T *p, const T *o) {
    typename Vec<T>::a v;
#pragma unroll
    for (int e = 0; e < Vec<T>::E; e++) v[e] = o[e];
    *reinterpret_cast<typename Vec<T>::a *>(p) = v;
}

template <typename T, int CH = 16> struct StageBuf {
    T xs[CH][68];
    T ys[CH][68];
};

template <typename T, int CH>
__device__ __forceinline__ void wg_gemm_lds(Quad<T> &q, const T *X, int mx, int64_t ldx, const T *Y, int my,
                                            int64_t ldy, int K, StageBuf<T, CH> &sb, int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int i = lane & 15, kq = lane >> 4;
    // 16 bytes per lane: E rows of one column, 4E columns per pass of the
    // workgroup, CH / 4E passes per chunk
    constexpr int E = Vec<T>::E, NP = CH / (4 * E);
    static_assert(CH % (4 * E) == 0, "chunk of whole passes");
    const int re = E * (tid % (64 / E)), ke = tid / (64 / E);
    T gx[NP][E], gy[NP][E];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NP; u++) {
            const int k = k0 + ke + 4 * E * u;
            ld_rows(X + re + (int64_t)k * ldx, re, k < K ? mx : 0, gx[u]);
            ld_rows(Y + re + (int64_t)k * ldy, re, k < K ? my : 0, gy[u]);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int u = 0; u < NP; u++) {
            st_rows(&sb.xs[ke + 4 * E * u][re], gx[u]);
            st_rows(&sb.ys[ke + 4 * E * u][re], gy[u]);
        }
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

// The same product for a 64 x 128 tile: q0, q1 += X[0:64, 0:K] * Y[0:128,
// 0:K]^T, wave w owning X rows (w>>1)*32 + [0, 32) against Y rows (w&1)*64 +
// [0, 32) (q0) and [32, 64) (q1).  Each staged X fragment feeds twice as
// many MFMAs as in the 64 x 64 form.  Must be called by all 256 threads.
template <typename T, int CH = 16> struct StageBuf2 {
    T xs[CH][68];
    T ys[CH][132];
};

template <typename T, int CH>
__device__ __forceinline__ void wg_gemm_lds_64x128(Quad<T> &q0, Quad<T> &q1, const T *X, int mx, int64_t ldx,
                                                   const T *Y, int my, int64_t ldy, int K, StageBuf2<T, CH> &sb,
                                                   int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;
    const int i = lane & 15, kq = lane >> 4;
    // 16 bytes per lane: X (64 rows) 4E columns per pass, Y (128 rows) 2E
    constexpr int E = Vec<T>::E, NPX = CH / (4 * E), NPY = CH / (2 * E);
    static_assert(CH % (4 * E) == 0, "chunk of whole passes");
    const int rex = E * (tid % (64 / E)), kex = tid / (64 / E);
    const int rey = E * (tid % (128 / E)), key = tid / (128 / E);
    T gx[NPX][E], gy[NPY][E];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NPX; u++) {
            const int k = k0 + kex + 4 * E * u;
            ld_rows(X + rex + (int64_t)k * ldx, rex, k < K ? mx : 0, gx[u]);
        }
#pragma unroll
        for (int u = 0; u < NPY; u++) {
            const int k = k0 + key + 2 * E * u;
            ld_rows(Y + rey + (int64_t)k * ldy, rey, k < K ? my : 0, gy[u]);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int u = 0; u < NPX; u++) st_rows(&sb.xs[kex + 4 * E * u][rex], gx[u]);
#pragma unroll
        for (int u = 0; u < NPY; u++) st_rows(&sb.ys[key + 2 * E * u][rey], gy[u]);
    };
    const bool act0 = (wr < mx) && (wc < my), act1 = act0 && (wc + 32 < my);
    gload(0);
    sstore();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += CH) {
        const bool more = k0 + CH < K;
        if (more) gload(k0 + CH);
        if (act0) {
#pragma unroll
            for (int s = 0; s < CH / 4; s++) {
                if (k0 + 4 * s < K) {
                    const int kr = 4 * s + kq;
                    const T a0 = sb.xs[kr][wr + i], a1 = sb.xs[kr][wr + 16 + i];
                    const T b0 = sb.ys[kr][wc + i], b1 = sb.ys[kr][wc + 16 + i];
                    q0.c[0][0] = Mfma<T>::mma(a0, b0, q0.c[0][0]);
                    q0.c[0][1] = Mfma<T>::mma(a0, b1, q0.c[0][1]);
                    q0.c[1][0] = Mfma<T>::mma(a1, b0, q0.c[1][0]);
                    q0.c[1][1] = Mfma<T>::mma(a1, b1, q0.c[1][1]);
                    if (act1) {
                        const T b2 = sb.ys[kr][wc + 32 + i], b3 = sb.ys[kr][wc + 48 + i];
                        q1.c[0][0] = Mfma<T>::mma(a0, b2, q1.c[0][0]);
                        q1.c[0][1] = Mfma<T>::mma(a0, b3, q1.c[0][1]);
                        q1.c[1][0] = Mfma<T>::mma(a1, b2, q1.c[1][0]);
                        q1.c[1][1] = Mfma<T>::mma(a1, b3, q1.c[1][1]);
                    }
                }
            }
        }
        __syncthreads();
        if (more) sstore();
        __syncthreads();
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of work items so neighbouring items share its L2.  The grid
// size must be a multiple of 8 (the map is then a bijection of [0, grid));
// callers drop ids beyond their item count.
__device__ __forceinline__ int xcd_remap(int bid, int grid) {
    int per = grid >> 3;
    return (bid & 7) * per + (bid >> 3);
}

}  // namespace pxa
