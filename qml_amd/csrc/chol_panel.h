// Cholesky panel building blocks shared by the single-matrix factorisation (potrf.hip) and the
// batched one (batched.hip): the factor of a <= 64 x 64 diagonal block in LDS, and the
// row-parallel solve of the panel rows below it.
#pragma once

#include "common.h"

namespace qmlh {

static constexpr int CP_B = 64;  // diagonal block / sub-panel width

static constexpr int CP_S = 16;  // step of the blocked diagonal factor

// LDS (FP64) <- the lower triangle of the w x w block at A (column-major, lda), zero elsewhere;
// TRANS stores element (i, j) at s[j][i]. All 256 lanes take part; each issues its 16 loads
// back to back before any LDS store (a loop with a runtime bound would wait on every load).
template <bool TRANS, int LDS_LD, typename T>
__device__ __forceinline__ void stage_lower64(double (*s)[LDS_LD], const T* __restrict__ A,
                                              int64_t lda, int w) {
    const int t = threadIdx.x;
    double v[CP_B * CP_B / 256];
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;  // consecutive lanes: consecutive rows
        v[q] = (i < w && j < w && i >= j) ? (double)A[i + (int64_t)j * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        if (TRANS) s[j][i] = v[q]; else s[i][j] = v[q];
    }
}

// Cholesky of the w x w (w <= 64) lower block in s (FP64, LDS), all 256 lanes, in four
// 16-wide steps: the 16 x 16 diagonal block is factored in the registers of lanes 0..15 of
// wave 0 (pivot and column entries broadcast by lane shuffles), the rows below it are solved
// against it (one lane per row), and the trailing lower block gets the rank-16 update (lanes
// spread over its entries). On exit s holds L in its lower triangle. Returns 0, or c + 1 for
// the first non-positive (or NaN) pivot c, with the Schur value left on that diagonal entry
// (as LAPACK leaves it); the value is the same in every lane.
__device__ __forceinline__ int potf2_block64(double (*s)[CP_B + 1], int w) {
    __shared__ int s_fail;
    const int t = threadIdx.x;
    if (t == 0) s_fail = 0;
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < CP_B; kb += CP_S) {
        if (kb >= w) break;
        const int bw = w - kb < CP_S ? w - kb : CP_S;
        // 1. factor the diagonal 16 x 16 block
        if (t < 64) {
            const int lane = t;
            double a[CP_S];
#pragma unroll
            for (int j = 0; j < CP_S; ++j)
                a[j] = (lane < bw && j <= lane) ? s[kb + lane][kb + j] : 0.0;
            int f = 0;
#pragma unroll
            for (int c = 0; c < CP_S; ++c) {
                if (c < bw && f == 0) {
                    const double piv = __shfl(a[c], c, 64);
                    if (!(piv > 0.0)) {
                        f = c + 1;
                    } else {
                        const double l = sqrt(piv);
                        const double rl = 1.0 / l;
                        a[c] = lane == c ? l : (lane > c ? a[c] * rl : a[c]);
#pragma unroll
                        for (int j = c + 1; j < CP_S; ++j) {
                            const double lj = __shfl(a[c], j, 64);  // L(kb + j, kb + c)
                            a[j] = fma(-a[c], lj, a[j]);  // rows above j: unused upper values
                        }
                    }
                }
            }
            if (lane < bw) {
#pragma unroll
                for (int j = 0; j < CP_S; ++j)
                    if (j <= lane) s[kb + lane][kb + j] = a[j];
            }
            if (lane == 0 && f) s_fail = kb + f;
        }
        __syncthreads();
        if (s_fail) return s_fail;
        const int r0 = kb + bw, rem = w - r0;  // rows below the block
        // 2. rows below: X := B L^-T (one lane per row)
        if (t < rem) {
            const int i = r0 + t;
            double x[CP_S];
#pragma unroll
            for (int c = 0; c < CP_S; ++c) x[c] = c < bw ? s[i][kb + c] : 0.0;
#pragma unroll
            for (int c = 0; c < CP_S; ++c) {
                if (c < bw) {
                    x[c] = x[c] / s[kb + c][kb + c];
#pragma unroll
                    for (int j = c + 1; j < CP_S; ++j) x[j] = fma(-x[c], s[kb + j][kb + c], x[j]);
                }
            }
#pragma unroll
            for (int c = 0; c < CP_S; ++c)
                if (c < bw) s[i][kb + c] = x[c];
        }
        __syncthreads();
        // 3. trailing lower block: s[i][j] -= sum_c s[i][kb + c] s[j][kb + c], r0 <= j <= i < w
        constexpr int RM = CP_B - CP_S;  // largest trailing order (48); entries beyond rem skip
        for (int idx = t; idx < RM * RM; idx += 256) {
            const int io = idx % RM, jo = idx / RM;
            if (io < rem && jo <= io) {
                const int i = r0 + io, j = r0 + jo;
                double v = s[i][j];
#pragma unroll
                for (int c = 0; c < CP_S; ++c)
                    if (c < bw) v = fma(-s[i][kb + c], s[j][kb + c], v);
                s[i][j] = v;
            }
        }
        __syncthreads();
    }
    return 0;
}

// Row `row` of the m x w panel B (column-major, ldb) := B L^-T with L^T in sT (sT[c][j] =
// L(j, c)) and the reciprocal pivots in inv; FP64 arithmetic, storage T. The row is loaded
// whole first; then, column by column, x_c is scaled and subtracted from the later entries
// (independent FMAs; per entry the same operations in the same order as a row substitution);
// then the row is stored.
template <typename T>
__device__ __forceinline__ void trsm_row(T* __restrict__ B, int64_t ldb, int64_t row, int w,
                                         const double (*sT)[CP_B + 2],
                                         const double* __restrict__ inv) {
    double x[CP_B];
#pragma unroll
    for (int c = 0; c < CP_B; ++c) x[c] = c < w ? (double)B[row + c * ldb] : 0.0;
#pragma unroll
    for (int c = 0; c < CP_B; ++c) {
        if (c < w) {
            x[c] = x[c] * inv[c];
#pragma unroll
            for (int j = c + 1; j < CP_B; ++j) x[j] = fma(-x[c], sT[c][j], x[j]);
        }
    }
#pragma unroll
    for (int c = 0; c < CP_B; ++c)
        if (c < w) B[row + c * ldb] = (T)x[c];
}

// Diagonal block of a sub-panel, one workgroup: factor the w x w block at A in wave 0, write
// L back, set info = col_base + c + 1 on a failed pivot. `info` already set: nothing is done.
template <typename T>
__device__ __forceinline__ void potf2_body(T* __restrict__ A, int64_t lda, int w,
                                           int64_t col_base, int32_t* __restrict__ info) {
    __shared__ double s[CP_B][CP_B + 1];
    if (*info) return;
    const int t = threadIdx.x;
    stage_lower64<false, CP_B + 1>(s, A, lda, w);
    __syncthreads();
    const int fail = potf2_block64(s, w);
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        if (i < w && j < w && i >= j) A[i + (int64_t)j * lda] = (T)s[i][j];
    }
    if (t == 0 && fail) *info = (int32_t)(col_base + fail);
}

// Rows of a sub-panel, workgroup `local` of the launch: rows 256 local .. of the m rows below
// the factored w x w block L at A become B L^-T. Runs after potf2_body in stream order, so L
// is final when it is read. `info` set: nothing is done.
template <typename T>
__device__ __forceinline__ void trsm_body(T* __restrict__ A, int64_t lda, int w, int64_t m,
                                          const int32_t* __restrict__ info, int64_t local) {
    __shared__ __attribute__((aligned(16))) double sT[CP_B][CP_B + 2];  // sT[c][j] = L(j, c)
    __shared__ double inv[CP_B];
    if (*info) return;
    const int t = threadIdx.x;
    stage_lower64<true, CP_B + 2>(sT, A, lda, w);
    __syncthreads();
    if (t < w) inv[t] = 1.0 / sT[t][t];
    __syncthreads();
    const int64_t row = local * 256 + t;
    if (row < m) trsm_row<T>(A + w, lda, row, w, sT, inv);
}

}  // namespace qmlh
