// Cholesky panel building blocks shared by the single-matrix factorisation (potrf.hip) and the
// batched one (batched.hip): the unblocked factor of a <= 64 x 64 diagonal block held in the
// registers of one wavefront, and the row-parallel solve of the panel rows below it.
#pragma once

#include "common.h"

namespace qmlh {

static constexpr int CP_B = 64;  // diagonal block / sub-panel width

// Right-looking unblocked Cholesky of the w x w (w <= 64) lower block in s (FP64), by wave 0
// of the workgroup only: lane i keeps row i in registers; column c is scaled by 1 / L(c, c)
// and its entries are broadcast with lane shuffles for the rank-1 update of the rows below,
// so no LDS round trips or barriers sit on the 64-step critical path. On exit s holds L in
// its lower triangle. Returns 0, or c + 1 for the first non-positive (or NaN) pivot c, with
// the Schur value left on that diagonal entry (as LAPACK leaves it). The result is the same
// in every lane (the pivot is read from one lane).
__device__ __forceinline__ int potf2_wave(double (*s)[CP_B + 1], int w) {
    const int lane = threadIdx.x & 63;
    double a[CP_B];
#pragma unroll
    for (int j = 0; j < CP_B; ++j) a[j] = (lane < w && j <= lane) ? s[lane][j] : 0.0;
    int fail = 0;
#pragma unroll
    for (int c = 0; c < CP_B; ++c) {
        if (c < w && fail == 0) {
            const double piv = __shfl(a[c], c, 64);
            if (!(piv > 0.0)) {
                fail = c + 1;
            } else {
                const double l = sqrt(piv);
                const double rl = 1.0 / l;
                a[c] = lane == c ? l : (lane > c ? a[c] * rl : a[c]);
#pragma unroll
                for (int j = c + 1; j < CP_B; ++j) {
                    if (j < w) {
                        const double lj = __shfl(a[c], j, 64);  // L(j, c)
                        // Rows i < j receive upper-triangle values that are never used.
                        a[j] = fma(-a[c], lj, a[j]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < CP_B; ++j)
        if (lane < w && j <= lane) s[lane][j] = a[j];
    return fail;
}

// Row `row` of the m x w panel B (column-major, ldb) := B L^-T with L (w x w) in s and the
// reciprocal pivots in inv; FP64 arithmetic, storage T.
template <typename T>
__device__ __forceinline__ void trsm_row(T* __restrict__ B, int64_t ldb, int64_t row, int w,
                                         const double (*s)[CP_B + 1],
                                         const double* __restrict__ inv) {
    double x[CP_B];
#pragma unroll
    for (int c = 0; c < CP_B; ++c) {
        if (c < w) {
            double v = (double)B[row + c * ldb];
#pragma unroll
            for (int l = 0; l < c; ++l) v = v - x[l] * s[c][l];
            x[c] = v * inv[c];
            B[row + c * ldb] = (T)x[c];
        }
    }
}

// One launch per 64-wide sub-panel: every workgroup loads the w x w diagonal block at A,
// factors it in wave 0 (all workgroups compute the same factor), workgroup 0 writes it back
// and sets info = col_base + c + 1 on a failed pivot, then each workgroup solves its 256 rows
// of the m rows below the block. `info` already set: nothing is done.
template <typename T>
__device__ __forceinline__ void panel64_body(T* __restrict__ A, int64_t lda, int w, int64_t m,
                                             int64_t col_base, int32_t* __restrict__ info,
                                             int64_t local) {
    __shared__ double s[CP_B][CP_B + 1];
    __shared__ double inv[CP_B];
    __shared__ int s_fail;
    if (*info) return;
    const int t = threadIdx.x;
    for (int idx = t; idx < w * w; idx += 256) {
        const int i = idx % w, j = idx / w;
        if (i >= j) s[i][j] = (double)A[i + j * lda];
    }
    __syncthreads();
    if (t < 64) {
        const int f = potf2_wave(s, w);
        if (t == 0) s_fail = f;
    }
    __syncthreads();
    const int fail = s_fail;
    if (local == 0) {
        for (int idx = t; idx < w * w; idx += 256) {
            const int i = idx % w, j = idx / w;
            if (i >= j) A[i + j * lda] = (T)s[i][j];
        }
        if (fail && t == 0) *info = (int32_t)(col_base + fail);
    }
    if (fail) return;
    if (t < w) inv[t] = 1.0 / s[t][t];
    __syncthreads();
    const int64_t row = local * 256 + t;
    if (row < m) trsm_row<T>(A + w, lda, row, w, s, inv);
}

}  // namespace qmlh
