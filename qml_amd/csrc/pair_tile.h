// FP64 VALU pairwise tile shared by the pairwise kernels and the batched sweep.
//
// Tiling: a 256-lane workgroup owns a 128 x 128 output tile; each lane owns an 8 x 8 register
// tile (rows 2*tx + 32*r + {0,1}, columns 2*ty + 32*c + {0,1}, tx = t % 16, ty = t / 16).
// Features stream through LDS in chunks of 16, stored k-major so a lane fetches its two
// adjacent rows (or columns) for one feature with a single 16-byte LDS read. Each lane keeps
// the feature sum of every entry in feature order k = 0, 1, ..., d-1, so an entry equals a
// serial sum over the features.
#pragma once

#include "common.h"

namespace qmlh {

static constexpr int PT_BM = 128;
static constexpr int PT_BN = 128;
static constexpr int PT_BK = 16;
static constexpr int PT_LDS = PT_BM + 2;  // padded k-major row (doubles), keeps 16 B alignment
static constexpr int PT_THREADS = 256;

enum { OP_L1 = 0, OP_SQ = 1 };
enum { EPI_IDENTITY = 0, EPI_SQRT = 1, EPI_EXP = 2 };

template <int OP>
__device__ __forceinline__ double pair_term(double acc, double a, double b) {
    if (OP == OP_L1) {
        return acc + fabs(a - b);
    } else {
        double t = a - b;
        return fma(t, t, acc);
    }
}

template <int EPI>
__device__ __forceinline__ double pair_epilogue(double acc, double scale) {
    if (EPI == EPI_EXP) return exp(scale * acc);
    if (EPI == EPI_SQRT) return sqrt(acc);
    return acc;
}

// Load a 128-row x 16-feature chunk of X (row-major, ld) into LDS as s[k][row].
// Rows >= n and features >= d read as zero; zero features add exactly zero to either sum.
__device__ __forceinline__ void pair_load_chunk(double (*s)[PT_LDS], const double* __restrict__ X,
                                                int64_t ld, int64_t n, int64_t d, int64_t row0,
                                                int64_t k0, int t) {
    const int k = t % PT_BK;
    const int rbase = t / PT_BK;  // 0..15
    const int64_t kk = k0 + k;
#pragma unroll
    for (int q = 0; q < PT_BM / 16; ++q) {
        int r = rbase + 16 * q;
        int64_t row = row0 + r;
        double v = 0.0;
        if (row < n && kk < d) v = X[row * ld + kk];
        s[k][r] = v;
    }
}

// acc[r][c] = sum over features of term(A(row0 + ri(r)), B(col0 + cj(c))) with
// ri(r) = 2 tx + 32 (r / 2) + r % 2 and cj(c) = 2 ty + 32 (c / 2) + c % 2. Has barriers.
template <int OP>
__device__ __forceinline__ void pair_tile_accumulate(double (&acc)[8][8], double (*sA)[PT_LDS],
                                                     double (*sB)[PT_LDS],
                                                     const double* __restrict__ A, int64_t lda,
                                                     int64_t na, const double* __restrict__ B,
                                                     int64_t ldb, int64_t nb, int64_t d,
                                                     int64_t row0, int64_t col0) {
    const int t = threadIdx.x;
    const int tx = t % 16;
    const int ty = t / 16;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = 0.0;

    for (int64_t k0 = 0; k0 < d; k0 += PT_BK) {
        pair_load_chunk(sA, A, lda, na, d, row0, k0, t);
        pair_load_chunk(sB, B, ldb, nb, d, col0, k0, t);
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < PT_BK; ++k) {
            double a[8], b[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double2 va = *reinterpret_cast<const double2*>(&sA[k][2 * tx + 32 * r]);
                a[2 * r] = va.x;
                a[2 * r + 1] = va.y;
                const double2 vb = *reinterpret_cast<const double2*>(&sB[k][2 * ty + 32 * r]);
                b[2 * r] = vb.x;
                b[2 * r + 1] = vb.y;
            }
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[r][c] = pair_term<OP>(acc[r][c], a[r], b[c]);
        }
        __syncthreads();
    }
}

}  // namespace qmlh
