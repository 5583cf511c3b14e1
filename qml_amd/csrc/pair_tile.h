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

// Per-feature terms: |a - b|, (a - b)^2, |a - b|^3 and |a - b|^4 by the products binary
// powering forms (x (x x), (x x)(x x)), |a - b|^p for any other p (pow).
enum { OP_L1 = 0, OP_SQ = 1, OP_POW3 = 2, OP_POW4 = 3, OP_POWR = 4 };
// Epilogues of the feature sum s: s, sqrt(s), exp(s0 s), the Sargan polynomial-times-exp in
// the L1 distance, the Matern forms in the L2 distance, s^s0 (the p-norm root).
enum { EPI_IDENTITY = 0, EPI_SQRT = 1, EPI_EXP = 2, EPI_SARGAN = 3, EPI_MATERN = 4, EPI_ROOT = 5 };

static constexpr int PAIR_MAX_GAMMAS = 16;

struct PairEpi {
    double s0, s1;                 // scales (see the epilogues)
    double g[PAIR_MAX_GAMMAS];     // Sargan coefficients gamma_1 ..
    int ng;                        // their count
    int order;                     // Matern order 0, 1, 2
};

// x^n for n >= 0 by binary powering.
__device__ __forceinline__ double pair_powi(double x, int n) {
    double r = (n & 1) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1) r = r * x;
    }
    return r;
}

template <int OP>
__device__ __forceinline__ double pair_term(double acc, double a, double b, double p = 0.0) {
    if (OP == OP_L1) {
        return acc + fabs(a - b);
    } else if (OP == OP_SQ) {
        double t = a - b;
        return fma(t, t, acc);
    } else if (OP == OP_POW3) {
        const double x = fabs(a - b);
        return acc + x * (x * x);
    } else if (OP == OP_POW4) {
        const double x = fabs(a - b);
        const double x2 = x * x;
        return acc + x2 * x2;
    } else {
        // |a - b|^p as exp(p log|a - b|): log and exp instead of pow's extended-precision path
        // (relative error about p |log x| ulp, far inside the 1e-12 the kernel is held to);
        // x = 0 gives exp(-inf) = 0 as pow does for p > 0.
        return acc + exp(p * log(fabs(a - b)));
    }
}

template <int EPI>
__device__ __forceinline__ double pair_epilogue(double acc, const PairEpi& e) {
    if (EPI == EPI_EXP) return exp(e.s0 * acc);
    if (EPI == EPI_SQRT) return sqrt(acc);
    if (EPI == EPI_ROOT) return pow(acc, e.s0);
    if (EPI == EPI_SARGAN) {
        // qml/fkernels.f90:513-518: sum_m gamma_m (-inv d)^m, then exp(inv d) (1 + sum).
        const double x = -e.s0 * acc;
        double sum = 0.0;
        for (int m = 0; m < e.ng; ++m) sum = sum + e.g[m] * pair_powi(x, m + 1);
        return exp(e.s0 * acc) * (1 + sum);
    }
    if (EPI == EPI_MATERN) {
        // qml/fkernels.f90:442-482 with acc the squared distance.
        const double d = sqrt(acc);
        if (e.order == 0) return exp(e.s0 * d);
        if (e.order == 1) return exp(e.s0 * d) * (1.0 - e.s0 * d);
        return exp(e.s0 * d) * (1.0 - e.s0 * d + e.s1 * acc);
    }
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
// The next chunk's global loads are issued into registers before the current chunk's
// products and written to LDS after them, so their latency hides behind the arithmetic.
template <int OP>
__device__ __forceinline__ void pair_tile_accumulate(double (&acc)[8][8], double (*sA)[PT_LDS],
                                                     double (*sB)[PT_LDS],
                                                     const double* __restrict__ A, int64_t lda,
                                                     int64_t na, const double* __restrict__ B,
                                                     int64_t ldb, int64_t nb, int64_t d,
                                                     int64_t row0, int64_t col0, double p = 0.0) {
    const int t = threadIdx.x;
    const int tx = t % 16;
    const int ty = t / 16;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = 0.0;

    // chunk staging: lane t holds feature t % 16 of rows t / 16 + 16 q (q < 8), as
    // pair_load_chunk; rows >= n and features >= d read as zero
    const int kl = t % PT_BK;
    const int rbase = t / PT_BK;
    double pa[PT_BM / 16], pb[PT_BN / 16];
    auto fetch = [&](int64_t k0) {
        const int64_t kk = k0 + kl;
#pragma unroll
        for (int q = 0; q < PT_BM / 16; ++q) {
            const int64_t ra = row0 + rbase + 16 * q, rb = col0 + rbase + 16 * q;
            pa[q] = (ra < na && kk < d) ? A[ra * lda + kk] : 0.0;
            pb[q] = (rb < nb && kk < d) ? B[rb * ldb + kk] : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < PT_BM / 16; ++q) {
            sA[kl][rbase + 16 * q] = pa[q];
            sB[kl][rbase + 16 * q] = pb[q];
        }
    };
    fetch(0);
    stash();
    __syncthreads();
    for (int64_t k0 = 0; k0 < d; k0 += PT_BK) {
        const bool more = k0 + PT_BK < d;
        if (more) fetch(k0 + PT_BK);
#pragma unroll 2
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
                for (int c = 0; c < 8; ++c)
                    acc[r][c] = pair_term<OP>(acc[r][c], a[r], b[c], p);
        }
        __syncthreads();  // every lane has read the chunk
        if (more) {
            stash();
            __syncthreads();
        }
    }
}

}  // namespace qmlh
