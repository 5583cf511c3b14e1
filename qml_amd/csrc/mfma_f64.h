// FP64 MFMA tile core: C(i, j) = sum_k A(i, k) * B(j, k) on v_mfma_f64_16x16x4f64.
//
// One wavefront owns a 64 x 64 tile held as 4 x 4 MFMA blocks (16 x double4 accumulators).
// Operand layout of v_mfma_f64_16x16x4f64: lane l supplies A(l % 16, l / 16) and
// B(l / 16, l % 16); the result block D(i, j) sits in lane j + 16 * (i % 4), element i / 4
// (rows interleave across the four lane groups; checked on gfx950 by the parity tests).
//
// Fragments are read straight from global memory (the FP64 MFMA issues once per 64 cycles, so
// a 16-deep k chunk keeps a wave busy ~4k cycles, long enough to cover the loads of the next
// chunk). For a 16-deep chunk, lane (r = l % 16, g = l / 16) reads the four features
// k0 + 4g .. k0 + 4g + 3 of its row, and MFMA number s (0..3) consumes feature 4g + s: the
// four MFMAs together cover all 16 features once, with the same permutation on A and B, so
// the product is unchanged while each lane's row-major read is one contiguous 32-byte access.
#pragma once

#include "common.h"

namespace qmlh {

enum { LAYOUT_ROW = 0, LAYOUT_COL = 1 };

// Operand view: element (i, k) at p[i * ld + k] (ROW) or p[i + k * ld] (COL).
template <int LAYOUT>
struct Operand {
    const double* __restrict__ p;
    int64_t ld;
    int64_t rows;  // valid rows; row indices are clamped to rows - 1 when loading
    int64_t kdim;  // valid features; features >= kdim read as zero
};

struct Frag4 {
    double v[4];
};

// Load the 4 features of one 16-row block for this lane.
template <int LAYOUT>
__device__ __forceinline__ Frag4 load_frag(const Operand<LAYOUT>& op, int64_t row, int64_t k0,
                                           bool full) {
    Frag4 f;
    const int64_t r = row < op.rows ? row : op.rows - 1;
    if (LAYOUT == LAYOUT_ROW) {
        const double* src = op.p + r * op.ld + k0;
        if (full) {
            const double2 x = reinterpret_cast<const double2*>(src)[0];
            const double2 y = reinterpret_cast<const double2*>(src)[1];
            f.v[0] = x.x; f.v[1] = x.y; f.v[2] = y.x; f.v[3] = y.y;
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) f.v[s] = (k0 + s < op.kdim) ? src[s] : 0.0;
        }
    } else {
        const double* src = op.p + r + k0 * op.ld;
#pragma unroll
        for (int s = 0; s < 4; ++s)
            f.v[s] = (full || k0 + s < op.kdim) ? src[s * op.ld] : 0.0;
    }
    return f;
}

struct WaveTile {
    double4_t acc[4][4];
};

__device__ __forceinline__ void wave_tile_zero(WaveTile& t) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) t.acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
}

// Accumulate A(row0 .. row0+63, 0..kdim) * B(col0 .. col0+63, 0..kdim)^T into t.
// Requires (ROW layout) ld % 2 == 0 and 16-byte aligned rows for the vector loads.
template <int LA, int LB>
__device__ __forceinline__ void wave_tile_mma(WaveTile& t, const Operand<LA>& A,
                                              const Operand<LB>& B, int64_t row0, int64_t col0,
                                              int64_t kdim) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15;
    const int g = lane >> 4;
    Frag4 fa[4], fb[4];
    int64_t k0 = 0;
    if (kdim <= 0) return;
    {
        const bool full = 16 <= kdim;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            fa[q] = load_frag<LA>(A, row0 + 16 * q + r, 4 * g, full);
            fb[q] = load_frag<LB>(B, col0 + 16 * q + r, 4 * g, full);
        }
    }
    for (; k0 < kdim; k0 += 16) {
        Frag4 na[4], nb[4];
        const int64_t kn = k0 + 16;
        const bool more = kn < kdim;
        if (more) {
            const bool full = kn + 16 <= kdim;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                na[q] = load_frag<LA>(A, row0 + 16 * q + r, kn + 4 * g, full);
                nb[q] = load_frag<LB>(B, col0 + 16 * q + r, kn + 4 * g, full);
            }
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    t.acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a].v[s], fb[b].v[s],
                                                                       t.acc[a][b], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                fa[q] = na[q];
                fb[q] = nb[q];
            }
        }
    }
}

// Row/column of accumulator element e (0..3) of block (a, b) for this lane.
__device__ __forceinline__ int tile_row(int a, int e) {
    const int lane = threadIdx.x & 63;
    return 16 * a + (lane >> 4) + 4 * e;
}
__device__ __forceinline__ int tile_col(int b) {
    const int lane = threadIdx.x & 63;
    return 16 * b + (lane & 15);
}

}  // namespace qmlh
