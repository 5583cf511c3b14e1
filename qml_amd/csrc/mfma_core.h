// MFMA tile cores shared by the kernel-matrix builds, the Cholesky updates and the predict.
//
// Element types: double on v_mfma_f64_16x16x4f64, float on v_mfma_f32_16x16x4f32. Both take
// one A value and one B value per lane per 4-deep k step: lane l supplies A(l % 16, l / 16)
// and B(l / 16, l % 16). The 16 x 16 result block is 4 values per lane:
//   f64: D(i, j) in lane j + 16 * (i % 4), element i / 4 (rows interleave across lane groups)
//   f32: D(i, j) in lane j + 16 * (i / 4), element i % 4
// (both pinned on gfx950 by the parity tests; MfmaTraits<T>::acc_row encodes them).
//
// Feature permutation: in a 16-deep chunk, lane group g = l / 16 holds features 4g .. 4g+3 and
// MFMA number s (0..3) consumes feature 4g + s. The four MFMAs cover all 16 features once,
// with the same permutation on both operands, so the product is unchanged.
#pragma once

#include "common.h"

namespace qmlh {

enum { LAYOUT_ROW = 0, LAYOUT_COL = 1 };

template <typename T>
struct MfmaTraits;

template <>
struct MfmaTraits<double> {
    typedef double4_t acc_t;
    typedef double2 pair_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int acc_row(int lane, int e) { return (lane >> 4) + 4 * e; }
};

template <>
struct MfmaTraits<float> {
    typedef float4_t acc_t;
    typedef float2 pair_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int acc_row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

// Operand view: element (i, k) at p[i * ld + k] (ROW) or p[i + k * ld] (COL).
template <typename T, int LAYOUT>
struct Operand {
    const T* __restrict__ p;
    int64_t ld;
    int64_t rows;  // valid rows; row indices are clamped to rows - 1 when loading
    int64_t kdim;  // valid features; features >= kdim read as zero
};

// One wavefront's 64 x 64 block: 4 x 4 MFMA blocks.
template <typename T>
struct WaveTile {
    typename MfmaTraits<T>::acc_t acc[4][4];
};

template <typename T>
__device__ __forceinline__ void wave_tile_zero(WaveTile<T>& t) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) t.acc[a][b][e] = T(0);
}

// Output row (within the wave's 64 rows of the first operand) of accumulator element e of
// first-operand block a for this lane; the second-operand index is 16 b + (lane & 15).
template <typename T>
__device__ __forceinline__ int tile_row(int a, int e) {
    return 16 * a + MfmaTraits<T>::acc_row(threadIdx.x & 63, e);
}

// ------------------------------------------------------------------------------------------
// Workgroup tile: 256 lanes own 128 x 128 (2 x 2 waves of 64 x 64). Both operands are staged
// through LDS in 16-feature chunks, double buffered: while the waves run the 64 MFMAs of chunk
// c from one buffer, every lane holds its share of chunk c + 1 in registers (coalesced global
// loads) and writes it to the other buffer after the MFMAs; one barrier per chunk. LDS layout
// of a panel: feature k = 4g + 2h + e of row r at ((2g + h) * 128 + r) * 2 + e, so a lane's
// pair (row r, features 4g + 2h + {0,1}) is one read and the 16 lanes of a group read
// consecutive pairs.
// ------------------------------------------------------------------------------------------
static constexpr int WG_T = 128;
static constexpr int WG_KC = 16;
static constexpr int WG_PANEL = WG_T * WG_KC;        // elements per operand panel
static constexpr int WG_LDS_ELEMS = 2 * 2 * WG_PANEL;  // two buffers x two operands

__device__ __forceinline__ int lds_index(int row, int k) {
    const int g = k >> 2, h = (k >> 1) & 1, e = k & 1;
    return (((g * 2 + h) * WG_T) + row) * 2 + e;
}

template <typename T>
struct PanelRegs {
    T v[8];  // 128 x 16 panel / 256 lanes
};

template <typename T, int LAYOUT>
__device__ __forceinline__ bool operand_vec_ok(const Operand<T, LAYOUT>& op) {
    return ((reinterpret_cast<uintptr_t>(op.p) & (2 * sizeof(T) - 1)) == 0) && ((op.ld & 1) == 0);
}

// Registers <- features k0 .. k0+15 of panel rows row0 .. row0+127 (clamped rows, zero
// features beyond kdim). COL: lane holds rows 2(t%64)+{0,1} of features t/64 + 4u.
// ROW: lane holds row t/2, features 8(t%2) .. 8(t%2)+7.
template <typename T, int LAYOUT>
__device__ __forceinline__ void panel_load(PanelRegs<T>& R, const Operand<T, LAYOUT>& op,
                                           int64_t row0, int64_t k0, bool vec) {
    typedef typename MfmaTraits<T>::pair_t P2;
    const int t = threadIdx.x;
    if (LAYOUT == LAYOUT_COL) {
        const int64_t ra = row0 + 2 * (t & 63);
        const int64_t r0 = ra < op.rows ? ra : op.rows - 1;
        const int64_t r1 = ra + 1 < op.rows ? ra + 1 : op.rows - 1;
        const bool pair = vec && (ra + 1 < op.rows);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t k = k0 + (t >> 6) + 4 * u;
            T x = T(0), y = T(0);
            if (k < op.kdim) {
                const T* col = op.p + k * op.ld;
                if (pair) {
                    const P2 v = *reinterpret_cast<const P2*>(col + r0);
                    x = v.x;
                    y = v.y;
                } else {
                    x = col[r0];
                    y = col[r1];
                }
            }
            R.v[2 * u] = x;
            R.v[2 * u + 1] = y;
        }
    } else {
        const int64_t ra = row0 + (t >> 1);
        const int64_t r = ra < op.rows ? ra : op.rows - 1;
        const T* src = op.p + r * op.ld;
        const int64_t kb = k0 + 8 * (t & 1);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t k = kb + 2 * u;
            T x = T(0), y = T(0);
            if (vec && k + 1 < op.kdim) {
                const P2 v = *reinterpret_cast<const P2*>(src + k);
                x = v.x;
                y = v.y;
            } else {
                if (k < op.kdim) x = src[k];
                if (k + 1 < op.kdim) y = src[k + 1];
            }
            R.v[2 * u] = x;
            R.v[2 * u + 1] = y;
        }
    }
}

template <typename T, int LAYOUT>
__device__ __forceinline__ void panel_store(T* __restrict__ panel, const PanelRegs<T>& R) {
    typedef typename MfmaTraits<T>::pair_t P2;
    const int t = threadIdx.x;
    if (LAYOUT == LAYOUT_COL) {
        const int r = 2 * (t & 63);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = (t >> 6) + 4 * u;
            panel[lds_index(r, k)] = R.v[2 * u];
            panel[lds_index(r + 1, k)] = R.v[2 * u + 1];
        }
    } else {
        const int r = t >> 1;
        const int kb = 8 * (t & 1);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            P2 v;
            v.x = R.v[2 * u];
            v.y = R.v[2 * u + 1];
            *reinterpret_cast<P2*>(panel + lds_index(r, kb + 2 * u)) = v;
        }
    }
}

// t += F(f0 .. f0+127, :) * S(s0 .. s0+127, :)^T restricted to this wave's 64 x 64 block
// (t -= ... with SUB: the first operand's fragments are negated, which is exact):
// wave w owns first-operand rows f0 + 64 (w & 1) and second-operand rows s0 + 64 (w >> 1).
// `lds` holds WG_LDS_ELEMS elements. All 256 lanes must call this (it has barriers).
template <typename T, int LF, int LS, bool SUB = false>
__device__ __forceinline__ void wg_tile_mma_lds(WaveTile<T>& t, T* __restrict__ lds,
                                                const Operand<T, LF>& F,
                                                const Operand<T, LS>& S, int64_t f0, int64_t s0,
                                                int64_t kdim) {
    typedef MfmaTraits<T> M;
    typedef typename M::pair_t P2;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r = lane & 15;
    const int g = lane >> 4;
    const int fr = 64 * (w & 1);
    const int sr = 64 * (w >> 1);
    const bool vf = operand_vec_ok(F), vs = operand_vec_ok(S);
    if (kdim <= 0) return;
    PanelRegs<T> pf, ps;
    panel_load(pf, F, f0, 0, vf);
    panel_load(ps, S, s0, 0, vs);
    panel_store<T, LF>(lds, pf);
    panel_store<T, LS>(lds + WG_PANEL, ps);
    __syncthreads();
    int stage = 0;
    for (int64_t k0 = 0; k0 < kdim; k0 += WG_KC) {
        const bool more = k0 + WG_KC < kdim;
        if (more) {
            panel_load(pf, F, f0, k0 + WG_KC, vf);
            panel_load(ps, S, s0, k0 + WG_KC, vs);
        }
        const T* Fp = lds + stage * 2 * WG_PANEL;
        const T* Sp = Fp + WG_PANEL;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            P2 fa[4], fb[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                fa[q] = *reinterpret_cast<const P2*>(Fp + lds_index(fr + 16 * q + r, 4 * g + 2 * h));
                fb[q] = *reinterpret_cast<const P2*>(Sp + lds_index(sr + 16 * q + r, 4 * g + 2 * h));
                if (SUB) {
                    fa[q].x = -fa[q].x;
                    fa[q].y = -fa[q].y;
                }
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) t.acc[a][b] = M::mma(fa[a].x, fb[b].x, t.acc[a][b]);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) t.acc[a][b] = M::mma(fa[a].y, fb[b].y, t.acc[a][b]);
        }
        if (more) {
            T* Fn = lds + (stage ^ 1) * 2 * WG_PANEL;
            panel_store<T, LF>(Fn, pf);
            panel_store<T, LS>(Fn + WG_PANEL, ps);
        }
        __syncthreads();
        stage ^= 1;
    }
}

// ------------------------------------------------------------------------------------------
// Accumulators <-> a column-major C block: element e of block (a, b) of this lane is
// C(row0 + 16 b + (lane & 15), col0 + tile_row(a, e)) (first operand = output columns).
// `live`: the wave's block has entries to update; `interior`: every entry of its 64 x 64 block
// is inside the M x N matrix (and on or below the diagonal when LOWER), so no masks are needed
// and all loads are issued back to back. Masked entries load as zero and are not stored.
// ------------------------------------------------------------------------------------------
template <typename T, bool LOWER>
__device__ __forceinline__ void wave_tile_load_c(WaveTile<T>& t, const T* __restrict__ C,
                                                 int64_t ldc, int64_t row0, int64_t col0,
                                                 int64_t M, int64_t N, bool live, bool interior) {
    const int lane = threadIdx.x & 63;
    if (interior) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    t.acc[a][b][e] = C[(row0 + 16 * b + (lane & 15)) + (col0 + tile_row<T>(a, e)) * ldc];
        return;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = row0 + 16 * b + (lane & 15);
                const int64_t j = col0 + tile_row<T>(a, e);
                const bool ok = live && i < M && j < N && (!LOWER || j <= i);
                t.acc[a][b][e] = ok ? C[i + j * ldc] : T(0);
            }
}

template <typename T, bool LOWER>
__device__ __forceinline__ void wave_tile_store_c(const WaveTile<T>& t, T* __restrict__ C,
                                                  int64_t ldc, int64_t row0, int64_t col0,
                                                  int64_t M, int64_t N, bool interior) {
    const int lane = threadIdx.x & 63;
    if (interior) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    C[(row0 + 16 * b + (lane & 15)) + (col0 + tile_row<T>(a, e)) * ldc] = t.acc[a][b][e];
        return;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t i = row0 + 16 * b + (lane & 15);
            if (i >= M) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t j = col0 + tile_row<T>(a, e);
                if (j >= N || (LOWER && j > i)) continue;
                C[i + j * ldc] = t.acc[a][b][e];
            }
        }
}

// ------------------------------------------------------------------------------------------
// Wave-level FP64 tile read straight from global memory (no LDS, no barriers): used where a
// workgroup's waves work on different operand rows (the fused predict). Lane (r, g) reads the
// four features k0 + 4g .. k0 + 4g + 3 of its row as two 16-byte loads (ROW layout only).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void load_row_frag4(double f[4], const Operand<double, LAYOUT_ROW>& op,
                                               int64_t row, int64_t k0, bool full) {
    const int64_t r = row < op.rows ? row : op.rows - 1;
    const double* src = op.p + r * op.ld + k0;
    if (full) {
        const double2 x = reinterpret_cast<const double2*>(src)[0];
        const double2 y = reinterpret_cast<const double2*>(src)[1];
        f[0] = x.x; f[1] = x.y; f[2] = y.x; f[3] = y.y;
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) f[s] = (k0 + s < op.kdim) ? src[s] : 0.0;
    }
}

__device__ __forceinline__ void wave_tile_mma_rows(WaveTile<double>& t,
                                                   const Operand<double, LAYOUT_ROW>& A,
                                                   const Operand<double, LAYOUT_ROW>& B,
                                                   int64_t row0, int64_t col0, int64_t kdim) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15;
    const int g = lane >> 4;
    if (kdim <= 0) return;
    double fa[4][4], fb[4][4];
    {
        const bool full = 16 <= kdim;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            load_row_frag4(fa[q], A, row0 + 16 * q + r, 4 * g, full);
            load_row_frag4(fb[q], B, col0 + 16 * q + r, 4 * g, full);
        }
    }
    for (int64_t k0 = 0; k0 < kdim; k0 += 16) {
        double na[4][4], nb[4][4];
        const int64_t kn = k0 + 16;
        const bool more = kn < kdim;
        if (more) {
            const bool full = kn + 16 <= kdim;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                load_row_frag4(na[q], A, row0 + 16 * q + r, kn + 4 * g, full);
                load_row_frag4(nb[q], B, col0 + 16 * q + r, kn + 4 * g, full);
            }
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    t.acc[a][b] = MfmaTraits<double>::mma(fa[a][s], fb[b][s], t.acc[a][b]);
        if (more) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    fa[q][s] = na[q][s];
                    fb[q][s] = nb[q][s];
                }
        }
    }
}

}  // namespace qmlh
