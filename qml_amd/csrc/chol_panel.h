// Cholesky panel building blocks shared by the single-matrix factorisation (potrf.hip) and the
// batched one (batched.hip): the factor of a <= 64 x 64 diagonal block in LDS, and the
// row-parallel solve of the panel rows below it.
#pragma once

#include "common.h"

namespace qmlh {

static constexpr int CP_B = 64;  // diagonal block / sub-panel width

// Lane `lane`'s value of v in every lane (two scalar reads; lane is a compile-time constant
// where this is used, so no LDS round trip).
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Packed lower triangle of a 64 x 64 block in LDS, column-major: element (i, j), i >= j, at
// pk(i, j); column j holds rows j .. 63. 2080 doubles (16.6 KB) instead of a padded square
// (33 KB), so the panel kernels fit in the LDS the trailing-update workgroups leave free and run
// beside them. Wherever the loops are unrolled the index is a compile-time constant (+ lane).
static constexpr int CP_PACK = CP_B * (CP_B + 1) / 2;
__host__ __device__ constexpr int pk(int i, int j) { return j * CP_B - (j * (j - 1)) / 2 + (i - j); }

// Every lane loads unconditionally from an index clamped into the w x w block and selects
// zero afterwards: a load (and its conversion) under a condition makes the compiler wait for
// it inside the branch, which serialises a run of loads.
__device__ __forceinline__ int cp_clamp(int i, int w) { return i < w ? i : w - 1; }

// Packed LDS <- the lower triangle of the w x w block at A (zero outside w); 256 lanes, loads
// issued back to back.
template <typename T>
__device__ __forceinline__ void stage_packed64(double* __restrict__ sp, const T* __restrict__ A,
                                               int64_t lda, int w) {
    const int t = threadIdx.x;
    T raw[CP_B * CP_B / 256];
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        raw[q] = A[cp_clamp(i, w) + (int64_t)cp_clamp(j, w) * lda];
    }
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        if (i >= j) sp[pk(i, j)] = (i < w && j < w) ? (double)raw[q] : 0.0;
    }
}

// LDS (FP64) <- the lower triangle of the w x w block at A (column-major, lda), zero elsewhere;
// TRANS stores element (i, j) at s[j][i]. All 256 lanes take part; each issues its 16 loads
// back to back before any LDS store.
template <bool TRANS, int LDS_LD, typename T>
__device__ __forceinline__ void stage_lower64(double (*s)[LDS_LD], const T* __restrict__ A,
                                              int64_t lda, int w) {
    const int t = threadIdx.x;
    T raw[CP_B * CP_B / 256];
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;  // consecutive lanes: consecutive rows
        raw[q] = A[cp_clamp(i, w) + (int64_t)cp_clamp(j, w) * lda];
    }
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        const double v = (i < w && j < w && i >= j) ? (double)raw[q] : 0.0;
        if (TRANS) s[j][i] = v; else s[i][j] = v;
    }
}

// Cholesky of the w x w (w <= 64) lower block in packed LDS sp (FP64) by one wavefront, all
// in registers: lane i holds row i. Per column c: the pivot is read from lane c, lanes below
// scale their entry, the column goes through a 64-entry LDS buffer, and every lane subtracts
// L(i, c) L(j, c) from its entries j > c (independent FMAs; entries above the diagonal are
// scratch). Called by all threads of the workgroup (barriers); waves 1.. only wait. On exit sp
// holds L in its lower triangle. Returns 0, or c + 1 for the first non-positive (or NaN) pivot
// c, with the Schur value left on that diagonal entry (as LAPACK leaves it); the value is the
// same in every lane. floor > 0 (static pivoting, for factors that only precondition): a
// finite pivot at or below floor is replaced by floor instead of failing; NaN still fails.
// fail_at >= 0: a pivot at or below it fails (0: LAPACK's rule).
__device__ __forceinline__ int potf2_wave64(double* __restrict__ sp, int w, double floor = 0.0,
                                            double fail_at = 0.0) {
    __shared__ double colbuf[CP_B];
    __shared__ int s_fail;
    const int t = threadIdx.x;
    if (t < CP_B) {
        const int lane = t;
        double a[CP_B];
#pragma unroll
        for (int j = 0; j < CP_B; ++j) {
            const double v = sp[pk(lane > j ? lane : j, j)];  // in range for every lane
            a[j] = (lane >= j && lane < w) ? v : 0.0;
        }
        int f = 0;
#pragma unroll
        for (int c = 0; c < CP_B; ++c) {
            if (c >= w || f) break;
            double piv = readlane_f64(a[c], c);
            if (!(piv > floor) && floor > 0.0 && piv == piv) piv = floor;
            if (!(piv > fail_at)) {
                f = c + 1;
                break;
            }
            // 1/sqrt(piv) from the hardware estimate and two Newton steps (each squares the
            // relative error; about one unit in the last place after the second), then
            // sqrt(piv) = piv / sqrt(piv): five dependent operations on the critical path
            // instead of a correctly rounded square root followed by a division.
            double rl = __builtin_amdgcn_rsq(piv);
            const double hp = 0.5 * piv;
            rl = rl * fma(-hp * rl, rl, 1.5);
            rl = rl * fma(-hp * rl, rl, 1.5);
            const double l = piv * rl;
            a[c] = lane == c ? l : (lane > c ? a[c] * rl : a[c]);
            if (c + 1 < CP_B) {
                colbuf[lane] = a[c];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int j = c + 1; j < CP_B; ++j) a[j] = fma(-a[c], colbuf[j], a[j]);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();  // every lane has read before the next write
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        if (lane < w) {
#pragma unroll
            for (int j = 0; j < CP_B; ++j)
                if (j <= lane) sp[pk(lane, j)] = a[j];
        }
        if (lane == 0) s_fail = f;
    }
    __syncthreads();
    return s_fail;
}

// Row `row` of the m x w panel B (column-major, ldb) := B L^-T with L packed by rows in sp
// and the reciprocal pivots in inv; FP64 arithmetic, storage T. The row is loaded whole first
// (64 loads in flight together), solved right-looking (below), and stored. Addresses are a
// uniform column pointer plus the lane's 32-bit byte offset (rows below 2^31 / 8).
template <typename T>
__device__ __forceinline__ const T& at_byte(const T* base, uint32_t off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ T& at_byte(T* base, uint32_t off) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}

// Row-major packed lower triangle (row i holds columns 0 .. i at rp(i, 0) ..): the 8 entries
// L(j, c .. c + 7) a lane needs together are adjacent.
__host__ __device__ constexpr int rp(int i, int j) { return i * (i + 1) / 2 + j; }

// LDS (row-major packed) <- the lower triangle of the w x w block at A, zero outside w.
template <typename T>
__device__ __forceinline__ void stage_packed64_rows(double* __restrict__ sp,
                                                    const T* __restrict__ A, int64_t lda, int w) {
    const int t = threadIdx.x;
    T raw[CP_B * CP_B / 256];
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        raw[q] = A[cp_clamp(i, w) + (int64_t)cp_clamp(j, w) * lda];
    }
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        if (i >= j) sp[rp(i, j)] = (i < w && j < w) ? (double)raw[q] : 0.0;
    }
}

// The substitution runs in blocks of 8 columns: the 8 columns of a block are finished among
// themselves, then every later entry takes the block's 8 contributions at once from 8
// adjacent LDS values -- independent per entry, so the reads stay in flight together instead
// of one or two at a time. Per entry the updates come in the same order as column by column.
template <typename T>
__device__ __forceinline__ void trsm_row(T* __restrict__ B, int64_t ldb, uint32_t row, int w,
                                         const double* __restrict__ sp,
                                         const double* __restrict__ inv) {
    constexpr int CB = 8;
    const uint32_t off = row * (uint32_t)sizeof(T);
    T raw[CP_B];
#pragma unroll
    for (int c = 0; c < CP_B; ++c) raw[c] = at_byte<T>(B + cp_clamp(c, w) * ldb, off);
    double x[CP_B];
#pragma unroll
    for (int c = 0; c < CP_B; ++c) x[c] = c < w ? (double)raw[c] : 0.0;
#pragma unroll
    for (int cb = 0; cb < CP_B; cb += CB) {
        if (cb < w) {  // inv and sp are zero past w, which keeps x zero there
#pragma unroll
            for (int c = cb; c < cb + CB; ++c) {
                x[c] = x[c] * inv[c];
#pragma unroll
                for (int j = c + 1; j < cb + CB; ++j) x[j] = fma(-x[c], sp[rp(j, c)], x[j]);
            }
            // four entries at a time: their 32 LDS values loaded together, then four
            // independent FMA chains interleaved
#pragma unroll
            for (int j = cb + CB; j < CP_B; j += 4) {
                double l[4][CB];
#pragma unroll
                for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                    for (int c = 0; c < CB; ++c) l[jj][c] = sp[rp(j + jj, cb + c)];
#pragma unroll
                for (int c = 0; c < CB; ++c)
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj)
                        x[j + jj] = fma(-x[cb + c], l[jj][c], x[j + jj]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < CP_B; ++c)
        if (c < w) at_byte<T>(B + c * ldb, off) = (T)x[c];
}

// Diagonal block of a sub-panel, one workgroup: factor the w x w block at A in wave 0, write
// L back, set info = col_base + c + 1 on a failed pivot. `info` already set: nothing is done.
// sp: CP_PACK doubles of LDS (potf2_body brings its own; a kernel that has LDS to spare at
// that point passes it).
// Shift on breakdown (shift_acc non-null, bshift > 0; for factors that only precondition): a
// pivot at or below bpivot counts as failed, the block's diagonal first gets *shift_acc (the
// shift earlier blocks settled on), a failed pivot re-factors the block from A with its whole
// diagonal raised by bshift, then 10 bshift (no further: a larger shift would precondition a
// different matrix), and the shift that succeeds is left in *shift_acc for the later blocks. The factor is then the exact Cholesky factor of A + diag(0, .., 0, s1, .., s2, ..):
// zero before the first failing block, nondecreasing after. A block that fails at 10 bshift
// sets info. Diagonal blocks are factored one at a time in column order, each after the
// one before it, so *shift_acc needs no atomics.
template <typename T>
__device__ __forceinline__ void potf2_body_sp(T* __restrict__ A, int64_t lda, int w,
                                              int64_t col_base, int32_t* __restrict__ info,
                                              double* __restrict__ sp, double floor = 0.0,
                                              double bshift = 0.0,
                                              double* __restrict__ shift_acc = nullptr,
                                              double bpivot = 0.0) {
    if (*info) return;
    const int t = threadIdx.x;
    const bool shifting = shift_acc != nullptr && bshift > 0.0;
    double shift = shifting ? *shift_acc : 0.0;
    int fail = 0;
    for (int attempt = 0;; ++attempt) {
        stage_packed64(sp, A, lda, w);
        __syncthreads();
        if (shift > 0.0 && t < w) sp[pk(t, t)] += shift;
        __syncthreads();
        fail = potf2_wave64(sp, w, floor, shifting ? bpivot : 0.0);
        if (!fail || !shifting || shift >= 10.0 * bshift) break;
        shift = shift > 0.0 ? 10.0 * shift : bshift;  // uniform: every lane sees the same fail
        __syncthreads();  // every lane has staged before sp is reloaded
    }
#pragma unroll
    for (int q = 0; q < CP_B * CP_B / 256; ++q) {
        const int idx = t + 256 * q;
        const int i = idx & (CP_B - 1), j = idx / CP_B;
        if (i < w && j < w && i >= j) A[i + (int64_t)j * lda] = (T)sp[pk(i, j)];
    }
    if (t == 0 && fail) *info = (int32_t)(col_base + fail);
    if (t == 0 && shifting && !fail) *shift_acc = shift;
}

template <typename T>
__device__ __forceinline__ void potf2_body(T* __restrict__ A, int64_t lda, int w,
                                           int64_t col_base, int32_t* __restrict__ info,
                                           double floor = 0.0, double bshift = 0.0,
                                           double* __restrict__ shift_acc = nullptr,
                                           double bpivot = 0.0) {
    __shared__ double sp[CP_PACK];
    potf2_body_sp<T>(A, lda, w, col_base, info, sp, floor, bshift, shift_acc, bpivot);
}

// Rows of a sub-panel, workgroup `local` of the launch: rows 256 local .. of the m rows below
// the factored w x w block L at A become B L^-T. Runs after potf2_body in stream order, so L
// is final when it is read. `info` set: nothing is done.
template <typename T>
__device__ __forceinline__ void trsm_body(T* __restrict__ A, int64_t lda, int w, int64_t m,
                                          const int32_t* __restrict__ info, int64_t local) {
    __shared__ double sp[CP_PACK];
    __shared__ double inv[CP_B];
    if (*info) return;
    const int t = threadIdx.x;
    stage_packed64_rows(sp, A, lda, w);
    __syncthreads();
    if (t < CP_B) inv[t] = t < w ? 1.0 / sp[rp(t, t)] : 0.0;
    __syncthreads();
    const int64_t row = local * 256 + t;
    if (row < m) trsm_row<T>(A + w, lda, (uint32_t)row, w, sp, inv);
}

}  // namespace qmlh
