// Blocked right-looking Cholesky (lower, column-major) and the two triangular solves.
//
// Replaces LAPACK dpotrf("U") / dpotrs("U", n, 1) as called at qml/fcho_solve.f90:35,46 (see
// the header for why "U" on the Fortran view is the lower triangle here).
//
// Factorisation, outer block NB = 256, per step k:
//   1. diagonal block (kb x kb): itself blocked by 64 -- potf2 in LDS (one workgroup, FP64),
//      a row-parallel triangular solve for the rows below, an MFMA update of the rest;
//   2. panel A[k+kb:n, k:k+kb] := A[k+kb:n, k:k+kb] L_kk^-T: recursive halving down to
//      64-wide row-parallel substitution, MFMA updates in between;
//   3. trailing update A[k+kb:n, k+kb:n] -= P P^T on the lower tiles only (FP64 MFMA).
// A device int32 `info` follows LAPACK: the first non-positive (or NaN) pivot j sets
// info = j + 1, its Schur-complement value is left on the diagonal, and every later kernel of
// the factorisation returns without work.
#include "mfma_f64.h"
#include "timer.h"

namespace qmlh {

static constexpr int PF_NB = 256;
static constexpr int PF_IB = 64;
static constexpr int UP_TILE = 128;

// ------------------------------------------------------------------------------------------
// C(i, j) += alpha * sum_k A(i, k) B(j, k); A, B, C column-major. LOWER: only tiles with
// bi >= bj run and only i >= j is written (a SYRK when A == B).
// ------------------------------------------------------------------------------------------
template <bool LOWER>
__global__ __launch_bounds__(256) void gemm_update_kernel(
    const double* __restrict__ A, int64_t lda, const double* __restrict__ B, int64_t ldb,
    int64_t M, int64_t N, int64_t Kd, double alpha, double* __restrict__ C, int64_t ldc,
    const int32_t* __restrict__ info, int64_t tiles_m) {
    if (info && *info) return;
    int64_t bi, bj;
    const int64_t bid = blockIdx.x;
    if (LOWER) {
        int64_t i = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > bid) --i;
        while ((i + 1) * (i + 2) / 2 <= bid) ++i;
        bi = i;
        bj = bid - i * (i + 1) / 2;
    } else {
        bi = bid % tiles_m;
        bj = bid / tiles_m;
    }
    const int w = threadIdx.x >> 6;
    const int64_t row0 = bi * UP_TILE + 64 * (w & 1);
    const int64_t col0 = bj * UP_TILE + 64 * (w >> 1);
    if (row0 >= M || col0 >= N) return;
    if (LOWER && col0 > row0 + 63) return;

    Operand<LAYOUT_COL> opA{A, lda, M, Kd};
    Operand<LAYOUT_COL> opB{B, ldb, N, Kd};
    WaveTile t;
    wave_tile_zero(t);
    wave_tile_mma(t, opB, opA, col0, row0, Kd);

    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t i = row0 + 16 * b + (lane & 15);
            if (i >= M) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t j = col0 + tile_row(a, e);
                if (j >= N || (LOWER && j > i)) continue;
                double* c = C + i + j * ldc;
                *c = *c + alpha * t.acc[a][b][e];
            }
        }
    }
}

static int launch_update(bool lower, const double* A, int64_t lda, const double* B, int64_t ldb,
                         int64_t M, int64_t N, int64_t Kd, double alpha, double* C, int64_t ldc,
                         const int32_t* info, hipStream_t s) {
    if (M <= 0 || N <= 0 || Kd <= 0) return QMLH_OK;
    int64_t tm = ceil_div(M, UP_TILE), tn = ceil_div(N, UP_TILE);
    TimerScope ts(lower ? QMLH_TIMER_SYRK : QMLH_TIMER_GEMM,
                  lower ? (double)M * (double)(M + 1) * (double)Kd
                        : 2.0 * (double)M * (double)N * (double)Kd,
                  s);
    if (lower) {
        int64_t nt = tm * (tm + 1) / 2;
        hipLaunchKernelGGL(gemm_update_kernel<true>, dim3((unsigned)nt), dim3(256), 0, s, A, lda,
                           B, ldb, M, N, Kd, alpha, C, ldc, info, tm);
    } else {
        hipLaunchKernelGGL(gemm_update_kernel<false>, dim3((unsigned)(tm * tn)), dim3(256), 0, s,
                           A, lda, B, ldb, M, N, Kd, alpha, C, ldc, info, tm);
    }
    QMLH_CHECK_LAUNCH("gemm_update launch");
    return QMLH_OK;
}

// ------------------------------------------------------------------------------------------
// Unblocked right-looking Cholesky of a w x w (w <= 64) diagonal block held in LDS.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void potf2_kernel(double* __restrict__ A, int64_t lda, int w,
                                                    int64_t col_base, int32_t* __restrict__ info) {
    if (*info) return;
    __shared__ double s[PF_IB][PF_IB + 1];
    const int t = threadIdx.x;
    for (int idx = t; idx < w * w; idx += blockDim.x) {
        int i = idx % w, j = idx / w;
        if (i >= j) s[i][j] = A[i + j * lda];
    }
    __syncthreads();
    int fail = 0;
    for (int c = 0; c < w; ++c) {
        const double ajj = s[c][c];
        if (!(ajj > 0.0)) {  // non-positive or NaN pivot: LAPACK stops here
            fail = c + 1;
            break;
        }
        const double l = sqrt(ajj);
        __syncthreads();
        if (t == 0) s[c][c] = l;
        for (int i = c + 1 + t; i < w; i += blockDim.x) s[i][c] = s[i][c] / l;
        __syncthreads();
        const int m = w - c - 1;  // trailing order
        for (int idx = t; idx < m * m; idx += blockDim.x) {
            int i = c + 1 + idx % m, j = c + 1 + idx / m;
            if (i >= j) s[i][j] = s[i][j] - s[i][c] * s[j][c];
        }
        __syncthreads();
    }
    for (int idx = t; idx < w * w; idx += blockDim.x) {
        int i = idx % w, j = idx / w;
        if (i >= j) A[i + j * lda] = s[i][j];
    }
    if (fail && t == 0) *info = (int32_t)(col_base + fail);
}

// ------------------------------------------------------------------------------------------
// Row-parallel triangular solve X := B L^-T for a panel with w <= 64 columns: every lane
// owns one row and runs the substitution in registers; L sits in LDS.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void trsm_base_kernel(double* __restrict__ B, int64_t ldb,
                                                        int64_t m, const double* __restrict__ L,
                                                        int64_t ldl, int w,
                                                        const int32_t* __restrict__ info) {
    if (info && *info) return;
    __shared__ double sL[PF_IB][PF_IB + 1];
    __shared__ double sInv[PF_IB];
    for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
        int i = idx % w, j = idx / w;
        if (i >= j) sL[i][j] = L[i + j * ldl];
    }
    __syncthreads();
    if (threadIdx.x < w) sInv[threadIdx.x] = 1.0 / sL[threadIdx.x][threadIdx.x];
    __syncthreads();
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m) return;
    double x[PF_IB];
#pragma unroll
    for (int c = 0; c < PF_IB; ++c) {
        if (c < w) {
            double s = B[row + c * ldb];
#pragma unroll
            for (int l = 0; l < c; ++l) s = s - x[l] * sL[c][l];
            x[c] = s * sInv[c];
            B[row + c * ldb] = x[c];
        }
    }
}

static int launch_trsm_base(double* B, int64_t ldb, int64_t m, const double* L, int64_t ldl,
                            int w, const int32_t* info, hipStream_t s) {
    if (m <= 0) return QMLH_OK;
    hipLaunchKernelGGL(trsm_base_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, B,
                       ldb, m, L, ldl, w, info);
    QMLH_CHECK_LAUNCH("trsm_base launch");
    return QMLH_OK;
}

// X := B L^-T, B (m x w) and L (w x w, lower) column-major, by recursive halving.
static int trsm_panel(double* B, int64_t ldb, int64_t m, const double* L, int64_t ldl, int64_t w,
                      const int32_t* info, hipStream_t s) {
    if (w <= PF_IB) return launch_trsm_base(B, ldb, m, L, ldl, (int)w, info, s);
    int64_t h = round_up(w / 2, PF_IB);
    QMLH_TRY(trsm_panel(B, ldb, m, L, ldl, h, info, s));
    // B2 -= X1 L21^T
    QMLH_TRY(launch_update(false, B, ldb, L + h, ldl, m, w - h, h, -1.0, B + h * ldb, ldb, info,
                           s));
    return trsm_panel(B + h * ldb, ldb, m, L + h + h * ldl, ldl, w - h, info, s);
}

// Factor a kb x kb (kb <= NB) diagonal block: 64-wide potf2 / trsm / syrk steps.
static int factor_diag(double* A, int64_t lda, int64_t kb, int64_t col_base, int32_t* info,
                       hipStream_t s) {
    for (int64_t kk = 0; kk < kb; kk += PF_IB) {
        int w = (int)(kb - kk < PF_IB ? kb - kk : PF_IB);
        double* Akk = A + kk + kk * lda;
        hipLaunchKernelGGL(potf2_kernel, dim3(1), dim3(256), 0, s, Akk, lda, w, col_base + kk,
                           info);
        QMLH_CHECK_LAUNCH("potf2 launch");
        int64_t rest = kb - kk - w;
        if (rest > 0) {
            QMLH_TRY(launch_trsm_base(Akk + w, lda, rest, Akk, lda, w, info, s));
            QMLH_TRY(launch_update(true, Akk + w, lda, Akk + w, lda, rest, rest, w, -1.0,
                                   Akk + w + w * lda, lda, info, s));
        }
    }
    return QMLH_OK;
}

// ------------------------------------------------------------------------------------------
// Triangular solves for potrs, blocks of 64 rows.
// ------------------------------------------------------------------------------------------
// x_blk := L_blk^-1 x_blk (TRANS = false) or L_blk^-T x_blk (TRANS = true); one wavefront,
// lane i holds x_i.
template <bool TRANS>
__global__ __launch_bounds__(64) void trsv_block_kernel(const double* __restrict__ L,
                                                        int64_t ldl, int w,
                                                        double* __restrict__ x) {
    const int i = threadIdx.x;
    double xi = (i < w) ? x[i] : 0.0;
    if (!TRANS) {
        for (int c = 0; c < w; ++c) {
            double xc = __shfl(xi / L[c + c * ldl], c, 64);
            if (i == c) xi = xc;
            if (i > c && i < w) xi = xi - L[i + c * ldl] * xc;
        }
    } else {
        for (int c = w - 1; c >= 0; --c) {
            double xc = __shfl(xi / L[c + c * ldl], c, 64);
            if (i == c) xi = xc;
            if (i < c) xi = xi - L[c + i * ldl] * xc;
        }
    }
    if (i < w) x[i] = xi;
}

// y[r] -= sum_c L(r, c) xb[c] for r < m, c < w (forward substitution update).
__global__ __launch_bounds__(256) void gemv_fwd_kernel(const double* __restrict__ L, int64_t ldl,
                                                       int64_t m, int w,
                                                       const double* __restrict__ xb,
                                                       double* __restrict__ y) {
    __shared__ double sx[PF_IB];
    if (threadIdx.x < w) sx[threadIdx.x] = xb[threadIdx.x];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    double s = 0.0;
    for (int c = 0; c < w; ++c) s = fma(L[r + c * ldl], sx[c], s);
    y[r] = y[r] - s;
}

// xb[c] -= sum_r L(r, c) y[r] for c < w, r < m: one workgroup per column c.
__global__ __launch_bounds__(256) void gemv_bwd_kernel(const double* __restrict__ L, int64_t ldl,
                                                       int64_t m, const double* __restrict__ y,
                                                       double* __restrict__ xb) {
    __shared__ double red[256];
    const int c = blockIdx.x;
    const double* col = L + c * ldl;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < m; r += 256) s = fma(col[r], y[r], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) xb[c] = xb[c] - red[0];
}

__global__ void clear_info_kernel(int32_t* info) { *info = 0; }

__global__ void add_diagonal_kernel(double* A, int64_t n, int64_t lda, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i + i * lda] += v;
}

}  // namespace qmlh

extern "C" int qmlh_potrf(double* A, int64_t n, int64_t lda, int32_t* info,
                          qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0) return set_error(QMLH_ERR_ARG, "potrf: n < 0");
    if (lda < (n > 1 ? n : 1)) return set_error(QMLH_ERR_ARG, "potrf: lda < n");
    if (!info) return set_error(QMLH_ERR_ARG, "potrf: info is null");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(clear_info_kernel, dim3(1), dim3(1), 0, s, info);
    QMLH_CHECK_LAUNCH("potrf: clear info");
    if (n == 0) return QMLH_OK;
    if (!A) return set_error(QMLH_ERR_ARG, "potrf: A is null");
    TimerScope ts(QMLH_TIMER_POTRF, (double)n * (double)n * (double)n / 3.0, s);
    for (int64_t k = 0; k < n; k += PF_NB) {
        int64_t kb = n - k < PF_NB ? n - k : PF_NB;
        double* Akk = A + k + k * lda;
        QMLH_TRY(factor_diag(Akk, lda, kb, k, info, s));
        int64_t m = n - k - kb;
        if (m > 0) {
            double* P = Akk + kb;
            QMLH_TRY(trsm_panel(P, lda, m, Akk, lda, kb, info, s));
            QMLH_TRY(launch_update(true, P, lda, P, lda, m, m, kb, -1.0, P + kb * lda, lda, info,
                                   s));
        }
    }
    return QMLH_OK;
}

extern "C" int qmlh_potrs(const double* L, int64_t n, int64_t lda, double* x,
                          qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0) return set_error(QMLH_ERR_ARG, "potrs: n < 0");
    if (lda < (n > 1 ? n : 1)) return set_error(QMLH_ERR_ARG, "potrs: lda < n");
    if (n == 0) return QMLH_OK;
    if (!L || !x) return set_error(QMLH_ERR_ARG, "potrs: null pointer");
    hipStream_t s = (hipStream_t)stream;
    TimerScope ts(QMLH_TIMER_POTRS, 2.0 * (double)n * (double)n, s);
    // Forward: L z = y.
    for (int64_t k = 0; k < n; k += PF_IB) {
        int w = (int)(n - k < PF_IB ? n - k : PF_IB);
        const double* Lkk = L + k + k * lda;
        hipLaunchKernelGGL(trsv_block_kernel<false>, dim3(1), dim3(64), 0, s, Lkk, lda, w, x + k);
        QMLH_CHECK_LAUNCH("trsv launch");
        int64_t m = n - k - w;
        if (m > 0) {
            hipLaunchKernelGGL(gemv_fwd_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s,
                               Lkk + w, lda, m, w, x + k, x + k + w);
            QMLH_CHECK_LAUNCH("gemv_fwd launch");
        }
    }
    // Backward: L^T x = z.
    int64_t last = ((n - 1) / PF_IB) * PF_IB;
    for (int64_t k = last; k >= 0; k -= PF_IB) {
        int w = (int)(n - k < PF_IB ? n - k : PF_IB);
        const double* Lkk = L + k + k * lda;
        int64_t m = n - k - w;
        if (m > 0) {
            hipLaunchKernelGGL(gemv_bwd_kernel, dim3((unsigned)w), dim3(256), 0, s, Lkk + w, lda,
                               m, x + k + w, x + k);
            QMLH_CHECK_LAUNCH("gemv_bwd launch");
        }
        hipLaunchKernelGGL(trsv_block_kernel<true>, dim3(1), dim3(64), 0, s, Lkk, lda, w, x + k);
        QMLH_CHECK_LAUNCH("trsv launch");
    }
    return QMLH_OK;
}

extern "C" int qmlh_add_diagonal(double* A, int64_t n, int64_t lda, double value,
                                 qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || lda < n) return set_error(QMLH_ERR_ARG, "add_diagonal: bad shape");
    if (n == 0) return QMLH_OK;
    if (!A) return set_error(QMLH_ERR_ARG, "add_diagonal: A is null");
    hipLaunchKernelGGL(add_diagonal_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, A, n, lda, value);
    QMLH_CHECK_LAUNCH("add_diagonal launch");
    return QMLH_OK;
}
