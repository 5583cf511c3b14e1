// Gaussian kernel matrix in the expanded form on FP64 MFMA:
//   K(i, j) = exp(scale * max(|a_i|^2 + |b_j|^2 - 2 a_i.b_j, 0)),  scale = -0.5 / (sigma*sigma),
// the same quantity qml/fkernels.f90:340-372 computes with direct differences. The cross term
// is the one dense contraction of the path; norms and the exp argument stay FP64.
//
// A 256-lane workgroup owns a 128 x 128 output tile (2 x 2 wavefronts of 64 x 64) with both
// operands staged through LDS. The MFMA is fed (B rows, A rows) so that the 16 lanes holding
// one accumulator element run down 16 consecutive output rows: every store is a 128-byte run
// of the column-major K. In symmetric mode only tiles with bi <= bj run; a diagonal tile
// writes i <= j and mirrors, an off-diagonal tile writes both (i, j) and (j, i), so K is
// exactly symmetric with unit diagonal.
//
// The deflated build (mixed-precision KRR) writes, from the same cross term c = x_i.x_j:
//   K64(i, j) = exp(scale * max(n_i + n_j - 2c, 0)) + lambda [i == j]             (lower, FP64)
//   B32(i, j) = u_i u_j (expm1(c / s2) - c / s2) + (lambda + shift) [i == j]     (lower, FP32)
//   V(i, :)   = u_i [1, x_i / sigma]                                  (n x (d+1), FP32)
// with s2 = sigma^2 and u_i = exp(scale * n_i). Since exp(c/s2) = 1 + c/s2 + (expm1(c/s2) -
// c/s2), K = V V^T + B: a rank-(d+1) term (the constant and linear parts of the expansion)
// plus a positive semidefinite remainder (the Hadamard powers >= 2 of the Gram matrix / s2),
// whose entries are O((|x|^2 / s2)^2) -- about 1e-3 for Coulomb matrices at sigma = 4000,
// against O(1) for K. A low-precision Cholesky of B + lambda I stays accurate relative to
// lambda where one of K + lambda I does not, and V is kept exact in the preconditioner.
#include "mfma_core.h"
#include "timer.h"

namespace qmlh {

static constexpr int GK_TILE = WG_T;

__global__ __launch_bounds__(256) void row_sqnorms_kernel(const double* __restrict__ X,
                                                          int64_t n, int64_t ld, int64_t d,
                                                          double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* x = X + i * ld;
    double s = 0.0;
    for (int64_t k = 0; k < d; ++k) s = fma(x[k], x[k], s);
    out[i] = s;
}

__global__ void exp_scale_kernel(const double* __restrict__ norms, int64_t n, double scale,
                                 double* __restrict__ u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = exp(scale * norms[i]);
}

// V(i, 0) = u_i, V(i, 1 + k) = u_i x_ik / sigma (column-major, FP32).
__global__ void lowrank_kernel(const double* __restrict__ X, int64_t ldx, int64_t n, int64_t d,
                               const double* __restrict__ u, double inv_sigma,
                               float* __restrict__ V, int64_t ldv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ui = u[i];
    V[i] = (float)ui;
    const double* x = X + i * ldx;
    for (int64_t k = 0; k < d; ++k) V[i + (k + 1) * ldv] = (float)(ui * x[k] * inv_sigma);
}

// V(i, 0) = u_i, V(i, 1 + k) = u_i x_ik / sigma (column-major, FP64): the rank-(d+1) term of
// the deflation at full precision, for preconditioners that keep it in FP64.
__global__ void lowrank64_kernel(const double* __restrict__ X, int64_t ldx, int64_t n, int64_t d,
                                 const double* __restrict__ u, double inv_sigma,
                                 double* __restrict__ V, int64_t ldv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ui = u[i];
    V[i] = ui;
    const double* x = X + i * ldx;
    for (int64_t k = 0; k < d; ++k) V[i + (k + 1) * ldv] = ui * x[k] * inv_sigma;
}

__device__ __forceinline__ void tile_coords(int64_t bid, int symmetric, int64_t tiles_m,
                                            int64_t& bi, int64_t& bj) {
    if (symmetric) {
        int64_t j = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
        while (j * (j + 1) / 2 > bid) --j;
        while ((j + 1) * (j + 2) / 2 <= bid) ++j;
        bj = j;
        bi = bid - j * (j + 1) / 2;
    } else {
        bi = bid % tiles_m;
        bj = bid / tiles_m;
    }
}

__global__ __launch_bounds__(256, 2) void gaussian_mfma_kernel(
    const double* __restrict__ A, int64_t lda, int64_t na, const double* __restrict__ nA,
    const double* __restrict__ B, int64_t ldb, int64_t nb, const double* __restrict__ nB,
    int64_t d, double scale, double* __restrict__ K, int64_t ldk, int symmetric,
    int64_t tiles_m, int lower) {
    __shared__ __attribute__((aligned(16))) double lds[WG_LDS_ELEMS];
    int64_t bi, bj;
    tile_coords(blockIdx.x, symmetric, tiles_m, bi, bj);
    if (lower) {  // the transposed tile: on or below the diagonal, written as is
        const int64_t tmp = bi;
        bi = bj;
        bj = tmp;
    }
    Operand<double, LAYOUT_ROW> opA{A, lda, na, d};
    Operand<double, LAYOUT_ROW> opB{B, ldb, nb, d};
    WaveTile<double> t;
    wave_tile_zero(t);
    // First operand indexes output columns (B), second output rows (A).
    wg_tile_mma_lds(t, lds, opB, opA, bj * GK_TILE, bi * GK_TILE, d);

    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t col0 = bj * GK_TILE + 64 * (w & 1);
    const int64_t row0 = bi * GK_TILE + 64 * (w >> 1);
    if (row0 >= na || col0 >= nb) return;
    const bool diag_tile = symmetric && bi == bj;
    const bool mirror = symmetric && !lower && bi != bj;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t i = row0 + 16 * b + (lane & 15);
            if (i >= na) continue;
            const double ni = nA[i];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t j = col0 + tile_row<double>(a, e);
                if (j >= nb) continue;
                if (diag_tile && (lower ? i < j : i > j)) continue;
                double d2 = ni + nB[j] - 2.0 * t.acc[a][b][e];
                d2 = d2 > 0.0 ? d2 : 0.0;
                if (symmetric && i == j) d2 = 0.0;
                const double v = exp(scale * d2);
                K[i + j * ldk] = v;
                if (mirror || (diag_tile && !lower && i < j)) K[j + i * ldk] = v;
            }
        }
    }
}

// Linear kernel K(i, j) = a_i . b_j (qml/fkernels.f90:403-424) from the same cross-term tiles;
// symmetric mode computes the upper tiles and mirrors.
__global__ __launch_bounds__(256, 2) void linear_mfma_kernel(
    const double* __restrict__ A, int64_t lda, int64_t na, const double* __restrict__ B,
    int64_t ldb, int64_t nb, int64_t d, double* __restrict__ K, int64_t ldk, int symmetric,
    int64_t tiles_m) {
    __shared__ __attribute__((aligned(16))) double lds[WG_LDS_ELEMS];
    int64_t bi, bj;
    tile_coords(blockIdx.x, symmetric, tiles_m, bi, bj);
    Operand<double, LAYOUT_ROW> opA{A, lda, na, d};
    Operand<double, LAYOUT_ROW> opB{B, ldb, nb, d};
    WaveTile<double> t;
    wave_tile_zero(t);
    wg_tile_mma_lds(t, lds, opB, opA, bj * GK_TILE, bi * GK_TILE, d);
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t col0 = bj * GK_TILE + 64 * (w & 1);
    const int64_t row0 = bi * GK_TILE + 64 * (w >> 1);
    if (row0 >= na || col0 >= nb) return;
    const bool diag_tile = symmetric && bi == bj;
    const bool mirror = symmetric && bi != bj;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t i = row0 + 16 * b + (lane & 15);
            if (i >= na) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t j = col0 + tile_row<double>(a, e);
                if (j >= nb) continue;
                if (diag_tile && i > j) continue;
                const double v = t.acc[a][b][e];
                K[i + j * ldk] = v;
                if (mirror || (diag_tile && i < j)) K[j + i * ldk] = v;
            }
        }
    }
}

__global__ __launch_bounds__(256, 2) void gaussian_deflated_kernel(
    const double* __restrict__ X, int64_t ldx, int64_t n, const double* __restrict__ nrm,
    const double* __restrict__ u, int64_t d, double scale, double lambda, double shift,
    double* __restrict__ K, int64_t ldk, float* __restrict__ B32, int64_t ldb) {
    __shared__ __attribute__((aligned(16))) double lds[WG_LDS_ELEMS];
    int64_t bi, bj;
    tile_coords(blockIdx.x, 1, 0, bi, bj);  // bi <= bj
    Operand<double, LAYOUT_ROW> op{X, ldx, n, d};
    WaveTile<double> t;
    wave_tile_zero(t);
    // First operand: block bi (the column index i of the lower-triangle entry), second: block
    // bj (its row index j), so the 16 lanes of an accumulator element run down 16 consecutive
    // rows j of the column-major outputs.
    wg_tile_mma_lds(t, lds, op, op, bi * GK_TILE, bj * GK_TILE, d);

    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t i0 = bi * GK_TILE + 64 * (w & 1);
    const int64_t j0 = bj * GK_TILE + 64 * (w >> 1);
    if (i0 >= n || j0 >= n || i0 > j0 + 63) return;
    const double inv_s2 = -2.0 * scale;  // 1 / sigma^2
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int64_t j = j0 + 16 * b + (lane & 15);  // row of the lower triangle
        if (j >= n) continue;
        const double nj = nrm[j], uj = u[j];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = i0 + tile_row<double>(a, e);  // column, i <= j
                if (i > j) continue;
                const double c = t.acc[a][b][e];
                double kv, bv;
                const double xv = c * inv_s2;
                // B'' = u_i u_j (expm1(c/s^2) - c/s^2) = k_ij - u_i u_j (1 + c/s^2), since
                // k_ij = u_i u_j exp(c/s^2): one exponential per entry. The subtraction loses
                // about three of FP64's sixteen digits (B'' ~ 1e-3 k), far below the FP32 it
                // is stored in.
                if (i == j) {
                    kv = 1.0 + lambda;
                    bv = (1.0 - uj * uj * (1.0 + xv)) + (lambda + shift);
                } else {
                    double d2 = nrm[i] + nj - 2.0 * c;
                    d2 = d2 > 0.0 ? d2 : 0.0;
                    kv = exp(scale * d2);
                    bv = kv - u[i] * uj * (1.0 + xv);
                }
                if (K) K[j + i * ldk] = kv;
                B32[j + i * ldb] = (float)bv;
            }
        }
    }
}

}  // namespace qmlh

extern "C" int qmlh_row_sqnorms(const double* X, int64_t n, int64_t ldx, int64_t d,
                                double* norms, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || d < 0 || ldx < d) return set_error(QMLH_ERR_ARG, "row_sqnorms: bad shape");
    if (n == 0) return QMLH_OK;
    if (!X || !norms) return set_error(QMLH_ERR_ARG, "row_sqnorms: null pointer");
    hipLaunchKernelGGL(row_sqnorms_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, X, n, ldx, d, norms);
    QMLH_CHECK_LAUNCH("row_sqnorms launch");
    return QMLH_OK;
}

extern "C" int qmlh_gaussian_mfma(const double* A, int64_t na, int64_t lda, const double* normsA,
                                  const double* B, int64_t nb, int64_t ldb,
                                  const double* normsB, int64_t d, double sigma, double* K,
                                  int64_t ldk, int32_t flags, qmlh_stream_t stream) {
    using namespace qmlh;
    if (na < 0 || nb < 0 || d < 0) return set_error(QMLH_ERR_ARG, "gaussian_mfma: negative size");
    if (lda % 4 || ldb % 4) return set_error(QMLH_ERR_ARG, "gaussian_mfma: lda, ldb must be multiples of 4");
    if (lda < d || ldb < d) return set_error(QMLH_ERR_ARG, "gaussian_mfma: ld < d");
    if (ldk < na) return set_error(QMLH_ERR_ARG, "gaussian_mfma: ldk < na");
    if ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15)
        return set_error(QMLH_ERR_ARG, "gaussian_mfma: A and B must be 16-byte aligned");
    bool sym = (flags & QMLH_SYMMETRIC) != 0;
    const bool lower = (flags & QMLH_LOWER) != 0;
    if (sym && na != nb) return set_error(QMLH_ERR_ARG, "gaussian_mfma: symmetric needs na == nb");
    if (lower && !sym) return set_error(QMLH_ERR_ARG, "gaussian_mfma: QMLH_LOWER needs QMLH_SYMMETRIC");
    if (na == 0 || nb == 0) return QMLH_OK;
    if (!A || !B || !K || !normsA || !normsB)
        return set_error(QMLH_ERR_ARG, "gaussian_mfma: null pointer");
    int64_t tm = ceil_div(na, GK_TILE), tn = ceil_div(nb, GK_TILE);
    int64_t ntiles = sym ? tm * (tm + 1) / 2 : tm * tn;
    if (ntiles > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "gaussian_mfma: grid too large");
    // Work actually performed: full tiles, or the upper tile triangle in symmetric mode.
    double work = 2.0 * (double)d * (sym ? (double)na * (double)(na + 1) / 2.0
                                         : (double)na * (double)nb);
    TimerScope ts(QMLH_TIMER_GAUSS, work, (hipStream_t)stream);
    hipLaunchKernelGGL(gaussian_mfma_kernel, dim3((unsigned)ntiles), dim3(256), 0,
                       (hipStream_t)stream, A, lda, na, normsA, B, ldb, nb, normsB, d,
                       -0.5 / (sigma * sigma), K, ldk, sym ? 1 : 0, tm, lower ? 1 : 0);
    QMLH_CHECK_LAUNCH("gaussian_mfma launch");
    return QMLH_OK;
}

extern "C" int qmlh_linear_mfma(const double* A, int64_t na, int64_t lda, const double* B,
                                int64_t nb, int64_t ldb, int64_t d, double* K, int64_t ldk,
                                int32_t flags, qmlh_stream_t stream) {
    using namespace qmlh;
    if (na < 0 || nb < 0 || d < 0) return set_error(QMLH_ERR_ARG, "linear_mfma: negative size");
    if (lda % 4 || ldb % 4) return set_error(QMLH_ERR_ARG, "linear_mfma: lda, ldb must be multiples of 4");
    if (lda < d || ldb < d) return set_error(QMLH_ERR_ARG, "linear_mfma: ld < d");
    if (ldk < na) return set_error(QMLH_ERR_ARG, "linear_mfma: ldk < na");
    if ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15)
        return set_error(QMLH_ERR_ARG, "linear_mfma: A and B must be 16-byte aligned");
    const bool sym = (flags & QMLH_SYMMETRIC) != 0;
    if (sym && na != nb) return set_error(QMLH_ERR_ARG, "linear_mfma: symmetric needs na == nb");
    if (na == 0 || nb == 0) return QMLH_OK;
    if (!A || !B || !K) return set_error(QMLH_ERR_ARG, "linear_mfma: null pointer");
    const int64_t tm = ceil_div(na, GK_TILE), tn = ceil_div(nb, GK_TILE);
    const int64_t ntiles = sym ? tm * (tm + 1) / 2 : tm * tn;
    if (ntiles > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "linear_mfma: grid too large");
    const double work = 2.0 * (double)d * (sym ? (double)na * (double)(na + 1) / 2.0
                                               : (double)na * (double)nb);
    TimerScope ts(QMLH_TIMER_GAUSS, work, (hipStream_t)stream);
    hipLaunchKernelGGL(linear_mfma_kernel, dim3((unsigned)ntiles), dim3(256), 0,
                       (hipStream_t)stream, A, lda, na, B, ldb, nb, d, K, ldk, sym ? 1 : 0, tm);
    QMLH_CHECK_LAUNCH("linear_mfma launch");
    return QMLH_OK;
}

extern "C" int qmlh_gaussian_deflated(const double* X, int64_t n, int64_t ldx,
                                      const double* norms, int64_t d, double sigma,
                                      double lambda, double shift, double* K, int64_t ldk,
                                      float* B32, int64_t ldb, float* V, int64_t ldv, double* u,
                                      qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || d < 0 || ldx < d || ldx % 4)
        return set_error(QMLH_ERR_ARG, "gaussian_deflated: bad shape (ldx multiple of 4, >= d)");
    if ((K && ldk < n) || ldb < n || (V && ldv < n))
        return set_error(QMLH_ERR_ARG, "gaussian_deflated: ld < n");
    if (reinterpret_cast<uintptr_t>(X) & 15)
        return set_error(QMLH_ERR_ARG, "gaussian_deflated: X must be 16-byte aligned");
    if (n == 0) return QMLH_OK;
    if (!X || !norms || !B32 || !u) return set_error(QMLH_ERR_ARG, "gaussian_deflated: null pointer");
    hipStream_t s = (hipStream_t)stream;
    const double scale = -0.5 / (sigma * sigma);
    hipLaunchKernelGGL(exp_scale_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, norms,
                       n, scale, u);
    QMLH_CHECK_LAUNCH("gaussian_deflated scale launch");
    if (V) {
        hipLaunchKernelGGL(lowrank_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, X, ldx,
                           n, d, u, 1.0 / sigma, V, ldv);
        QMLH_CHECK_LAUNCH("gaussian_deflated low-rank launch");
    }
    int64_t tm = ceil_div(n, GK_TILE);
    int64_t ntiles = tm * (tm + 1) / 2;
    if (ntiles > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "gaussian_deflated: grid too large");
    TimerScope ts(QMLH_TIMER_GAUSS, 2.0 * (double)d * (double)n * (double)(n + 1) / 2.0, s);
    hipLaunchKernelGGL(gaussian_deflated_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, X, ldx,
                       n, norms, u, d, scale, lambda, shift, K, ldk, B32, ldb);
    QMLH_CHECK_LAUNCH("gaussian_deflated launch");
    return QMLH_OK;
}

extern "C" int qmlh_gaussian_lowrank(const double* X, int64_t n, int64_t ldx, int64_t d,
                                     double sigma, const double* u, double* V, int64_t ldv,
                                     qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || d < 0 || ldx < d || ldv < n)
        return set_error(QMLH_ERR_ARG, "gaussian_lowrank: bad shape");
    if (n == 0) return QMLH_OK;
    if (!X || !u || !V) return set_error(QMLH_ERR_ARG, "gaussian_lowrank: null pointer");
    hipLaunchKernelGGL(lowrank64_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, X, ldx, n, d, u, 1.0 / sigma, V, ldv);
    QMLH_CHECK_LAUNCH("gaussian_lowrank launch");
    return QMLH_OK;
}
