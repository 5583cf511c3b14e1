// Mixed-precision solve: FP64 preconditioned conjugate gradients on A x = y, preconditioned by
// M = L L^T + V V^T, with L a low-precision Cholesky factor (FP32 storage) and V (n x r) a
// low-rank term kept out of the factorisation. With Y = L^-1 V and C = I + Y^T Y = R R^T
// (Woodbury),
//   M^-1 r = L^-T (q - Y C^-1 Y^T q),   q = L^-1 r,
// i.e. two triangular solves with L plus two skinny products with Y per application.
// Every residual r = y - A x is formed from the FP64 matrix, so the iteration converges to the
// FP64 solution; the preconditioner only sets how fast. (Iterative refinement with the
// correction equation solved by CG: CG-IR.)
//
// Stopping. The CG recurrence updates r recursively (r -= alpha A p), and that vector drifts
// from y - A x in finite precision. When the recursive residual meets the tolerance, the true
// residual y - A x is recomputed from A; if it is above the tolerance the CG restarts from it
// (preconditioner applied to the true residual) for as long as each restart at least halves it
// and iterations remain. The reported relative residual is always the recomputed one. When it
// stays above the tolerance -- the attainable floor of an FP64 residual, about
// u |A| |x| / |y|, lies above a tight tolerance when |x| is large -- the normwise backward
// error |y - A x| / (|A|_inf |x| + |y|) is computed as well, so that a caller can tell a
// solution at the FP64 floor (backward error below the unit roundoff, as a backward-stable
// FP64 Cholesky solve reaches) from one that did not converge.
//
// Reductions are deterministic: a fixed number of workgroups each reduce a fixed contiguous
// range in a fixed tree order, and one workgroup adds the partials in order. PCG scalars stay
// on the device; the host reads one residual norm per iteration to decide convergence.
#include <cfloat>

#include "linalg.h"
#include "timer.h"

namespace qmlh {

static constexpr int WB_MAX_RANK = 1024;

static constexpr int RED_BLOCKS = 256;

__device__ __forceinline__ double block_reduce_256(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    return red[0];
}

__global__ __launch_bounds__(256) void dot_partial_kernel(const double* __restrict__ a,
                                                          const double* __restrict__ b, int64_t n,
                                                          double* __restrict__ partial) {
    __shared__ double red[256];
    const int64_t chunk = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s = fma(a[i], b[i], s);
    const double tot = block_reduce_256(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void dot_final_kernel(const double* __restrict__ partial,
                                                        double* __restrict__ out) {
    __shared__ double red[256];
    const double tot = block_reduce_256(partial[threadIdx.x], red);
    if (threadIdx.x == 0) *out = tot;
}

// ------------------------------------------------------------------------------------------
// y = A x for symmetric A reading only its lower triangle (column-major), deterministic.
// Workgroup (strip J, chunk) takes the 64 columns 64 J .. and up to SY_CHUNK row blocks of 64
// from block J + SY_CHUNK chunk down. Each 64 x 64 tile is staged in LDS once and used twice:
// row sums sum_c A(r, c) x_c (c <= r on the diagonal tile) go to part_r[J][row] -- one
// partial per (strip, row) --, column sums sum_r A(r, c) x_r (r > c on the diagonal tile)
// accumulate over the chunk into part_c[J][chunk][c]. symv_reduce_kernel adds, for every i,
// the row partials of strips 0 .. i/64 and the column partials of strip i/64, in a fixed order.
// ------------------------------------------------------------------------------------------
static constexpr int SY_B = 64;
static constexpr int SY_CHUNK = 8;

__host__ __device__ inline int64_t symv_blocks(int64_t n) { return (n + SY_B - 1) / SY_B; }
__host__ __device__ inline int64_t symv_chunks(int64_t n) {
    return (symv_blocks(n) + SY_CHUNK - 1) / SY_CHUNK;
}
// doubles of workspace: row partials + column partials
static inline size_t symv_ws_doubles(int64_t n) {
    return (size_t)(symv_blocks(n) * n + symv_blocks(n) * symv_chunks(n) * SY_B);
}

// ABS: the same product with |A| and |x| (the scale of a backward error).
template <bool ABS>
__global__ __launch_bounds__(256) void symv_lower_kernel(const double* __restrict__ A, int64_t lda,
                                                         int64_t n, const double* __restrict__ x,
                                                         double* __restrict__ part_r,
                                                         double* __restrict__ part_c) {
    __shared__ double tile[SY_B][SY_B + 1];  // tile[r][c] = A(row0 + r, col0 + c)
    __shared__ double xj[SY_B], xr[SY_B];
    __shared__ double redr[4][SY_B];
    __shared__ double redc[4][SY_B];
    const int64_t J = blockIdx.y, chunk = blockIdx.x;
    const int64_t nblk = symv_blocks(n);
    const int64_t first = J + chunk * SY_CHUNK;
    if (first >= nblk) return;
    const int64_t last = first + SY_CHUNK < nblk ? first + SY_CHUNK : nblk;
    const int t = threadIdx.x;
    const int lane = t & 63, q = t >> 6;
    const int64_t col0 = J * SY_B;
    if (t < SY_B) xj[t] = col0 + t < n ? (ABS ? fabs(x[col0 + t]) : x[col0 + t]) : 0.0;
    double csum = 0.0;  // column lane, rows 16 q .. 16 q + 15 of every tile, in tile order
    for (int64_t blk = first; blk < last; ++blk) {
        const int64_t row0 = blk * SY_B;
        const bool diag = blk == J;
        double v[SY_B * SY_B / 256];
#pragma unroll
        for (int k = 0; k < SY_B * SY_B / 256; ++k) {
            const int idx = t + 256 * k;
            const int r = idx & (SY_B - 1), c = idx / SY_B;  // consecutive lanes: consecutive rows
            const bool ok = row0 + r < n && col0 + c < n && (!diag || r >= c);
            v[k] = ok ? A[(row0 + r) + (col0 + c) * lda] : 0.0;
        }
        __syncthreads();  // previous tile's readers are done
#pragma unroll
        for (int k = 0; k < SY_B * SY_B / 256; ++k) {
            const int idx = t + 256 * k;
            tile[idx & (SY_B - 1)][idx / SY_B] = v[k];
        }
        if (t < SY_B) xr[t] = row0 + t < n ? (ABS ? fabs(x[row0 + t]) : x[row0 + t]) : 0.0;
        __syncthreads();
        // row part: lane = row, columns 16 q ..
        double rs = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            // |a| where the tile is read: next to the load the compiler would move it into the
            // load's branch and wait for each load in turn
            const double a = ABS ? fabs(tile[lane][16 * q + c]) : tile[lane][16 * q + c];
            rs = fma(a, xj[16 * q + c], rs);
        }
        redr[q][lane] = rs;
        // column part: lane = column, rows 16 q .. (strictly below the diagonal on the diagonal tile)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int rr = 16 * q + r;
            const double a = (diag && rr <= lane) ? 0.0 : (ABS ? fabs(tile[rr][lane]) : tile[rr][lane]);
            csum = fma(a, xr[rr], csum);
        }
        __syncthreads();
        if (t < SY_B && row0 + t < n)
            part_r[J * n + row0 + t] = ((redr[0][t] + redr[1][t]) + redr[2][t]) + redr[3][t];
    }
    redc[q][lane] = csum;
    __syncthreads();
    if (t < SY_B)
        part_c[(J * symv_chunks(n) + chunk) * SY_B + t] =
            ((redc[0][t] + redc[1][t]) + redc[2][t]) + redc[3][t];
}

__global__ void symv_reduce_kernel(const double* __restrict__ part_r,
                                   const double* __restrict__ part_c, int64_t n,
                                   double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t I = i / SY_B;
    double s = 0.0;
    for (int64_t J = 0; J <= I; ++J) s += part_r[J * n + i];
    const int64_t nch = (symv_blocks(n) - I + SY_CHUNK - 1) / SY_CHUNK;  // chunks of strip I
    const int64_t base = I * symv_chunks(n) * SY_B + (i - I * SY_B);
    for (int64_t c = 0; c < nch; ++c) s += part_c[base + c * SY_B];
    y[i] = s;
}

static int symv_lower(const double* A, int64_t lda, int64_t n, const double* x, double* y,
                      double* ws, hipStream_t s, bool abs_values = false) {
    double* part_r = ws;
    double* part_c = ws + symv_blocks(n) * n;
    const dim3 grid((unsigned)symv_chunks(n), (unsigned)symv_blocks(n));
    if (abs_values)
        hipLaunchKernelGGL(symv_lower_kernel<true>, grid, dim3(256), 0, s, A, lda, n, x, part_r,
                           part_c);
    else
        hipLaunchKernelGGL(symv_lower_kernel<false>, grid, dim3(256), 0, s, A, lda, n, x, part_r,
                           part_c);
    hipLaunchKernelGGL(symv_reduce_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                       part_r, part_c, n, y);
    QMLH_CHECK_LAUNCH("pcg symv launch");
    return QMLH_OK;
}

// sc layout: 0 rz, 1 pq, 2 rr, 3 (unused), 4 u.z, 5 den = 1 + u.Bu, 6 rz_new
__global__ void axpy2_kernel(double* __restrict__ x, double* __restrict__ r,
                             const double* __restrict__ p, const double* __restrict__ q,
                             const double* __restrict__ sc, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = sc[0] / sc[1];
    x[i] = fma(alpha, p[i], x[i]);
    r[i] = fma(-alpha, q[i], r[i]);
}

// r = y - q (the true residual once q = A x).
__global__ void residual_kernel(const double* __restrict__ y, const double* __restrict__ q,
                                double* __restrict__ r, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = y[i] - q[i];
}

// Largest entry of a (n entries, all >= 0): per-workgroup maxima, then one workgroup.
__global__ __launch_bounds__(256) void max_partial_kernel(const double* __restrict__ a, int64_t n,
                                                          double* __restrict__ partial) {
    __shared__ double red[256];
    const int64_t chunk = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    double m = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) m = fmax(m, a[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void max_final_kernel(const double* __restrict__ partial,
                                                        double* __restrict__ out) {
    __shared__ double red[256];
    red[threadIdx.x] = partial[threadIdx.x];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// t_a = sum_i Y(i, a) q_i: one workgroup per column a, fixed-order reduction (Y FP32 or FP64).
template <typename TY>
__global__ __launch_bounds__(256) void yt_q_kernel(const TY* __restrict__ Y, int64_t ldy,
                                                   int64_t n, const double* __restrict__ q,
                                                   double* __restrict__ t) {
    __shared__ double red[256];
    const TY* col = Y + (int64_t)blockIdx.x * ldy;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s = fma((double)col[i], q[i], s);
    const double tot = block_reduce_256(s, red);
    if (threadIdx.x == 0) t[blockIdx.x] = tot;
}

// q_i -= sum_a Y(i, a) t_a (rank r <= WB_MAX_RANK).
template <typename TY>
__global__ __launch_bounds__(256) void q_minus_yt_kernel(const TY* __restrict__ Y, int64_t ldy,
                                                         int64_t n, int64_t r,
                                                         const double* __restrict__ t,
                                                         double* __restrict__ q) {
    __shared__ double st[WB_MAX_RANK];
    for (int64_t a = threadIdx.x; a < r; a += 256) st[a] = t[a];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t a = 0; a < r; ++a) s = fma((double)Y[i + a * ldy], st[a], s);
    q[i] = q[i] - s;
}

// C = I + Y^T Y (r x r, FP64): a 16 x 16 block per workgroup, rows streamed through LDS.
__global__ __launch_bounds__(256) void gram_kernel(const float* __restrict__ Y, int64_t ldy,
                                                   int64_t n, int64_t r, double* __restrict__ C,
                                                   int64_t ldc) {
    __shared__ double sa[64][17], sb[64][17];
    const int ta = threadIdx.x & 15, tb = threadIdx.x >> 4;
    const int64_t a0 = (int64_t)blockIdx.x * 16, b0 = (int64_t)blockIdx.y * 16;
    double acc = 0.0;
    for (int64_t i0 = 0; i0 < n; i0 += 64) {
        for (int idx = threadIdx.x; idx < 64 * 16; idx += 256) {
            const int ii = idx % 64, cc = idx / 64;
            const int64_t i = i0 + ii;
            sa[ii][cc] = (i < n && a0 + cc < r) ? (double)Y[i + (a0 + cc) * ldy] : 0.0;
            sb[ii][cc] = (i < n && b0 + cc < r) ? (double)Y[i + (b0 + cc) * ldy] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int ii = 0; ii < 64; ++ii) acc = fma(sa[ii][ta], sb[ii][tb], acc);
        __syncthreads();
    }
    const int64_t a = a0 + ta, b = b0 + tb;
    if (a < r && b < r) C[a + b * ldc] = acc + (a == b ? 1.0 : 0.0);
}

// C = I + Y^T Y in two deterministic passes. Pass 1: workgroup (block, slice) takes one
// 32 x 32 block of the lower triangle (block index over the lower blocks) and one slice of the
// rows of Y, stages 64-row chunks of the two column groups in LDS as FP64, and accumulates a
// 2 x 2 micro tile per lane; partial[slice] holds the slice's block. Pass 2 adds the slices in
// order, adds I, and writes both triangles from the lower one.
static constexpr int GR_T = 32;
static constexpr int GR_SLICES_MAX = 16;

static int64_t gram_slices(int64_t n) {
    const int64_t s = ceil_div(n, (int64_t)1024);
    return s < 1 ? 1 : (s > GR_SLICES_MAX ? GR_SLICES_MAX : s);
}

template <typename TY>
__global__ __launch_bounds__(256) void gram_partial_kernel(const TY* __restrict__ Y,
                                                           int64_t ldy, int64_t n, int64_t r,
                                                           int64_t slice_rows,
                                                           double* __restrict__ partial) {
    __shared__ double sa[64][GR_T + 1], sb[64][GR_T + 1];
    int64_t bi = (int64_t)((sqrt(8.0 * (double)blockIdx.x + 1.0) - 1.0) * 0.5);
    while (bi * (bi + 1) / 2 > (int64_t)blockIdx.x) --bi;
    while ((bi + 1) * (bi + 2) / 2 <= (int64_t)blockIdx.x) ++bi;
    const int64_t bj = blockIdx.x - bi * (bi + 1) / 2;
    const int64_t a0 = bi * GR_T, b0 = bj * GR_T;
    const int64_t lo = (int64_t)blockIdx.y * slice_rows;
    const int64_t hi = lo + slice_rows < n ? lo + slice_rows : n;
    const int t = threadIdx.x, ta = t & 15, tb = t >> 4;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int64_t i0 = lo; i0 < hi; i0 += 64) {
        TY ra[8], rb[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {  // element (ii, cc) of the 64 x 32 chunk, ii fastest
            const int idx = t + 256 * q, ii = idx & 63, cc = idx >> 6;
            const int64_t i = i0 + ii < hi ? i0 + ii : hi - 1;
            const int64_t ca = a0 + cc < r ? a0 + cc : r - 1, cb = b0 + cc < r ? b0 + cc : r - 1;
            ra[q] = Y[i + ca * ldy];
            rb[q] = Y[i + cb * ldy];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int idx = t + 256 * q, ii = idx & 63, cc = idx >> 6;
            const bool in = i0 + ii < hi;
            sa[ii][cc] = in && a0 + cc < r ? (double)ra[q] : 0.0;
            sb[ii][cc] = in && b0 + cc < r ? (double)rb[q] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int ii = 0; ii < 64; ++ii) {
            const double x0 = sa[ii][ta], x1 = sa[ii][ta + 16];
            const double y0 = sb[ii][tb], y1 = sb[ii][tb + 16];
            acc[0][0] = fma(x0, y0, acc[0][0]);
            acc[0][1] = fma(x0, y1, acc[0][1]);
            acc[1][0] = fma(x1, y0, acc[1][0]);
            acc[1][1] = fma(x1, y1, acc[1][1]);
        }
        __syncthreads();
    }
    double* P = partial + (int64_t)blockIdx.y * r * r;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int64_t a = a0 + ta + 16 * u, b = b0 + tb + 16 * v;
            if (a < r && b < r) P[a + b * r] = acc[u][v];
        }
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(const double* __restrict__ partial,
                                                          int64_t r, int64_t slices,
                                                          double* __restrict__ C, int64_t ldc) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= r * r) return;
    const int64_t a = idx % r, b = idx / r;
    if (a < b) return;  // the mirror entry writes it
    double v = 0.0;
    for (int64_t s = 0; s < slices; ++s) v += partial[s * r * r + idx];
    if (a == b) v += 1.0;
    C[a + b * ldc] = v;
    C[b + a * ldc] = v;
}

__global__ void p_update_kernel(double* __restrict__ p, const double* __restrict__ z,
                                double* __restrict__ sc, int64_t n, int first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double beta = first ? 0.0 : sc[6] / sc[0];
        p[i] = first ? z[i] : fma(beta, p[i], z[i]);
    }
}

__global__ void rz_roll_kernel(double* sc) { sc[0] = sc[6]; }

__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void fill_kernel(double* __restrict__ a, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void add_diagonal_kernel(double* A, int64_t n, int64_t lda, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i + i * lda] += v;
}

// B(i, j) = (float)(A(i, j) + shift [i == j]) for i >= j (column-major, lower triangle).
__global__ void lower_to_f32_kernel(const double* __restrict__ A, int64_t lda, int64_t n,
                                    float* __restrict__ B, int64_t ldb, double shift) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int64_t j = blockIdx.y; j <= i; j += gridDim.y) {
        double v = A[i + j * lda];
        if (i == j) v += shift;
        B[i + j * ldb] = (float)v;
    }
}

struct PcgCtx {
    hipStream_t s;
    int64_t n;
    double* partial;
    int32_t* sync;           // triangular-solve synchronisation words
    const double* dinv128;   // FP64 inverses of the factor's 128 x 128 diagonal blocks
};

static inline unsigned grid1(int64_t n) { return (unsigned)ceil_div(n, 256); }

static int dot(const PcgCtx& c, const double* a, const double* b, double* out) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(RED_BLOCKS), dim3(256), 0, c.s, a, b, c.n,
                       c.partial);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, c.s, c.partial, out);
    QMLH_CHECK_LAUNCH("pcg dot launch");
    return QMLH_OK;
}

static int copy(const PcgCtx& c, const double* a, double* b) {
    hipLaunchKernelGGL(copy_kernel, dim3(grid1(c.n)), dim3(256), 0, c.s, a, b, c.n);
    QMLH_CHECK_LAUNCH("pcg copy launch");
    return QMLH_OK;
}

struct Woodbury {
    const void* Y;  // float, or double with y64
    bool y64;
    int64_t ldy;
    int64_t r;
    const double* C;  // Cholesky factor of I + Y^T Y
    int64_t ldc;
    double* t;        // r
    double* tmp_r;    // r
};

// z = M^-1 r; tmp and q are n-vectors of scratch.
static int precond(const PcgCtx& c, const float* L, int64_t ldl, const Woodbury& wb,
                   const double* r, double* z, double* q, double* tmp) {
    QMLH_TRY(copy(c, r, tmp));
    const bool big = c.n >= TRSV_SYNCFREE_MIN;
    if (big)
        QMLH_TRY(trsv_syncfree128<float>(L, c.n, ldl, tmp, q, c.sync, true, c.s, c.dinv128));
    else
        QMLH_TRY(trsv_lower_fwd<float>(L, c.n, ldl, tmp, q, nullptr, c.s));
    if (wb.r > 0) {
        if (wb.y64)
            hipLaunchKernelGGL(yt_q_kernel<double>, dim3((unsigned)wb.r), dim3(256), 0, c.s,
                               static_cast<const double*>(wb.Y), wb.ldy, c.n, q, wb.t);
        else
            hipLaunchKernelGGL(yt_q_kernel<float>, dim3((unsigned)wb.r), dim3(256), 0, c.s,
                               static_cast<const float*>(wb.Y), wb.ldy, c.n, q, wb.t);
        QMLH_CHECK_LAUNCH("pcg Y^T q launch");
        QMLH_TRY(potrs_impl<double>(wb.C, wb.r, wb.ldc, wb.t, wb.tmp_r, nullptr, c.s));
        if (wb.y64)
            hipLaunchKernelGGL(q_minus_yt_kernel<double>, dim3(grid1(c.n)), dim3(256), 0, c.s,
                               static_cast<const double*>(wb.Y), wb.ldy, c.n, wb.r, wb.t, q);
        else
            hipLaunchKernelGGL(q_minus_yt_kernel<float>, dim3(grid1(c.n)), dim3(256), 0, c.s,
                               static_cast<const float*>(wb.Y), wb.ldy, c.n, wb.r, wb.t, q);
        QMLH_CHECK_LAUNCH("pcg q - Y t launch");
    }
    if (big) return trsv_syncfree128<float>(L, c.n, ldl, q, z, c.sync, false, c.s, c.dinv128);
    return trsv_lower_bwd<float>(L, c.n, ldl, q, z, nullptr, c.s);
}

}  // namespace qmlh

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

extern "C" int qmlh_lower_to_f32(const double* A, int64_t n, int64_t lda, float* B, int64_t ldb,
                                 double shift, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || lda < n || ldb < n) return set_error(QMLH_ERR_ARG, "lower_to_f32: bad shape");
    if (n == 0) return QMLH_OK;
    if (!A || !B) return set_error(QMLH_ERR_ARG, "lower_to_f32: null pointer");
    const unsigned gy = (unsigned)(n < 4096 ? n : 4096);
    hipLaunchKernelGGL(lower_to_f32_kernel, dim3((unsigned)ceil_div(n, 256), gy), dim3(256),
                       0, (hipStream_t)stream, A, lda, n, B, ldb, shift);
    QMLH_CHECK_LAUNCH("lower_to_f32 launch");
    return QMLH_OK;
}

extern "C" size_t qmlh_woodbury_workspace(int64_t n) {
    return n > 0 ? qmlh::trsv_dinv_doubles(n) * sizeof(double) : 0;
}

extern "C" int qmlh_woodbury_setup(const float* L, int64_t n, int64_t ldl, float* V, int64_t ldv,
                                   int64_t r, double* C, int64_t ldc, int32_t* info, void* ws,
                                   size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || r < 0 || ldl < n || ldv < n || ldc < r)
        return set_error(QMLH_ERR_ARG, "woodbury_setup: bad shape");
    if (r > WB_MAX_RANK)
        return set_error(QMLH_ERR_ARG, "woodbury_setup: rank %lld above %d", (long long)r,
                         WB_MAX_RANK);
    if (n == 0 || r == 0) return QMLH_OK;
    if (!L || !V || !C || !info || !ws) return set_error(QMLH_ERR_ARG, "woodbury_setup: null pointer");
    if (ws_bytes < qmlh_woodbury_workspace(n))
        return set_error(QMLH_ERR_WORKSPACE, "woodbury_setup: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    double* dinv = static_cast<double*>(ws);
    QMLH_TRY(trsv_diag_inverse<float>(L, n, ldl, dinv, s));
    QMLH_TRY(trsm_lower_fwd_multi(L, n, ldl, V, ldv, r, dinv, s));
    dim3 g((unsigned)ceil_div(r, 16), (unsigned)ceil_div(r, 16));
    hipLaunchKernelGGL(gram_kernel, g, dim3(256), 0, s, V, ldv, n, r, C, ldc);
    QMLH_CHECK_LAUNCH("woodbury gram launch");
    return potrf_impl<double>(C, r, ldc, info, s);
}

extern "C" size_t qmlh_woodbury_capacitance_workspace(int64_t n, int64_t r) {
    if (n <= 0 || r <= 0) return 0;
    return (size_t)qmlh::gram_slices(n) * (size_t)r * (size_t)r * sizeof(double);
}

namespace qmlh {
template <typename TY>
static int woodbury_capacitance_impl(const TY* Y, int64_t n, int64_t ldy, int64_t r, double* C,
                                     int64_t ldc, int32_t* info, void* ws, size_t ws_bytes,
                                     hipStream_t s) {
    if (n < 0 || r < 0 || ldy < (n > 1 ? n : 1) || ldc < r)
        return set_error(QMLH_ERR_ARG, "woodbury_capacitance: bad shape");
    if (r > WB_MAX_RANK)
        return set_error(QMLH_ERR_ARG, "woodbury_capacitance: rank %lld above %d", (long long)r,
                         WB_MAX_RANK);
    if (r == 0) return QMLH_OK;
    if (!C || !info || (n > 0 && (!Y || !ws)))
        return set_error(QMLH_ERR_ARG, "woodbury_capacitance: null pointer");
    if (ws_bytes < qmlh_woodbury_capacitance_workspace(n, r))
        return set_error(QMLH_ERR_WORKSPACE, "woodbury_capacitance: workspace too small");
    double* partial = static_cast<double*>(ws);
    const int64_t slices = n > 0 ? gram_slices(n) : 0;
    if (slices > 0) {
        const int64_t nbk = ceil_div(r, (int64_t)GR_T);
        const int64_t slice_rows = ceil_div(ceil_div(n, slices), (int64_t)64) * 64;
        hipLaunchKernelGGL(gram_partial_kernel<TY>, dim3((unsigned)(nbk * (nbk + 1) / 2), (unsigned)slices),
                           dim3(256), 0, s, Y, ldy, n, r, slice_rows, partial);
        QMLH_CHECK_LAUNCH("woodbury gram launch");
    }
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)ceil_div(r * r, (int64_t)256)), dim3(256),
                       0, s, partial, r, slices, C, ldc);
    QMLH_CHECK_LAUNCH("woodbury gram reduce launch");
    return potrf_impl<double>(C, r, ldc, info, s);
}
}  // namespace qmlh

extern "C" int qmlh_woodbury_capacitance(const float* Y, int64_t n, int64_t ldy, int64_t r,
                                         double* C, int64_t ldc, int32_t* info, void* ws,
                                         size_t ws_bytes, qmlh_stream_t stream) {
    return qmlh::woodbury_capacitance_impl<float>(Y, n, ldy, r, C, ldc, info, ws, ws_bytes,
                                                  (hipStream_t)stream);
}

extern "C" int qmlh_woodbury_capacitance_f64(const double* Y, int64_t n, int64_t ldy, int64_t r,
                                             double* C, int64_t ldc, int32_t* info, void* ws,
                                             size_t ws_bytes, qmlh_stream_t stream) {
    return qmlh::woodbury_capacitance_impl<double>(Y, n, ldy, r, C, ldc, info, ws, ws_bytes,
                                                   (hipStream_t)stream);
}

extern "C" size_t qmlh_pcg_workspace(int64_t n, int64_t r) {
    if (n <= 0) return 0;
    if (r < 0) r = 0;
    return (size_t)(5 * n + 2 * r + qmlh::RED_BLOCKS + 8) * sizeof(double) +
           (qmlh::trsv_dinv_doubles(n) + qmlh::trsv_dinv128_doubles(n)) * sizeof(double) +
           qmlh::symv_ws_doubles(n) * sizeof(double) +
           qmlh::trsv_sync_ints(n) * sizeof(int32_t);
}

extern "C" int qmlh_pcg_solve(const double* A, int64_t n, int64_t lda, const float* L,
                              int64_t ldl, const void* Y, int64_t ldy, int64_t r, int32_t y_fp64,
                              const double* C, int64_t ldc, const double* y, double* x,
                              double tol, int32_t max_iter, void* ws, size_t ws_bytes,
                              int32_t* iters_h, double* relres_h, double* backward_h,
                              qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || lda < n || ldl < n || r < 0) return set_error(QMLH_ERR_ARG, "pcg_solve: bad shape");
    if (r > 0 && (!Y || !C || ldy < n || ldc < r || r > WB_MAX_RANK))
        return set_error(QMLH_ERR_ARG, "pcg_solve: bad low-rank term");
    if (iters_h) *iters_h = 0;
    if (relres_h) *relres_h = 0.0;
    if (backward_h) *backward_h = 0.0;
    if (n == 0) return QMLH_OK;
    if (!A || !L || !y || !x || !ws) return set_error(QMLH_ERR_ARG, "pcg_solve: null pointer");
    if (ws_bytes < qmlh_pcg_workspace(n, r))
        return set_error(QMLH_ERR_WORKSPACE, "pcg_solve: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    TimerScope timer(QMLH_TIMER_PCG, 0.0, s);
    double* w = static_cast<double*>(ws);
    double* rv = w;
    double* z = rv + n;
    double* p = z + n;
    double* q = p + n;
    double* tmp = q + n;
    double* t = tmp + n;
    double* tmp_r = t + r;
    double* partial = tmp_r + r;
    double* sc = partial + RED_BLOCKS;  // 8 scalars: 0 rz, 1 pq, 2 rr, 3 |A|_inf, 6 rz_new, 7 xx
    double* dinv = sc + 8;
    double* dinv128 = dinv + trsv_dinv_doubles(n);
    double* symws = dinv128 + trsv_dinv128_doubles(n);
    int32_t* sync = reinterpret_cast<int32_t*>(symws + symv_ws_doubles(n));
    PcgCtx c{s, n, partial, sync, dinv128};
    Woodbury wb{Y, y_fp64 != 0, ldy, r, C, ldc, t, tmp_r};
    auto read_scalar = [&](const double* dev, double* host, const char* what) -> int {
        QMLH_TRY(check_hip(hipMemcpyAsync(host, dev, sizeof(double), hipMemcpyDeviceToHost, s),
                           what));
        return check_hip(hipStreamSynchronize(s), "pcg: sync");
    };
    // The preconditioner's triangular solves use explicit FP64 inverses of the 128 x 128
    // diagonal blocks of the low-precision factor (computed once here), so each block's
    // diagonal solve is one parallel product instead of a sequential substitution.
    if (n >= TRSV_SYNCFREE_MIN) QMLH_TRY(trsv_inverse128<float>(L, n, ldl, dinv, dinv128, s));

    double yy_h = 0.0;
    QMLH_TRY(dot(c, y, y, sc + 2));
    QMLH_TRY(read_scalar(sc + 2, &yy_h, "pcg: read |y|"));
    const double ynorm = sqrt(yy_h);
    hipLaunchKernelGGL(fill_kernel, dim3(grid1(n)), dim3(256), 0, s, x, 0.0, n);
    QMLH_CHECK_LAUNCH("pcg fill launch");
    if (ynorm == 0.0) return QMLH_OK;
    // (Re)start from the residual in rv: z = M^-1 r, p = z, rz = r.z.
    auto restart = [&]() -> int {
        QMLH_TRY(precond(c, L, ldl, wb, rv, z, q, tmp));
        QMLH_TRY(dot(c, rv, z, sc + 0));
        hipLaunchKernelGGL(p_update_kernel, dim3(grid1(n)), dim3(256), 0, s, p, z, sc, n, 1);
        QMLH_CHECK_LAUNCH("pcg p launch");
        return QMLH_OK;
    };
    QMLH_TRY(copy(c, y, rv));
    QMLH_TRY(restart());
    double relres = 1.0, prev_true = HUGE_VAL, anorm = -1.0, backward = -1.0;
    int it = 0;
    for (;;) {
        while (it < max_iter) {
            ++it;
            QMLH_TRY(symv_lower(A, lda, n, p, q, symws, s));
            QMLH_TRY(dot(c, p, q, sc + 1));
            hipLaunchKernelGGL(axpy2_kernel, dim3(grid1(n)), dim3(256), 0, s, x, rv, p, q, sc, n);
            QMLH_CHECK_LAUNCH("pcg axpy launch");
            QMLH_TRY(dot(c, rv, rv, sc + 2));
            double rr_h = 0.0;
            QMLH_TRY(read_scalar(sc + 2, &rr_h, "pcg: read |r|"));
            relres = sqrt(rr_h) / ynorm;
            if (!(relres > tol)) break;  // converged (or NaN: stop and check)
            QMLH_TRY(precond(c, L, ldl, wb, rv, z, q, tmp));
            QMLH_TRY(dot(c, rv, z, sc + 6));
            hipLaunchKernelGGL(p_update_kernel, dim3(grid1(n)), dim3(256), 0, s, p, z, sc, n, 0);
            QMLH_CHECK_LAUNCH("pcg p launch");
            hipLaunchKernelGGL(rz_roll_kernel, dim3(1), dim3(1), 0, s, sc);
            QMLH_CHECK_LAUNCH("pcg roll launch");
        }
        // true residual r = y - A x
        QMLH_TRY(symv_lower(A, lda, n, x, q, symws, s));
        hipLaunchKernelGGL(residual_kernel, dim3(grid1(n)), dim3(256), 0, s, y, q, rv, n);
        QMLH_CHECK_LAUNCH("pcg residual launch");
        QMLH_TRY(dot(c, rv, rv, sc + 2));
        double rr_h = 0.0;
        QMLH_TRY(read_scalar(sc + 2, &rr_h, "pcg: read true |r|"));
        relres = sqrt(rr_h) / ynorm;
        if (!(relres > tol)) {
            backward = -1.0;  // not computed: the residual test passed
            break;
        }
        // normwise backward error; |A|_inf = max_i (|A| 1)_i, the lower triangle read as the
        // symmetric matrix (computed once)
        if (anorm < 0.0) {
            hipLaunchKernelGGL(fill_kernel, dim3(grid1(n)), dim3(256), 0, s, tmp, 1.0, n);
            QMLH_CHECK_LAUNCH("pcg fill launch");
            QMLH_TRY(symv_lower(A, lda, n, tmp, q, symws, s, true));
            hipLaunchKernelGGL(max_partial_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, q, n, partial);
            hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, s, partial, sc + 3);
            QMLH_CHECK_LAUNCH("pcg norm launch");
            QMLH_TRY(read_scalar(sc + 3, &anorm, "pcg: read |A|"));
        }
        QMLH_TRY(dot(c, x, x, sc + 7));
        double xx = 0.0;
        QMLH_TRY(read_scalar(sc + 7, &xx, "pcg: read |x|"));
        backward = relres * ynorm / (anorm * sqrt(xx) + ynorm);
        // At the FP64 floor another restart cannot lower the residual: a backward error below
        // the unit roundoff u = 2^-53 means x solves a system whose matrix is within the
        // rounding of A's own FP64 entries.
        const bool at_floor = backward <= 0.5 * DBL_EPSILON;
        if (at_floor || it >= max_iter || !(relres < 0.5 * prev_true)) break;
        prev_true = relres;
        QMLH_TRY(restart());
    }
    timer.set_work((double)it);
    if (iters_h) *iters_h = it;
    if (relres_h) *relres_h = relres;
    if (backward_h) *backward_h = backward;
    return QMLH_OK;
}
