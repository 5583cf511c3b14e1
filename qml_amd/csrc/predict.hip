// Fused KRR prediction: y_q = sum_j k(q, x_j) alpha_j without materialising the kernel,
// the Ks.T.dot(alpha) step of tests/test_energy_krr_cmat.py:108-109.
//
// Batch invariance: the training set is cut into fixed chunks (256 rows for the Gaussian,
// 128 for the Laplacian); a workgroup reduces one chunk for a block of queries in an order
// that depends only on the training index, writes one partial per (chunk, query), and a second
// kernel adds the partials of a query in chunk order. Nothing in a query's arithmetic depends
// on the other queries or on m, so a query's result is bitwise the same alone or in a batch.
#include <mutex>

#include "mfma_core.h"
#include "pair_tile.h"
#include "timer.h"

namespace qmlh {

static constexpr int PG_Q = 64;     // queries per Gaussian workgroup
static constexpr int PG_T = 256;    // training rows per Gaussian workgroup (4 waves x 64)
static constexpr int PL_Q = 128;    // queries per Laplacian workgroup
static constexpr int PL_T = 128;    // training rows per Laplacian workgroup

// Gaussian: MFMA cross term, FP64 epilogue, per-chunk partials part[chunk * m + q].
__global__ __launch_bounds__(256) void predict_gauss_kernel(
    const double* __restrict__ Q, int64_t ldq, int64_t m, const double* __restrict__ nQ,
    const double* __restrict__ T, int64_t ldt, int64_t n, const double* __restrict__ nT,
    const double* __restrict__ alpha, int64_t d, double scale, double* __restrict__ part) {
    __shared__ double red[4][PG_Q];
    const int64_t q0 = (int64_t)blockIdx.x * PG_Q;
    const int64_t chunk = blockIdx.y;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t t0 = chunk * PG_T + 64 * w;

    double qsum[4] = {0.0, 0.0, 0.0, 0.0};  // per query block b, this lane's query (lane & 15)
    if (t0 < n) {
        Operand<double, LAYOUT_ROW> opT{T, ldt, n, d};
        Operand<double, LAYOUT_ROW> opQ{Q, ldq, m, d};
        WaveTile<double> t;
        wave_tile_zero(t);
        // First operand: training rows (accumulator rows), second: queries (columns).
        wave_tile_mma_rows(t, opT, opQ, t0, q0, d);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t q = q0 + 16 * b + (lane & 15);
            const int64_t qc = q < m ? q : m - 1;
            const double nq = nQ[qc];
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t j = t0 + tile_row<double>(a, e);
                    if (j < n) {
                        double d2 = nq + nT[j] - 2.0 * t.acc[a][b][e];
                        d2 = d2 > 0.0 ? d2 : 0.0;
                        s = fma(exp(scale * d2), alpha[j], s);
                    }
                }
            }
            // Lanes l, l+16, l+32, l+48 hold the same query: fixed-order butterfly.
            s = s + __shfl_xor(s, 16, 64);
            s = s + __shfl_xor(s, 32, 64);
            qsum[b] = s;
        }
    }
    if (lane < 16) {
#pragma unroll
        for (int b = 0; b < 4; ++b) red[w][16 * b + lane] = qsum[b];
    }
    __syncthreads();
    if (threadIdx.x < PG_Q) {
        const int64_t q = q0 + threadIdx.x;
        if (q < m) {
            const int qi = threadIdx.x;
            double s = ((red[0][qi] + red[1][qi]) + red[2][qi]) + red[3][qi];
            part[chunk * m + q] = s;
        }
    }
}

// Laplacian: the pairwise kernel's VALU L1 tile (pair_tile_accumulate: LDS-staged chunks, the
// next chunk's loads in flight during the current chunk's sums), exp * alpha, per-chunk
// partials.
__global__ __launch_bounds__(256, 2) void predict_laplace_kernel(
    const double* __restrict__ Q, int64_t ldq, int64_t m, const double* __restrict__ T,
    int64_t ldt, int64_t n, const double* __restrict__ alpha, int64_t d, double scale,
    double* __restrict__ part) {
    static_assert(PL_Q == PT_BM && PL_T == PT_BN, "the predict tile is the pairwise tile");
    __shared__ __attribute__((aligned(16))) double sA[PT_BK][PT_LDS];
    __shared__ __attribute__((aligned(16))) double sB[PT_BK][PT_LDS];
    __shared__ double red[16][PL_Q];
    const int64_t row0 = (int64_t)blockIdx.x * PL_Q;  // queries
    const int64_t chunk = blockIdx.y;
    const int64_t col0 = chunk * PL_T;  // training rows
    const int t = threadIdx.x;
    const int tx = t % 16, ty = t / 16;
    double acc[8][8];
    pair_tile_accumulate<OP_L1>(acc, sA, sB, Q, ldq, m, T, ldt, n, d, row0, col0);
    // Row sums over this lane's 8 training columns, in column order.
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t j = col0 + 2 * ty + 32 * (c / 2) + (c % 2);
            if (j < n) s = fma(exp(scale * acc[r][c]), alpha[j], s);
        }
        red[ty][2 * tx + 32 * (r / 2) + (r % 2)] = s;
    }
    __syncthreads();
    if (t < PL_Q) {
        const int64_t q = row0 + t;
        if (q < m) {
            double s = 0.0;
            for (int g = 0; g < 16; ++g) s = s + red[g][t];
            part[chunk * m + q] = s;
        }
    }
}

__global__ void predict_reduce_kernel(const double* __restrict__ part, int64_t m, int64_t nchunks,
                                      double* __restrict__ y) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    double s = 0.0;
    for (int64_t c = 0; c < nchunks; ++c) s = s + part[c * m + q];
    y[q] = s;
}

__global__ void zero_kernel(double* __restrict__ y, int64_t m) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m) y[q] = 0.0;
}

// ------------------------------------------------------------------------------------------
// Scratch of qmlh_krr_predict_noalloc: the per-(chunk, query) partials in a file-scope device
// array instead of a caller's workspace. Each GPU holds its own copy for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name on the host is a stand-in, not a device address).
// ------------------------------------------------------------------------------------------
static constexpr int64_t PREDICT_SCRATCH_DOUBLES = (int64_t)2 << 20;  // 16 MiB
static __device__ __align__(256) double scratch[PREDICT_SCRATCH_DOUBLES];

static constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[MAX_DEVICES] = {};

// Resolve (once per device) the current device's copy of the scratch array.
int predict_scratch_setup(double** out) {
    int dev = 0;
    QMLH_TRY(check_hip(hipGetDevice(&dev), "predict scratch: hipGetDevice"));
    if (dev < 0 || dev >= MAX_DEVICES)
        return set_error(QMLH_ERR_ARG, "predict scratch: device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        QMLH_TRY(check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "predict scratch: hipGetSymbolAddress"));
        if (!p) return set_error(QMLH_ERR_HIP, "predict scratch: null symbol address");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    if (out) *out = g_scratch_ptr[dev];
    return QMLH_OK;
}

size_t predict_workspace_bytes(int64_t m, int64_t n) {
    if (m <= 0 || n <= 0) return 0;
    const int64_t chunks = ceil_div(n, PL_T);  // the finer of the two chunkings
    return (size_t)(chunks * m) * sizeof(double);
}

// The prediction with its partials in `part` (at least predict_workspace_bytes(m, n)).
static int predict_impl(int32_t kind, const double* Xq, int64_t m, int64_t ldq,
                        const double* normsQ, const double* Xt, int64_t n, int64_t ldt,
                        const double* normsT, const double* alpha, int64_t d, double sigma,
                        double* y, double* part, hipStream_t s) {
    if (n == 0) {  // empty training set: y = 0 (a kernel, not a memset, so capture is safe)
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, y, m);
        QMLH_CHECK_LAUNCH("krr_predict zero launch");
        return QMLH_OK;
    }
    int64_t nchunks;
    TimerScope ts(QMLH_TIMER_PREDICT, (double)m * (double)n, s);
    if (kind == QMLH_PAIR_GAUSSIAN) {
        nchunks = ceil_div(n, PG_T);
        dim3 grid((unsigned)ceil_div(m, PG_Q), (unsigned)nchunks);
        hipLaunchKernelGGL(predict_gauss_kernel, grid, dim3(256), 0, s, Xq, ldq, m, normsQ, Xt,
                           ldt, n, normsT, alpha, d, -0.5 / (sigma * sigma), part);
    } else {
        nchunks = ceil_div(n, PL_T);
        dim3 grid((unsigned)ceil_div(m, PL_Q), (unsigned)nchunks);
        hipLaunchKernelGGL(predict_laplace_kernel, grid, dim3(256), 0, s, Xq, ldq, m, Xt, ldt, n,
                           alpha, d, -1.0 / sigma, part);
    }
    QMLH_CHECK_LAUNCH("krr_predict launch");
    hipLaunchKernelGGL(predict_reduce_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s,
                       part, m, nchunks, y);
    QMLH_CHECK_LAUNCH("krr_predict reduce launch");
    return QMLH_OK;
}

// Argument checks shared by both entry points (everything but the workspace).
static int predict_check(int32_t kind, const double* Xq, int64_t m, int64_t ldq,
                         const double* normsQ, const double* Xt, int64_t n, int64_t ldt,
                         const double* normsT, const double* alpha, int64_t d, double* y,
                         const char* who) {
    if (m < 0 || n < 0 || d < 0) return set_error(QMLH_ERR_ARG, "%s: negative size", who);
    if (ldq < d || ldt < d) return set_error(QMLH_ERR_ARG, "%s: ld < d", who);
    if (kind != QMLH_PAIR_GAUSSIAN && kind != QMLH_PAIR_LAPLACIAN)
        return set_error(QMLH_ERR_ARG, "%s: kind must be Gaussian or Laplacian", who);
    if (m == 0) return QMLH_OK;
    if (!y) return set_error(QMLH_ERR_ARG, "%s: null pointer", who);
    if (n == 0) return QMLH_OK;
    if (!Xq || !Xt || !alpha) return set_error(QMLH_ERR_ARG, "%s: null pointer", who);
    if (kind == QMLH_PAIR_GAUSSIAN) {
        if (ldq % 4 || ldt % 4)
            return set_error(QMLH_ERR_ARG, "%s: ldq, ldt must be multiples of 4", who);
        if (!normsQ || !normsT) return set_error(QMLH_ERR_ARG, "%s: norms missing", who);
    }
    return QMLH_OK;
}

}  // namespace qmlh

extern "C" size_t qmlh_predict_workspace(int64_t m, int64_t n) {
    return qmlh::predict_workspace_bytes(m, n);
}

extern "C" int qmlh_krr_predict(int32_t kind, const double* Xq, int64_t m, int64_t ldq,
                                const double* normsQ, const double* Xt, int64_t n, int64_t ldt,
                                const double* normsT, const double* alpha, int64_t d,
                                double sigma, double* y, void* ws, size_t ws_bytes,
                                qmlh_stream_t stream) {
    using namespace qmlh;
    QMLH_TRY(predict_check(kind, Xq, m, ldq, normsQ, Xt, n, ldt, normsT, alpha, d, y,
                           "krr_predict"));
    if (m == 0) return QMLH_OK;
    if (n > 0 && (ws_bytes < predict_workspace_bytes(m, n) || !ws))
        return set_error(QMLH_ERR_WORKSPACE, "krr_predict: workspace too small");
    return predict_impl(kind, Xq, m, ldq, normsQ, Xt, n, ldt, normsT, alpha, d, sigma, y,
                        static_cast<double*>(ws), (hipStream_t)stream);
}

extern "C" size_t qmlh_krr_predict_noalloc_capacity(void) {
    return (size_t)qmlh::PREDICT_SCRATCH_DOUBLES * sizeof(double);
}

extern "C" int qmlh_krr_predict_noalloc(int32_t kind, const double* Xq, int64_t m, int64_t ldq,
                                        const double* normsQ, const double* Xt, int64_t n,
                                        int64_t ldt, const double* normsT, const double* alpha,
                                        int64_t d, double sigma, double* y,
                                        qmlh_stream_t stream) {
    using namespace qmlh;
    QMLH_TRY(predict_check(kind, Xq, m, ldq, normsQ, Xt, n, ldt, normsT, alpha, d, y,
                           "krr_predict_noalloc"));
    if (m == 0) return QMLH_OK;
    const size_t need = predict_workspace_bytes(m, n);
    if (need > qmlh_krr_predict_noalloc_capacity())
        return set_error(QMLH_ERR_ARG,
                         "krr_predict_noalloc: %lld queries x %lld training rows need %zu bytes "
                         "of scratch, above its %zu (ceil(n / 128) * m <= %lld)",
                         (long long)m, (long long)n, need, qmlh_krr_predict_noalloc_capacity(),
                         (long long)PREDICT_SCRATCH_DOUBLES);
    double* part = nullptr;
    QMLH_TRY(predict_scratch_setup(&part));
    return predict_impl(kind, Xq, m, ldq, normsQ, Xt, n, ldt, normsT, alpha, d, sigma, y, part,
                        (hipStream_t)stream);
}
