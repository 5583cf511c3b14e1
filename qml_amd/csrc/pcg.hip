// Mixed-precision solve: FP64 preconditioned conjugate gradients on A x = y, preconditioned by
// an FP32 Cholesky factor L (B = L L^T) and, when u is given, the rank-one correction
// M = B + u u^T applied with Sherman-Morrison:
//   M^-1 r = B^-1 r - (u . B^-1 r) / (1 + u . B^-1 u) B^-1 u.
// Every residual r = y - A x is formed from the FP64 matrix, so the iteration converges to the
// FP64 solution; the FP32 factor only sets how fast. (This is iterative refinement with the
// correction equation solved by CG, i.e. CG-IR.)
//
// Reductions are deterministic: a fixed number of workgroups each reduce a fixed contiguous
// range in a fixed tree order, and one workgroup adds the partials in order. PCG scalars stay
// on the device; the host reads one residual norm per iteration to decide convergence.
#include "common.h"
#include "timer.h"

namespace qmlh {

template <typename T>
int potrs_impl(const T* L, int64_t n, int64_t lda, double* x, double* tmp, hipStream_t s);

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

// y_j = sum_i A(i, j) x_i (column dots; A symmetric so this is A x). One wavefront per
// column, lanes stride down the column, fixed-order butterfly.
__global__ __launch_bounds__(256) void symv_cols_kernel(const double* __restrict__ A, int64_t lda,
                                                        int64_t n, const double* __restrict__ x,
                                                        double* __restrict__ y) {
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n) return;
    const double* col = A + j * lda;
    double s = 0.0;
    for (int64_t i = lane; i < n; i += 64) s = fma(col[i], x[i], s);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[j] = s;
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

__global__ void sm_correct_kernel(double* __restrict__ z, const double* __restrict__ Bu,
                                  const double* __restrict__ sc, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    z[i] = fma(-(sc[4] / sc[5]), Bu[i], z[i]);
}

__global__ void den_kernel(double* sc) { sc[5] = 1.0 + sc[5]; }

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

// z = M^-1 r
static int precond(const PcgCtx& c, const float* L, int64_t ldl, const double* u,
                   const double* Bu, const double* r, double* z, double* tmp, double* sc) {
    QMLH_TRY(copy(c, r, z));
    QMLH_TRY(potrs_impl<float>(L, c.n, ldl, z, tmp, c.s));
    if (u) {
        QMLH_TRY(dot(c, u, z, sc + 4));
        hipLaunchKernelGGL(sm_correct_kernel, dim3(grid1(c.n)), dim3(256), 0, c.s, z, Bu, sc, c.n);
        QMLH_CHECK_LAUNCH("pcg correction launch");
    }
    return QMLH_OK;
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

extern "C" size_t qmlh_pcg_workspace(int64_t n) {
    if (n <= 0) return 0;
    return (size_t)(6 * n + qmlh::RED_BLOCKS + 8) * sizeof(double);
}

extern "C" int qmlh_pcg_solve(const double* A, int64_t n, int64_t lda, const float* L,
                              int64_t ldl, const double* u, const double* y, double* x,
                              double tol, int32_t max_iter, void* ws, size_t ws_bytes,
                              int32_t* iters_h, double* relres_h, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || lda < n || ldl < n) return set_error(QMLH_ERR_ARG, "pcg_solve: bad shape");
    if (iters_h) *iters_h = 0;
    if (relres_h) *relres_h = 0.0;
    if (n == 0) return QMLH_OK;
    if (!A || !L || !y || !x || !ws) return set_error(QMLH_ERR_ARG, "pcg_solve: null pointer");
    if (ws_bytes < qmlh_pcg_workspace(n))
        return set_error(QMLH_ERR_WORKSPACE, "pcg_solve: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    double* w = static_cast<double*>(ws);
    double* r = w;
    double* z = r + n;
    double* p = z + n;
    double* q = p + n;
    double* tmp = q + n;
    double* Bu = tmp + n;
    double* partial = Bu + n;
    double* sc = partial + RED_BLOCKS;  // 8 scalars
    PcgCtx c{s, n, partial};

    double yy_h = 0.0;
    QMLH_TRY(dot(c, y, y, sc + 2));
    QMLH_TRY(check_hip(hipMemcpyAsync(&yy_h, sc + 2, sizeof(double), hipMemcpyDeviceToHost, s),
                       "pcg: read |y|"));
    QMLH_TRY(check_hip(hipStreamSynchronize(s), "pcg: sync"));
    const double ynorm = sqrt(yy_h);
    hipLaunchKernelGGL(fill_kernel, dim3(grid1(n)), dim3(256), 0, s, x, 0.0, n);
    QMLH_CHECK_LAUNCH("pcg fill launch");
    if (ynorm == 0.0) return QMLH_OK;
    if (u) {  // Bu = B^-1 u, den = 1 + u.Bu
        QMLH_TRY(copy(c, u, Bu));
        QMLH_TRY(potrs_impl<float>(L, n, ldl, Bu, tmp, s));
        QMLH_TRY(dot(c, u, Bu, sc + 5));
        hipLaunchKernelGGL(den_kernel, dim3(1), dim3(1), 0, s, sc);
        QMLH_CHECK_LAUNCH("pcg den launch");
    }
    QMLH_TRY(copy(c, y, r));
    QMLH_TRY(precond(c, L, ldl, u, Bu, r, z, tmp, sc));
    QMLH_TRY(dot(c, r, z, sc + 0));
    hipLaunchKernelGGL(p_update_kernel, dim3(grid1(n)), dim3(256), 0, s, p, z, sc, n, 1);
    QMLH_CHECK_LAUNCH("pcg p launch");
    double relres = 1.0;
    int it = 0;
    while (it < max_iter) {
        ++it;
        hipLaunchKernelGGL(symv_cols_kernel, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, s, A,
                           lda, n, p, q);
        QMLH_CHECK_LAUNCH("pcg symv launch");
        QMLH_TRY(dot(c, p, q, sc + 1));
        hipLaunchKernelGGL(axpy2_kernel, dim3(grid1(n)), dim3(256), 0, s, x, r, p, q, sc, n);
        QMLH_CHECK_LAUNCH("pcg axpy launch");
        QMLH_TRY(dot(c, r, r, sc + 2));
        double rr_h = 0.0;
        QMLH_TRY(check_hip(hipMemcpyAsync(&rr_h, sc + 2, sizeof(double), hipMemcpyDeviceToHost, s),
                           "pcg: read |r|"));
        QMLH_TRY(check_hip(hipStreamSynchronize(s), "pcg: sync"));
        relres = sqrt(rr_h) / ynorm;
        if (!(relres > tol)) break;  // converged (or NaN: stop and report)
        QMLH_TRY(precond(c, L, ldl, u, Bu, r, z, tmp, sc));
        QMLH_TRY(dot(c, r, z, sc + 6));
        hipLaunchKernelGGL(p_update_kernel, dim3(grid1(n)), dim3(256), 0, s, p, z, sc, n, 0);
        QMLH_CHECK_LAUNCH("pcg p launch");
        hipLaunchKernelGGL(rz_roll_kernel, dim3(1), dim3(1), 0, s, sc);
        QMLH_CHECK_LAUNCH("pcg roll launch");
    }
    if (iters_h) *iters_h = it;
    if (relres_h) *relres_h = relres;
    return QMLH_OK;
}
