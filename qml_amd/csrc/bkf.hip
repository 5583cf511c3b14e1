// Symmetric indefinite factorisation A = L D L^T with Bunch-Kaufman pivoting (lower storage), and
// solves with it: the device counterpart of LAPACK dsytrf("L") / dsytrs("L"), which the reference
// calls in fbkf_solve and fbkf_invert (qml/fcho_solve.f90:76-150). LAPACK is a third-party
// dependency of the reference; this file restates its published unblocked algorithm (dsytf2):
//   alpha = (1 + sqrt(17)) / 8; for column k: colmax = max_{i>k} |A(i,k)| (first index imax);
//   1 x 1 pivot without interchange if |A(k,k)| >= alpha colmax, or if
//   |A(k,k)| >= alpha colmax (colmax / rowmax) with rowmax the largest off-diagonal entry of
//   row/column imax; 1 x 1 pivot imax if |A(imax,imax)| >= alpha rowmax; else a 2 x 2 pivot on
//   (k, imax) with imax moved to k + 1. Interchanges act on the trailing matrix only; ipiv holds
//   1-based pivots, negative for both columns of a 2 x 2 block; info is the first zero pivot.
//
// Below order 256 the factorisation is one workgroup of 1024 lanes walking the columns (the pivot
// search of each column depends on every earlier update), the trailing rank-1 / rank-2 updates
// spread over its 16 waves; from 256 on each column is a step of three launches whose update
// covers the GPU. The solves run one workgroup per right-hand-side column, so an inverse (n
// right-hand sides) fills the GPU.
#include "common.h"
#include "timer.h"

#pragma clang fp contract(off)

namespace qmlh {

static constexpr int BK_THREADS = 1024;
static constexpr int BK_WAVES = BK_THREADS / 64;

// Block-wide argmax of v over lanes (first index on ties); every thread gets the result.
__device__ __forceinline__ void bk_argmax(double v, int64_t i, double* s_v, int64_t* s_i,
                                          double& out_v, int64_t& out_i) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int64_t oi = __shfl_xor(i, off, 64);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
    if (lane == 0) {
        s_v[w] = v;
        s_i[w] = i;
    }
    __syncthreads();
    double bv = s_v[0];
    int64_t bi = s_i[0];
    for (int q = 1; q < BK_WAVES; ++q) {
        const double ov = s_v[q];
        const int64_t oi = s_i[q];
        if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    __syncthreads();
    out_v = bv;
    out_i = bi;
}

// Pivot choice and interchange for column k (one workgroup): returns kstep (1 or 2) and kp
// (0-based), zero = a zero column (no pivot, no update). Every thread gets the same values.
__device__ __forceinline__ void bk_pivot(double* __restrict__ A, int64_t n, int64_t lda, int64_t k,
                                         double* s_v, int64_t* s_i, int& kstep, int64_t& kp,
                                         bool& zero) {
    const int t = threadIdx.x;
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    auto at = [&](int64_t i, int64_t j) -> double& { return A[i + j * lda]; };
    kstep = 1;
    kp = k;
    const double absakk = fabs(at(k, k));
    double colmax = 0.0;
    int64_t imax = k;
    if (k + 1 < n) {
        double v = -1.0;
        int64_t vi = n;
        for (int64_t i = k + 1 + t; i < n; i += BK_THREADS) {
            const double a = fabs(at(i, k));
            if (a > v) {
                v = a;
                vi = i;
            }
        }
        bk_argmax(v, vi, s_v, s_i, colmax, imax);
    }
    zero = (absakk > colmax ? absakk : colmax) == 0.0;
    if (zero) return;
    if (absakk < alpha * colmax) {
        // rowmax: largest |entry| of row imax in columns k .. imax-1, then of column imax below
        // the diagonal.
        double v = -1.0;
        int64_t vi = n;
        for (int64_t j = k + t; j < imax; j += BK_THREADS) {
            const double a = fabs(at(imax, j));
            if (a > v) {
                v = a;
                vi = j;
            }
        }
        for (int64_t i = imax + 1 + t; i < n; i += BK_THREADS) {
            const double a = fabs(at(i, imax));
            if (a > v) {
                v = a;
                vi = i;
            }
        }
        double rowmax;
        int64_t jmax;
        bk_argmax(v, vi, s_v, s_i, rowmax, jmax);
        if (absakk >= alpha * colmax * (colmax / rowmax)) {
            kp = k;
        } else if (fabs(at(imax, imax)) >= alpha * rowmax) {
            kp = imax;
        } else {
            kp = imax;
            kstep = 2;
        }
    }
    const int64_t kk = k + kstep - 1;
    if (kp != kk) {
        // Interchange rows and columns kk and kp of the trailing lower triangle.
        for (int64_t i = kp + 1 + t; i < n; i += BK_THREADS) {
            const double x = at(i, kk);
            at(i, kk) = at(i, kp);
            at(i, kp) = x;
        }
        for (int64_t j = kk + 1 + t; j < kp; j += BK_THREADS) {
            const double x = at(j, kk);
            at(j, kk) = at(kp, j);
            at(kp, j) = x;
        }
        if (t == 0) {
            const double x = at(kk, kk);
            at(kk, kk) = at(kp, kp);
            at(kp, kp) = x;
            if (kstep == 2) {
                const double y = at(k + 1, k);
                at(k + 1, k) = at(kp, k);
                at(kp, k) = y;
            }
        }
        __syncthreads();
    }
}

// The 2 x 2 pivot's multipliers of rows j >= k + 2 from the old columns k, k + 1 (into wsp).
__device__ __forceinline__ void bk_multipliers(const double* __restrict__ A, int64_t n,
                                               int64_t lda, int64_t k, double* __restrict__ wsp) {
    auto at = [&](int64_t i, int64_t j) -> double { return A[i + j * lda]; };
    double d21 = at(k + 1, k);
    const double d11 = at(k + 1, k + 1) / d21;
    const double d22 = at(k, k) / d21;
    const double tt = 1.0 / (d11 * d22 - 1.0);
    d21 = tt / d21;
    for (int64_t j = k + 2 + threadIdx.x; j < n; j += blockDim.x) {
        const double ajk = at(j, k), ajk1 = at(j, k + 1);
        wsp[2 * j] = d21 * (d11 * ajk - ajk1);
        wsp[2 * j + 1] = d21 * (d22 * ajk1 - ajk);
    }
}

// Whole factorisation in one workgroup (small orders).
__global__ __launch_bounds__(BK_THREADS) void bkf_factor_kernel(double* __restrict__ A, int64_t n,
                                                                 int64_t lda,
                                                                 int32_t* __restrict__ ipiv,
                                                                 int32_t* __restrict__ info,
                                                                 double* __restrict__ wsp) {
    __shared__ double s_v[BK_WAVES];
    __shared__ int64_t s_i[BK_WAVES];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int32_t first_zero = 0;
    auto at = [&](int64_t i, int64_t j) -> double& { return A[i + j * lda]; };

    int64_t k = 0;
    while (k < n) {
        int kstep;
        int64_t kp;
        bool zero;
        bk_pivot(A, n, lda, k, s_v, s_i, kstep, kp, zero);
        if (zero) {
            if (first_zero == 0) first_zero = (int32_t)(k + 1);
        } else if (kstep == 1) {
            // A22 := A22 - x x^T / d (lower), then column k := x / d (the dsyr + dscal pair).
            const double r1 = 1.0 / at(k, k);
            for (int64_t j = k + 1 + w; j < n; j += BK_WAVES) {
                const double xj = at(j, k);
                if (xj != 0.0) {
                    const double temp = -r1 * xj;
                    for (int64_t i = j + lane; i < n; i += 64) at(i, j) = at(i, j) + at(i, k) * temp;
                }
            }
            __syncthreads();
            for (int64_t i = k + 1 + t; i < n; i += BK_THREADS) at(i, k) = r1 * at(i, k);
        } else if (k + 2 < n) {
            // 2 x 2 pivot (k, k+1): multipliers from the old columns, the rank-2 update with
            // them, then the columns replaced by the multipliers.
            bk_multipliers(A, n, lda, k, wsp);
            __syncthreads();
            for (int64_t j = k + 2 + w; j < n; j += BK_WAVES) {
                const double wk = wsp[2 * j], wkp1 = wsp[2 * j + 1];
                for (int64_t i = j + lane; i < n; i += 64)
                    at(i, j) = at(i, j) - at(i, k) * wk - at(i, k + 1) * wkp1;
            }
            __syncthreads();
            for (int64_t j = k + 2 + t; j < n; j += BK_THREADS) {
                at(j, k) = wsp[2 * j];
                at(j, k + 1) = wsp[2 * j + 1];
            }
        }
        if (t == 0) {
            if (kstep == 1) {
                ipiv[k] = (int32_t)(kp + 1);
            } else {
                ipiv[k] = -(int32_t)(kp + 1);
                ipiv[k + 1] = -(int32_t)(kp + 1);
            }
        }
        __syncthreads();
        k += kstep;
    }
    if (t == 0) *info = first_zero;
}

// ------------------------------------------------------------------------------------------
// Larger orders: one step per column, the same arithmetic in three launches -- the pivot and
// interchange (one workgroup), the rank-1 / rank-2 update of the trailing triangle (64 x 64
// tiles, many workgroups), and the scaled or replaced pivot columns. The current column lives
// in device memory (a 2 x 2 pivot consumes two steps, so step s is idle when the column
// counter has passed s); the host enqueues n steps without reading anything back.
// ------------------------------------------------------------------------------------------
struct BkState {
    int64_t k;       // next column to factor
    int64_t step_k;  // column of the current step (-1: idle step)
    int32_t kstep;
    int32_t zero;
    double r1;       // 1 / D(k, k) of a 1 x 1 pivot
};

__global__ void bkf_init_kernel(BkState* st, int32_t* info) {
    st->k = 0;
    st->step_k = -1;
    *info = 0;
}

__global__ __launch_bounds__(BK_THREADS) void bkf_pivot_kernel(double* __restrict__ A, int64_t n,
                                                                int64_t lda,
                                                                int32_t* __restrict__ ipiv,
                                                                int32_t* __restrict__ info,
                                                                BkState* __restrict__ st,
                                                                double* __restrict__ wsp,
                                                                int64_t s) {
    __shared__ double s_v[BK_WAVES];
    __shared__ int64_t s_i[BK_WAVES];
    const int t = threadIdx.x;
    const int64_t k = st->k;
    if (k != s) {  // consumed by the previous step's 2 x 2 pivot
        __syncthreads();
        if (t == 0) st->step_k = -1;
        return;
    }
    int kstep;
    int64_t kp;
    bool zero;
    bk_pivot(A, n, lda, k, s_v, s_i, kstep, kp, zero);
    if (!zero && kstep == 2 && k + 2 < n) bk_multipliers(A, n, lda, k, wsp);
    __syncthreads();
    if (t == 0) {
        if (zero && *info == 0) *info = (int32_t)(k + 1);
        if (kstep == 1) {
            ipiv[k] = (int32_t)(kp + 1);
        } else {
            ipiv[k] = -(int32_t)(kp + 1);
            ipiv[k + 1] = -(int32_t)(kp + 1);
        }
        st->step_k = k;
        st->kstep = kstep;
        st->zero = zero ? 1 : 0;
        st->r1 = zero ? 0.0 : 1.0 / A[k + k * lda];
        st->k = k + kstep;
    }
}

// Trailing update of step s: tile (ti, tj), ti >= tj, of the triangle rows/columns s+1 .. n-1.
__global__ __launch_bounds__(256) void bkf_update_kernel(double* __restrict__ A, int64_t n,
                                                          int64_t lda,
                                                          const BkState* __restrict__ st,
                                                          const double* __restrict__ wsp,
                                                          int64_t s) {
    const int64_t k = st->step_k;
    if (k < 0 || st->zero) return;
    const int kstep = st->kstep;
    const int64_t j0 = k + kstep;  // first trailing column
    if (j0 >= n) return;
    const int64_t bid = blockIdx.x;
    int64_t tj = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while (tj * (tj + 1) / 2 > bid) --tj;
    while ((tj + 1) * (tj + 2) / 2 <= bid) ++tj;
    const int64_t ti = bid - tj * (tj + 1) / 2;  // ti <= tj: row tile tj, column tile ti
    const int64_t r0 = s + 1 + 64 * tj, c0 = s + 1 + 64 * ti;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t i = r0 + lane;
    auto at = [&](int64_t a, int64_t b) -> double& { return A[a + b * lda]; };
    if (i >= n) return;
    if (kstep == 1) {
        const double r1 = st->r1;
        const double aik = at(i, k);
        for (int q = w; q < 64; q += 4) {
            const int64_t j = c0 + q;
            if (j < j0 || j > i) continue;
            const double xj = at(j, k);
            if (xj != 0.0) {
                const double temp = -r1 * xj;
                at(i, j) = at(i, j) + aik * temp;
            }
        }
    } else {
        const double aik = at(i, k), aik1 = at(i, k + 1);
        for (int q = w; q < 64; q += 4) {
            const int64_t j = c0 + q;
            if (j < j0 || j > i) continue;
            at(i, j) = at(i, j) - aik * wsp[2 * j] - aik1 * wsp[2 * j + 1];
        }
    }
}

// The pivot columns of step s: scaled by 1 / D (1 x 1) or replaced by the multipliers (2 x 2).
__global__ __launch_bounds__(256) void bkf_finish_kernel(double* __restrict__ A, int64_t n,
                                                          int64_t lda,
                                                          const BkState* __restrict__ st,
                                                          const double* __restrict__ wsp) {
    const int64_t k = st->step_k;
    if (k < 0 || st->zero) return;
    const int64_t i = k + 1 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (st->kstep == 1) {
        if (i < n) A[i + k * lda] = st->r1 * A[i + k * lda];
    } else {
        const int64_t j = i + 1;  // rows k + 2 ..
        if (j < n) {
            A[j + k * lda] = wsp[2 * j];
            A[j + (k + 1) * lda] = wsp[2 * j + 1];
        }
    }
}

// Block-wide sum of v (fixed order: waves in order), every thread gets the result.
template <int NT>
__device__ __forceinline__ double bk_sum(double v, double* s_v) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) s_v[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < NT / 64; ++q) s += s_v[q];
    __syncthreads();
    return s;
}

// One workgroup per right-hand side column: B(:, c) := A^-1 B(:, c) with the factors of
// bkf_factor_kernel (the dsytrs("L") sequence: L D with the interchanges forward, then L^T
// backward).
static constexpr int BS_THREADS = 256;

__global__ __launch_bounds__(BS_THREADS) void bkf_solve_kernel(const double* __restrict__ A,
                                                                int64_t n, int64_t lda,
                                                                const int32_t* __restrict__ ipiv,
                                                                double* __restrict__ Bm,
                                                                int64_t ldb) {
    __shared__ double s_v[BS_THREADS / 64];
    double* b = Bm + (int64_t)blockIdx.x * ldb;
    const int t = threadIdx.x;
    auto at = [&](int64_t i, int64_t j) -> double { return A[i + j * lda]; };
    int64_t k = 0;
    while (k < n) {
        if (ipiv[k] > 0) {
            const int64_t kp = ipiv[k] - 1;
            if (t == 0 && kp != k) {
                const double x = b[k];
                b[k] = b[kp];
                b[kp] = x;
            }
            __syncthreads();
            const double bk = b[k];
            for (int64_t i = k + 1 + t; i < n; i += BS_THREADS) b[i] = b[i] - at(i, k) * bk;
            __syncthreads();
            if (t == 0) b[k] = b[k] * (1.0 / at(k, k));  // dscal by the reciprocal
            __syncthreads();
            k += 1;
        } else {
            const int64_t kp = -ipiv[k] - 1;
            if (t == 0 && kp != k + 1) {
                const double x = b[k + 1];
                b[k + 1] = b[kp];
                b[kp] = x;
            }
            __syncthreads();
            const double bk = b[k], bk1 = b[k + 1];
            for (int64_t i = k + 2 + t; i < n; i += BS_THREADS)
                b[i] = (b[i] - at(i, k) * bk) - at(i, k + 1) * bk1;
            __syncthreads();
            if (t == 0) {
                const double akm1k = at(k + 1, k);
                const double akm1 = at(k, k) / akm1k;
                const double ak = at(k + 1, k + 1) / akm1k;
                const double denom = akm1 * ak - 1.0;
                const double bkm1 = b[k] / akm1k;
                const double bkk = b[k + 1] / akm1k;
                b[k] = (ak * bkm1 - bkk) / denom;
                b[k + 1] = (akm1 * bkk - bkm1) / denom;
            }
            __syncthreads();
            k += 2;
        }
    }
    k = n - 1;
    while (k >= 0) {
        if (ipiv[k] > 0) {
            double part = 0.0;
            for (int64_t i = k + 1 + t; i < n; i += BS_THREADS) part += at(i, k) * b[i];
            const double dot = bk_sum<BS_THREADS>(part, s_v);
            if (t == 0) {
                b[k] = b[k] - dot;
                const int64_t kp = ipiv[k] - 1;
                if (kp != k) {
                    const double x = b[k];
                    b[k] = b[kp];
                    b[kp] = x;
                }
            }
            __syncthreads();
            k -= 1;
        } else {
            // 2 x 2 block (k-1, k)
            double p0 = 0.0, p1 = 0.0;
            for (int64_t i = k + 1 + t; i < n; i += BS_THREADS) {
                p0 += at(i, k) * b[i];
                p1 += at(i, k - 1) * b[i];
            }
            const double d0 = bk_sum<BS_THREADS>(p0, s_v);
            const double d1 = bk_sum<BS_THREADS>(p1, s_v);
            if (t == 0) {
                b[k] = b[k] - d0;
                b[k - 1] = b[k - 1] - d1;
                const int64_t kp = -ipiv[k] - 1;
                if (kp != k) {
                    const double x = b[k];
                    b[k] = b[kp];
                    b[kp] = x;
                }
            }
            __syncthreads();
            k -= 2;
        }
    }
}

}  // namespace qmlh

// Orders from which the stepped (many-workgroup) factorisation is used.
static constexpr int64_t BK_STEPPED_MIN = 256;

extern "C" size_t qmlh_bkf_workspace(int64_t n) {
    return n > 0 ? (size_t)(2 * n) * sizeof(double) + 256 : 0;
}

extern "C" int qmlh_bkf_factor(double* A, int64_t n, int64_t lda, int32_t* ipiv, int32_t* info,
                               void* ws, size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || lda < (n > 1 ? n : 1)) return set_error(QMLH_ERR_ARG, "bkf_factor: n / lda");
    if (n == 0) return QMLH_OK;
    if (n > 0x7fffffffLL - 1) return set_error(QMLH_ERR_ARG, "bkf_factor: n too large");
    if (!A || !ipiv || !info) return set_error(QMLH_ERR_ARG, "bkf_factor: null pointer");
    if (!ws || ws_bytes < qmlh_bkf_workspace(n))
        return set_error(QMLH_ERR_WORKSPACE, "bkf_factor: workspace too small");
    if (reinterpret_cast<uintptr_t>(ws) & 15)
        return set_error(QMLH_ERR_ARG, "bkf_factor: workspace must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    TimerScope ts(QMLH_TIMER_POTRF, (double)n * (double)n * (double)n / 3.0, s);
    BkState* st = static_cast<BkState*>(ws);
    double* wsp = reinterpret_cast<double*>(static_cast<char*>(ws) + 256);
    if (n < BK_STEPPED_MIN) {
        hipLaunchKernelGGL(bkf_factor_kernel, dim3(1), dim3(BK_THREADS), 0, s, A, n, lda, ipiv,
                           info, wsp);
        QMLH_CHECK_LAUNCH("bkf_factor launch");
        return QMLH_OK;
    }
    hipLaunchKernelGGL(bkf_init_kernel, dim3(1), dim3(1), 0, s, st, info);
    QMLH_CHECK_LAUNCH("bkf_factor init launch");
    for (int64_t step = 0; step < n; ++step) {
        hipLaunchKernelGGL(bkf_pivot_kernel, dim3(1), dim3(BK_THREADS), 0, s, A, n, lda, ipiv,
                           info, st, wsp, step);
        const int64_t m = n - step - 1;  // trailing order bound of this step
        if (m > 0) {
            const int64_t tiles = ceil_div(m, 64);
            hipLaunchKernelGGL(bkf_update_kernel, dim3((unsigned)(tiles * (tiles + 1) / 2)),
                               dim3(256), 0, s, A, n, lda, st, wsp, step);
            hipLaunchKernelGGL(bkf_finish_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0,
                               s, A, n, lda, st, wsp);
        }
    }
    QMLH_CHECK_LAUNCH("bkf_factor step launches");
    return QMLH_OK;
}

extern "C" int qmlh_bkf_solve(const double* A, int64_t n, int64_t lda, const int32_t* ipiv,
                              double* B, int64_t ldb, int64_t nrhs, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || nrhs < 0 || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1))
        return set_error(QMLH_ERR_ARG, "bkf_solve: sizes");
    if (n == 0 || nrhs == 0) return QMLH_OK;
    if (!A || !ipiv || !B) return set_error(QMLH_ERR_ARG, "bkf_solve: null pointer");
    if (nrhs > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "bkf_solve: too many columns");
    hipLaunchKernelGGL(bkf_solve_kernel, dim3((unsigned)nrhs), dim3(BS_THREADS), 0,
                       (hipStream_t)stream, A, n, lda, ipiv, B, ldb);
    QMLH_CHECK_LAUNCH("bkf_solve launch");
    return QMLH_OK;
}
