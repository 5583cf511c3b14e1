// Symmetric indefinite factorisation A = L D L^T with Bunch-Kaufman pivoting (lower storage), and
// solves with it: the device counterpart of LAPACK dsytrf("L") / dsytrs("L"), which the reference
// calls in fbkf_solve and fbkf_invert (qml/fcho_solve.f90:76-150). LAPACK is a third-party
// dependency of the reference; this file restates its published algorithms, the unblocked
// dsytf2 and the blocked dsytrf / dlasyf the reference's call runs (lwork = n nb). Pivoting:
//   alpha = (1 + sqrt(17)) / 8; for column k: colmax = max_{i>k} |A(i,k)| (first index imax);
//   1 x 1 pivot without interchange if |A(k,k)| >= alpha colmax, or if
//   |A(k,k)| >= alpha colmax (colmax / rowmax) with rowmax the largest off-diagonal entry of
//   row/column imax; 1 x 1 pivot imax if |A(imax,imax)| >= alpha rowmax; else a 2 x 2 pivot on
//   (k, imax) with imax moved to k + 1. Interchanges act on the trailing matrix only; ipiv holds
//   1-based pivots, negative for both columns of a 2 x 2 block; info is the first zero pivot.
//
// Below order 256 the factorisation is one workgroup of 1024 lanes walking the columns (the pivot
// search of each column depends on every earlier update), the trailing rank-1 / rank-2 updates
// spread over its 16 waves; from 256 on it is blocked as LAPACK's dsytrf is (64-column dlasyf
// panels, an FP64 MFMA trailing update per panel). The solves run one workgroup per
// right-hand-side column, so an inverse (n right-hand sides) fills the GPU.
#include "common.h"
#include "mfma_core.h"
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

// dsytf2 on the trailing submatrix A(k_begin:n, k_begin:n), in global indices (1-based global
// pivots): all BK_THREADS lanes of one workgroup call it. Returns the first zero pivot column
// + 1, or 0.
__device__ int32_t bk_unblocked(double* __restrict__ A, int64_t n, int64_t lda, int64_t k_begin,
                                int32_t* __restrict__ ipiv, double* __restrict__ wsp,
                                double* s_v, int64_t* s_i) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int32_t first_zero = 0;
    auto at = [&](int64_t i, int64_t j) -> double& { return A[i + j * lda]; };

    int64_t k = k_begin;
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
    return first_zero;
}

// Whole factorisation in one workgroup (small orders).
__global__ __launch_bounds__(BK_THREADS) void bkf_factor_kernel(double* __restrict__ A, int64_t n,
                                                                 int64_t lda,
                                                                 int32_t* __restrict__ ipiv,
                                                                 int32_t* __restrict__ info,
                                                                 double* __restrict__ wsp) {
    __shared__ double s_v[BK_WAVES];
    __shared__ int64_t s_i[BK_WAVES];
    const int32_t first_zero = bk_unblocked(A, n, lda, 0, ipiv, wsp, s_v, s_i);
    if (threadIdx.x == 0) *info = first_zero;
}

// ------------------------------------------------------------------------------------------
// Blocked factorisation (orders from BKB_MIN): LAPACK's dsytrf("L") with dlasyf("L") panels,
// restated. While more than BKB_NB columns remain, a panel of up to BKB_NB columns is factored
// column by column without touching the trailing matrix: the updated column k is
//   W(k:n, j) = A(k:n, k) - A(k:n, k0:k-1) W(k, 0:j-1)^T        (j = k - k0)
// (and, when the pivot test needs it, the updated column imax into W(:, j + 1)); the pivot test,
// the interchange and the stored L column(s) are those of the unblocked algorithm. The panel
// stops once j >= BKB_NB - 1 (so it holds BKB_NB - 1 or, when the last pivot is 2 x 2, BKB_NB
// columns: kb); then the trailing lower triangle gets A22 -= L21 W21^T as one FP64 MFMA launch,
// and the interchanges of the panel's later steps are undone in its earlier columns (the
// unblocked algorithm's storage convention). The last <= BKB_NB columns run unblocked.
//
// Each panel step is two launches over the rows: (1) the L column(s) of the previous step are
// stored (scaled, or 2 x 2 multipliers), then W(:, j) and the per-workgroup argmax; (2) the
// column imax and its argmax, when needed, then the pivot choice and the interchange (O(n)
// element moves) in the last workgroup. Decisions are made by the last workgroup to finish a
// launch (a ticket counter: every other workgroup is done, so it may rewrite the shared
// state). The host does not know where the panels start (kb is 63 or 64),
// so it enqueues the most steps and panels the order can need; steps after a panel is complete,
// and panels after the last blocked one, return at once.
// ------------------------------------------------------------------------------------------
static constexpr int BKB_NB = 64;     // panel width (LAPACK's block size for dsytrf)
static constexpr int BKB_ROWS = 256;  // rows per workgroup of the store launch
static constexpr int BKB_RB = 64;     // rows per workgroup of the column launches (4 lane groups
                                      // split the panel columns of each row)

struct BkbState {
    int64_t k;        // next column to factor
    int64_t k0;       // first column of the current panel
    int64_t imax;     // row of the column maximum of the current step
    int64_t kp;       // pivot row of the current step
    int64_t sk;       // first column of the step still to be stored (-1: none)
    double colmax, absakk;
    uint32_t ticket;  // last-workgroup counter
    int32_t j;        // k - k0
    int32_t active;   // the current panel is a blocked one
    int32_t done;     // the current panel has all its columns
    int32_t need2;    // the step needs the updated column imax
    int32_t kstep, zero;
    int32_t sj, skstep, szero;
};

struct BkbPart {
    double v;
    int64_t i;
};

template <int NW>
__device__ __forceinline__ void bkb_argmax(double& v, int64_t& i, double* s_v, int64_t* s_i) {
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
    v = s_v[0];
    i = s_i[0];
    for (int q = 1; q < NW; ++q) {
        if (s_v[q] > v || (s_v[q] == v && s_i[q] < i)) {
            v = s_v[q];
            i = s_i[q];
        }
    }
    __syncthreads();
}

// True in the workgroup that finishes last; the stores of every workgroup are visible to it.
// Resets the ticket for the next launch.
__device__ __forceinline__ bool bkb_last(uint32_t* ticket) {
    __shared__ int s_last;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t =
            __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = t == gridDim.x - 1;
        if (s_last) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (s_last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return s_last != 0;
}

// Reduction of the per-workgroup argmax partials (first index on ties), in the last workgroup.
__device__ __forceinline__ void bkb_reduce(const BkbPart* part, int np, double& v, int64_t& vi,
                                           double* s_v, int64_t* s_i) {
    v = -1.0;
    vi = INT64_MAX;
    for (int b = threadIdx.x; b < np; b += blockDim.x) {
        const BkbPart p = part[b];
        if (p.v > v || (p.v == v && p.i < vi)) {
            v = p.v;
            vi = p.i;
        }
    }
    bkb_argmax<BKB_ROWS / 64>(v, vi, s_v, s_i);
}

// Stores row i of the L column(s) of step (sk, sj, skstep) from W (dlasyf's copy + dscal, or
// the 2 x 2 multipliers and D block; a zero column is copied unscaled).
__device__ __forceinline__ void bkb_store_row(double* __restrict__ A, int64_t n, int64_t lda,
                                              const double* __restrict__ W, int64_t ldw,
                                              const BkbState& S, int64_t i) {
    const int64_t k = S.sk;
    const double* w0 = W + (int64_t)S.sj * ldw;
    if (i < k || i >= n) return;
    if (S.skstep == 1) {
        if (S.szero || i == k) {
            A[i + k * lda] = w0[i];
        } else {
            const double r1 = 1.0 / w0[k];
            A[i + k * lda] = r1 * w0[i];
        }
        return;
    }
    const double* w1 = w0 + ldw;
    if (i == k) {
        A[k + k * lda] = w0[k];
    } else if (i == k + 1) {
        A[k + 1 + k * lda] = w0[k + 1];
        A[k + 1 + (k + 1) * lda] = w1[k + 1];
    } else {
        double d21 = w0[k + 1];
        const double d11 = w1[k + 1] / d21;
        const double d22 = w0[k] / d21;
        const double t = 1.0 / (d11 * d22 - 1.0);
        d21 = t / d21;
        A[i + k * lda] = d21 * (d11 * w0[i] - w1[i]);
        A[i + (k + 1) * lda] = d21 * (d22 * w1[i] - w0[i]);
    }
}

__global__ void bkb_init_kernel(BkbState* st, int32_t* info) {
    st->k = 0;
    st->sk = -1;
    st->ticket = 0;
    st->active = 0;
    st->done = 1;
    *info = 0;
}

__global__ void bkb_begin_kernel(BkbState* st, int64_t n) {
    const int64_t k = st->k;
    st->active = n - k >= BKB_NB + 1;
    st->k0 = k;
    st->j = 0;
    st->done = 0;
    st->need2 = 0;
    st->sk = -1;
}

// sum over q in [q0, q1) of a[q lda] w[q ldw], in order, with the loads of 8 terms issued
// before their products (the terms are independent loads; the sum stays sequential).
__device__ __forceinline__ double bkb_dot(const double* __restrict__ a, int64_t lda,
                                          const double* __restrict__ w, int64_t ldw, int q0,
                                          int q1) {
    double s = 0.0;
    int q = q0;
    for (; q + 8 <= q1; q += 8) {
        double x[8], y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x[u] = a[(int64_t)(q + u) * lda];
            y[u] = w[(int64_t)(q + u) * ldw];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s = s + x[u] * y[u];
    }
    for (; q < q1; ++q) s = s + a[(int64_t)q * lda] * w[(int64_t)q * ldw];
    return s;
}

// x(i) = base - sum_{q < j} A(i, k0 + q) W(wrow, q) for the 64 rows of this workgroup: lane
// group g (4 of them) takes panel columns [g j / 4, (g + 1) j / 4), the partials are added in a
// fixed order. Returns x in the lanes of group 0 (other lanes: 0).
__device__ __forceinline__ double bkb_update_rows(const double* __restrict__ A, int64_t lda,
                                                  const double* __restrict__ W, int64_t ldw,
                                                  int64_t k0, int j, int64_t wrow, int64_t i,
                                                  bool row_ok, double base,
                                                  double (*red)[BKB_RB]) {
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int q0 = (j * g) / 4, q1 = (j * (g + 1)) / 4;
    double p = 0.0;
    if (row_ok && q1 > q0) p = bkb_dot(A + i + (k0 + q0) * lda, lda, W + wrow + (int64_t)q0 * ldw,
                                       ldw, 0, q1 - q0);
    red[g][r] = p;
    __syncthreads();
    double x = 0.0;
    if (g == 0) x = base - ((red[0][r] + red[1][r]) + (red[2][r] + red[3][r]));
    return x;
}

// Argmax over the 64 lanes of wave 0 (first index on ties).
__device__ __forceinline__ void bkb_wave_argmax(double& v, int64_t& i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int64_t oi = __shfl_xor(i, off, 64);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
}

// (1) store the previous step's column(s), then W(:, j) and colmax / imax.
__global__ __launch_bounds__(256) void bkb_column_kernel(double* __restrict__ A, int64_t n,
                                                         int64_t lda, double* __restrict__ W,
                                                         int64_t ldw, BkbState* st,
                                                         BkbPart* __restrict__ part,
                                                         int32_t* __restrict__ info) {
    __shared__ double red[4][BKB_RB];
    __shared__ double s_v[4];
    __shared__ int64_t s_i[4];
    const BkbState S = *st;
    if (!S.active || S.done) return;
    const int64_t k = S.k, k0 = S.k0;
    const int j = S.j;
    const int64_t lo = S.sk >= 0 ? S.sk : k;
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t i = lo + (int64_t)blockIdx.x * BKB_RB + r;
    if (S.sk >= 0 && g == 0) bkb_store_row(A, n, lda, W, ldw, S, i);
    __syncthreads();  // the stored column is read by every lane group below
    const bool row_ok = i >= k && i < n;
    const double x = bkb_update_rows(A, lda, W, ldw, k0, j, k, i, row_ok,
                                     row_ok ? A[i + k * lda] : 0.0, red);
    if (g == 0) {
        double v = -1.0;
        int64_t vi = INT64_MAX;
        if (row_ok) {
            W[i + (int64_t)j * ldw] = x;
            if (i > k) {
                v = fabs(x);
                vi = i;
            }
        }
        bkb_wave_argmax(v, vi);
        if (r == 0) part[blockIdx.x] = BkbPart{v, vi};
    }
    if (!bkb_last(&st->ticket)) return;
    double colmax;
    int64_t imax;
    bkb_reduce(part, gridDim.x, colmax, imax, s_v, s_i);
    if (threadIdx.x != 0) return;
    if (colmax < 0.0) colmax = 0.0;  // no rows below k
    const double absakk = fabs(W[k + (int64_t)j * ldw]);
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    st->sk = -1;
    st->colmax = colmax;
    st->absakk = absakk;
    st->imax = imax;
    if ((absakk > colmax ? absakk : colmax) == 0.0) {
        // zero (or underflowed) column: no pivot, recorded in info
        st->zero = 1;
        st->kp = k;
        st->kstep = 1;
        st->need2 = 0;
        if (*info == 0) *info = (int32_t)(k + 1);
    } else {
        st->zero = 0;
        st->kp = k;
        st->kstep = 1;
        st->need2 = absakk >= alpha * colmax ? 0 : 1;
    }
}

// (2) the updated column imax into W(:, j + 1) and rowmax (when the step needs them), then, in
// the last workgroup, the pivot choice and the interchange of rows / columns kk = k + kstep - 1
// and kp: the non-updated column kk copied to column kp of the trailing matrix (column kk is
// overwritten by the stored L column), rows kk and kp swapped in the panel's earlier columns of
// A and in W(:, 0..jj); then the pivots are recorded and the step advances.
__global__ __launch_bounds__(256) void bkb_pivot_kernel(double* __restrict__ A, int64_t n,
                                                        int64_t lda, double* __restrict__ W,
                                                        int64_t ldw, BkbState* st,
                                                        BkbPart* __restrict__ part,
                                                        int32_t* __restrict__ ipiv) {
    __shared__ double red[4][BKB_RB];
    __shared__ double s_v[4];
    __shared__ int64_t s_i[4];
    const BkbState S = *st;
    if (!S.active || S.done) return;  // uniform: no workgroup takes a ticket
    const int64_t k = S.k, k0 = S.k0, imax = S.imax;
    const int j = S.j;
    if (S.need2) {
        const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
        const int64_t i = k + (int64_t)blockIdx.x * BKB_RB + r;
        const bool row_ok = i < n;
        // row imax left of the diagonal, then column imax
        const double base = row_ok ? (i < imax ? A[imax + i * lda] : A[i + imax * lda]) : 0.0;
        const double x = bkb_update_rows(A, lda, W, ldw, k0, j, imax, i, row_ok, base, red);
        if (g == 0) {
            double v = -1.0;
            int64_t vi = INT64_MAX;
            if (row_ok) {
                W[i + (int64_t)(j + 1) * ldw] = x;
                if (i != imax) {
                    v = fabs(x);
                    vi = i;
                }
            }
            bkb_wave_argmax(v, vi);
            if (r == 0) part[blockIdx.x] = BkbPart{v, vi};
        }
    }
    if (!bkb_last(&st->ticket)) return;
    int64_t kp = S.kp;
    int kstep = S.kstep;
    bool copy = false;
    if (S.need2) {
        double rowmax;
        int64_t jmax;
        bkb_reduce(part, gridDim.x, rowmax, jmax, s_v, s_i);
        const double alpha = (1.0 + sqrt(17.0)) / 8.0;
        if (S.absakk >= alpha * S.colmax * (S.colmax / rowmax)) {
            kp = k;
            kstep = 1;
        } else if (fabs(W[imax + (int64_t)(j + 1) * ldw]) >= alpha * rowmax) {
            kp = imax;
            kstep = 1;
            copy = true;  // W(:, j) := W(:, j + 1)
        } else {
            kp = imax;
            kstep = 2;
        }
    }
    const int64_t kk = k + kstep - 1;
    const int jj = j + kstep - 1;
    const int t = threadIdx.x;
    if (!S.zero) {
        // Disjoint element sets: the row loop writes W(i, j) for i other than kk, kp, row kp
        // right of column kk, column kp and A(kp, kp); the lane loop rows kk and kp of W and of
        // the columns left of kk.
        for (int64_t i = k + t; i < n; i += 256) {
            if (copy && i != kk && i != kp) W[i + (int64_t)j * ldw] = W[i + (int64_t)(j + 1) * ldw];
            if (kp != kk) {
                if (i > kk && i < kp) {
                    A[kp + i * lda] = A[i + kk * lda];
                } else if (i > kp) {
                    A[i + kp * lda] = A[i + kk * lda];
                } else if (i == kp) {
                    A[kp + kp * lda] = A[kk + kk * lda];
                }
            }
        }
        if (kp != kk && t < BKB_NB) {
            const int q = t;
            if (q <= jj) {
                const int src = (copy && q == j) ? j + 1 : q;
                const double a = W[kk + (int64_t)src * ldw], b = W[kp + (int64_t)src * ldw];
                W[kk + (int64_t)q * ldw] = b;
                W[kp + (int64_t)q * ldw] = a;
            }
            const int64_t c = k0 + q;
            if (c < kk) {
                const double a = A[kk + c * lda], b = A[kp + c * lda];
                A[kk + c * lda] = b;
                A[kp + c * lda] = a;
            }
        }
    }
    if (t != 0) return;
    if (kstep == 1) {
        ipiv[k] = (int32_t)(kp + 1);
    } else {
        ipiv[k] = -(int32_t)(kp + 1);
        ipiv[k + 1] = -(int32_t)(kp + 1);
    }
    st->sk = k;
    st->sj = j;
    st->skstep = kstep;
    st->szero = S.zero;
    st->k = k + kstep;
    st->j = j + kstep;
    st->done = j + kstep >= BKB_NB - 1;
    st->need2 = 0;
}

// The last step's column(s) of the panel.
__global__ __launch_bounds__(BKB_ROWS) void bkb_store_kernel(double* __restrict__ A, int64_t n,
                                                             int64_t lda,
                                                             const double* __restrict__ W,
                                                             int64_t ldw, const BkbState* st) {
    const BkbState S = *st;
    if (!S.active || S.sk < 0) return;
    bkb_store_row(A, n, lda, W, ldw, S, S.sk + (int64_t)blockIdx.x * BKB_ROWS + threadIdx.x);
}

// Trailing update A22 -= L21 W21^T (lower 128 x 128 tiles), L21 = A(s:n, k0:s), W21 =
// W(s:n, 0:kb), s = k0 + kb.
__global__ __launch_bounds__(256, 2) void bkb_update_kernel(double* __restrict__ A, int64_t n,
                                                            int64_t lda,
                                                            const double* __restrict__ W,
                                                            int64_t ldw, const BkbState* st) {
    __shared__ __attribute__((aligned(16))) double lds[WG_LDS_ELEMS];
    const int64_t k0 = st->k0, s0 = st->k;
    const int kb = st->j;
    if (!st->active) return;
    const int64_t m = n - s0;
    const int64_t bid = blockIdx.x;
    int64_t bi = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while (bi * (bi + 1) / 2 > bid) --bi;
    while ((bi + 1) * (bi + 2) / 2 <= bid) ++bi;
    const int64_t bj = bid - bi * (bi + 1) / 2;
    if (m <= 0 || bi * WG_T >= m) return;  // uniform over the workgroup
    Operand<double, LAYOUT_COL> opA{A + s0 + k0 * lda, lda, m, kb};
    Operand<double, LAYOUT_COL> opB{W + s0, ldw, m, kb};
    double* C = A + s0 + s0 * lda;
    const int w = threadIdx.x >> 6;
    const int64_t col0 = bj * WG_T + 64 * (w & 1);
    const int64_t row0 = bi * WG_T + 64 * (w >> 1);
    const bool live = row0 < m && col0 < m && col0 <= row0 + 63;
    const bool interior = live && row0 + 64 <= m && col0 + 64 <= m && col0 + 63 <= row0;
    WaveTile<double> t;
    wave_tile_load_c<double, true>(t, C, lda, row0, col0, m, m, live, interior);
    wg_tile_mma_lds<double, LAYOUT_COL, LAYOUT_COL, true>(t, lds, opB, opA, bj * WG_T, bi * WG_T,
                                                          kb);
    if (live) wave_tile_store_c<double, true>(t, C, lda, row0, col0, m, m, interior);
}

// Undo the panel's later interchanges in its earlier columns, one lane per column: the step
// ending at panel column e (pivot row jp) had swapped rows k0 + e and jp in the columns before
// its pivot block; walking the steps backwards restores the unblocked storage.
__global__ void bkb_undo_kernel(double* __restrict__ A, int64_t lda, const int32_t* __restrict__ ipiv,
                                const BkbState* st) {
    if (!st->active) return;
    const int64_t k0 = st->k0;
    const int kb = st->j;
    const int q = threadIdx.x;
    if (q >= kb) return;
    const int64_t c = k0 + q;
    int e = kb - 1;
    while (e > q) {
        const int32_t p = ipiv[k0 + e];
        int first;  // first panel column of the step's pivot block
        int64_t jp;
        if (p < 0) {
            jp = -(int64_t)p - 1;
            first = e - 1;
        } else {
            jp = (int64_t)p - 1;
            first = e;
        }
        const int64_t jj = k0 + e;
        if (jp != jj && q < first) {
            const double a = A[jj + c * lda], b = A[jp + c * lda];
            A[jj + c * lda] = b;
            A[jp + c * lda] = a;
        }
        e = first - 1;
    }
}

// The last columns (fewer than BKB_NB + 1), unblocked, in one workgroup.
__global__ __launch_bounds__(BK_THREADS) void bkb_tail_kernel(double* __restrict__ A, int64_t n,
                                                              int64_t lda,
                                                              int32_t* __restrict__ ipiv,
                                                              int32_t* __restrict__ info,
                                                              const BkbState* st,
                                                              double* __restrict__ wsp) {
    __shared__ double s_v[BK_WAVES];
    __shared__ int64_t s_i[BK_WAVES];
    const int64_t k = st->k;
    if (k >= n) return;
    const int32_t first_zero = bk_unblocked(A, n, lda, k, ipiv, wsp, s_v, s_i);
    if (threadIdx.x == 0 && first_zero != 0 && *info == 0) *info = first_zero;
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

// Order from which the blocked factorisation is used (below it: one workgroup). Measured on
// MI355X: blocked 3.1 ms at order 256, 6.4 ms at 512 against 3.4 and 7.4 ms for a step of
// launches per column with an O(n^2) update each.
static constexpr int64_t BKB_MIN = 256;

namespace {
// Workspace: [256, 512) blocked state, argmax partials, the 2 x 2 multipliers (2n doubles),
// and for the blocked path the panel W (ldw x BKB_NB doubles).
struct BkLayout {
    size_t part, wsp, w, total;
    int64_t ldw;
};
BkLayout bk_layout(int64_t n) {
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    BkLayout L;
    const int64_t npart = (n + qmlh::BKB_RB - 1) / qmlh::BKB_RB;
    L.part = 512;
    L.wsp = al(L.part + (size_t)npart * sizeof(qmlh::BkbPart));
    L.w = al(L.wsp + (size_t)(2 * n) * sizeof(double));
    L.ldw = (n + 1) / 2 * 2;
    L.total = n >= BKB_MIN ? L.w + (size_t)L.ldw * qmlh::BKB_NB * sizeof(double) : L.w;
    return L;
}
}  // namespace

extern "C" size_t qmlh_bkf_workspace(int64_t n) { return n > 0 ? bk_layout(n).total : 0; }

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
    const BkLayout lay = bk_layout(n);
    char* base = static_cast<char*>(ws);
    double* wsp = reinterpret_cast<double*>(base + lay.wsp);
    if (n < BKB_MIN) {
        hipLaunchKernelGGL(bkf_factor_kernel, dim3(1), dim3(BK_THREADS), 0, s, A, n, lda, ipiv,
                           info, wsp);
        QMLH_CHECK_LAUNCH("bkf_factor launch");
        return QMLH_OK;
    }
    // Blocked: a panel holds at least BKB_NB - 1 columns, so panel p starts at or after column
    // (BKB_NB - 1) p; panels run while more than BKB_NB columns remain.
    BkbState* st = reinterpret_cast<BkbState*>(base + 256);
    BkbPart* part = reinterpret_cast<BkbPart*>(base + lay.part);
    double* W = reinterpret_cast<double*>(base + lay.w);
    const int64_t ldw = lay.ldw;
    hipLaunchKernelGGL(bkb_init_kernel, dim3(1), dim3(1), 0, s, st, info);
    QMLH_CHECK_LAUNCH("bkf_factor init launch");
    const int64_t panels = (n - BKB_NB - 1) / (BKB_NB - 1) + 1;
    for (int64_t p = 0; p < panels; ++p) {
        const int64_t lo = p * (BKB_NB - 1);  // lower bound of the panel's first column
        const unsigned g = (unsigned)ceil_div(n - lo, BKB_RB);
        const unsigned gs = (unsigned)ceil_div(n - lo, BKB_ROWS);
        hipLaunchKernelGGL(bkb_begin_kernel, dim3(1), dim3(1), 0, s, st, n);
        for (int step = 0; step < BKB_NB - 1; ++step) {
            hipLaunchKernelGGL(bkb_column_kernel, dim3(g), dim3(256), 0, s, A, n, lda, W, ldw, st,
                               part, info);
            hipLaunchKernelGGL(bkb_pivot_kernel, dim3(g), dim3(256), 0, s, A, n, lda, W, ldw, st,
                               part, ipiv);
        }
        hipLaunchKernelGGL(bkb_store_kernel, dim3(gs), dim3(BKB_ROWS), 0, s, A, n, lda, W, ldw, st);
        const int64_t m = n - lo - (BKB_NB - 1);  // trailing order bound
        if (m > 0) {
            const int64_t t = ceil_div(m, (int64_t)WG_T);
            hipLaunchKernelGGL(bkb_update_kernel, dim3((unsigned)(t * (t + 1) / 2)), dim3(256), 0,
                               s, A, n, lda, W, ldw, st);
        }
        hipLaunchKernelGGL(bkb_undo_kernel, dim3(1), dim3(BKB_NB), 0, s, A, lda, ipiv, st);
        QMLH_CHECK_LAUNCH("bkf_factor panel launches");
    }
    hipLaunchKernelGGL(bkb_tail_kernel, dim3(1), dim3(BK_THREADS), 0, s, A, n, lda, ipiv, info,
                       st, wsp);
    QMLH_CHECK_LAUNCH("bkf_factor tail launch");
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
