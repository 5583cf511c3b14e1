// Single-launch triangular solves with a lower-triangular L (column-major, storage T):
//   forward  z = L^-1 r        backward  x = L^-T z
// One workgroup per 64-row block. A workgroup takes its block from an atomic ticket, so
// blocks are claimed in launch order and a block only ever waits for blocks already claimed
// by running workgroups (no deadlock, whatever the dispatch order). It accumulates the
// contributions of the finished blocks as they are published -- loading the L tile before
// waiting, so the loads overlap the wait -- then solves its 64 x 64 diagonal block in FP64
// and publishes the result.
//
// Publication needs no flag: before the solve the output vector is filled with a NaN bit
// pattern that arithmetic never produces (TS_PENDING), each solved value is written with one
// 64-bit atomic store, and a consumer polls the values it needs with 64-bit atomic loads until
// none is pending. One memory round trip per dependency and no release fence on the producer.
// Every wait is bounded: past the spin limit the workgroup records an error and stops
// waiting, so the kernel always drains; the host turns the error into a failed status.
//
// The accumulation order is fixed (finished blocks in solve order, a fixed intra-thread
// order, a fixed final reduction tree), so results are deterministic.
#include "chol_panel.h"

namespace qmlh {

static constexpr int TS_B = 64;
static constexpr unsigned TS_SPIN_LIMIT = 1u << 22;

// Quiet NaN with a payload no arithmetic produces (computed NaNs are the canonical
// 0x7ff8000000000000; the timeout path below writes that one).
static constexpr unsigned long long TS_PENDING = 0x7ffcdeadbeef0001ull;

// sync: [0] ticket, [1] error (the remaining words are unused); out[0 .. n) := pending.
__global__ void trsv_prepare_kernel(int32_t* __restrict__ sync, int64_t count,
                                    double* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) sync[i] = 0;
    if (i < n) out[i] = __builtin_bit_cast(double, TS_PENDING);
}

template <typename T, bool FWD>
__device__ __forceinline__ void load_tile16(T (&v)[16], const T* __restrict__ L, int64_t ldl,
                                            int64_t n, int64_t r0, int w, int64_t j, int lane,
                                            int grp) {
    const int64_t c0 = j * TS_B;
    const int wj = (int)(n - c0 < TS_B ? n - c0 : TS_B);
    // Full tiles (all but the last row or column block) load without any condition, so the 16
    // loads stay in flight together; a condition on the lane would put each load in its own
    // branch and wait for it there.
    if (w == TS_B && wj == TS_B) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int cc = 16 * grp + c;
            v[c] = FWD ? L[(r0 + lane) + (c0 + cc) * ldl] : L[(c0 + lane) + (r0 + cc) * ldl];
        }
        return;
    }
    // Partial tiles (the last row or column block): unconditional loads from indices clamped
    // into the tile, with no mask. Entries past the matrix then repeat its last row or column
    // (finite values of a computed factor) and only meet zeros: FWD rows past w and BWD
    // columns past w feed residuals that are never published and that the diagonal step
    // multiplies by zero; FWD columns past wj and BWD rows past wj meet published values that
    // the loader set to zero. A mask (or select) would be compiled into a conditional load
    // that waits in its own branch.
    if (w <= 0 || wj <= 0) {
#pragma unroll
        for (int c = 0; c < 16; ++c) v[c] = T(0);
        return;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int cc = 16 * grp + c;
        v[c] = FWD ? L[(r0 + cp_clamp(lane, w)) + (c0 + cp_clamp(cc, wj)) * ldl]
                   : L[(c0 + cp_clamp(lane, wj)) + (r0 + cp_clamp(cc, w)) * ldl];
    }
}

// One dependency block in flight: the lane's 16 tile values and one published value (FWD:
// z of column c0 + lane of block j, shared by the wave through lane reads; BWD: x of row
// c0 + lane of block j, the lane's own).
template <typename T>
struct TsStage {
    T l[16];
    double v;
};

// Issue the loads of dependency q (tile + published value); nothing if q >= nj.
template <typename T, bool FWD>
__device__ __forceinline__ void ts_issue(TsStage<T>& st, const T* __restrict__ L, int64_t ldl,
                                         int64_t n, int64_t r0, int w, int64_t nblocks,
                                         int64_t q, int64_t nj, const double* out, int lane,
                                         int grp) {
    if (q >= nj) return;
    const int64_t j = FWD ? q : nblocks - 1 - q;
    load_tile16<T, FWD>(st.l, L, ldl, n, r0, w, j, lane, grp);
    const int64_t c0 = j * TS_B;
    const int wj = (int)(n - c0 < TS_B ? n - c0 : TS_B);
    st.v = lane < wj ? __hip_atomic_load(&out[c0 + lane], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT)
                     : 0.0;
}

// Wait (bounded) until no lane of the wave holds a pending value. Two polls are kept in
// flight, checked oldest first and re-issued at once, so a value that lands is seen about half
// a memory round trip after it lands instead of a full one. (Published values never change
// again, so re-reading a settled lane is harmless; lanes that are not live keep their value,
// which the loader set to zero.)
__device__ __forceinline__ bool ts_pending(double v, bool live) {
    return live && __builtin_bit_cast(unsigned long long, v) == TS_PENDING;
}

__device__ __forceinline__ void ts_settle(double& v, const double* p, bool live, int32_t* err,
                                          bool& ok) {
    if (!__any(ts_pending(v, live))) return;
    double a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_sleep(2);
    double b = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (unsigned spins = 0; spins < TS_SPIN_LIMIT; spins += 2) {
        if (!__any(ts_pending(a, live))) {
            if (live) v = a;
            return;
        }
        a = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!__any(ts_pending(b, live))) {
            if (live) v = b;
            return;
        }
        b = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ok = false;
    v = 0.0;
}

// Consume dependency q: settle its published values, then accumulate its contributions.
template <typename T, bool FWD>
__device__ __forceinline__ void ts_consume(TsStage<T>& st, int64_t n, int64_t nblocks,
                                           int64_t q, int64_t nj, const double* out, int lane,
                                           int grp_s, int32_t* err, bool& ok, double& acc1,
                                           double (&acc16)[16]) {
    if (q >= nj) return;
    const int64_t j = FWD ? q : nblocks - 1 - q;
    const int64_t c0 = j * TS_B;
    const int wj = (int)(n - c0 < TS_B ? n - c0 : TS_B);
    if (ok) ts_settle(st.v, &out[c0 + (lane < wj ? lane : 0)], lane < wj, err, ok);
    if (!ok) st.v = 0.0;
    if (FWD) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            acc1 = fma((double)st.l[c], readlane_f64(st.v, 16 * grp_s + c), acc1);
    } else {
#pragma unroll
        for (int c = 0; c < 16; ++c) acc16[c] = fma((double)st.l[c], st.v, acc16[c]);
    }
}

// Solve block b of one system: wait for the values of the blocks it depends on (the earlier
// blocks in solve order), accumulate their contributions, solve the diagonal block by FP64
// substitution, publish. `err` is the shared error word. All 256 lanes call this.
//
// Dependencies are pipelined TS_DEPTH deep: the tile and published values of the next blocks
// are in flight while the current one is accumulated, so a block whose dependencies are long
// finished pays about one memory round trip per TS_DEPTH dependencies, not one per dependency.
// The accumulation order is unchanged (dependencies in solve order).
static constexpr int TS_DEPTH = 4;

template <typename T, bool FWD>
__device__ __forceinline__ void trsv_block(const T* __restrict__ L, int64_t ldl, int64_t n,
                                           const double* __restrict__ rhs, double* out,
                                           int32_t* __restrict__ err, int64_t b) {
    __shared__ double sL[TS_B][TS_B + 1];
    __shared__ double sInv[TS_B];
    __shared__ double red[TS_B][TS_B + 1];
    __shared__ double redp[4][TS_B];
    __shared__ double sv[TS_B];
    __shared__ int s_fail;
    const int t = threadIdx.x;
    const int64_t nblocks = (n + TS_B - 1) / TS_B;
    if (t == 0) s_fail = 0;
    const int64_t r0 = b * TS_B;
    const int w = (int)(n - r0 < TS_B ? n - r0 : TS_B);
    const int lane = t & 63, grp = t >> 6;  // 4 groups of 16 columns
    const int grp_s = __builtin_amdgcn_readfirstlane(grp);  // wave-uniform

    // The diagonal block does not depend on other blocks: stage it first.
    stage_lower64<false, TS_B + 1>(sL, L + r0 + r0 * ldl, ldl, w);
    __syncthreads();  // s_fail initialised before any lane may set it

    // Forward: thread (row lane, columns 16 grp ..) keeps one partial sum.
    // Backward: thread (row lane of block j, columns 16 grp .. of block b) keeps 16.
    const double rhs_v = rhs[r0 + cp_clamp(lane, w)];  // used after the dependencies
    double acc1 = 0.0;
    double acc16[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc16[c] = 0.0;
    bool ok = true;
    const int64_t nj = FWD ? b : nblocks - 1 - b;
    static_assert(TS_DEPTH == 4, "the ring below is written out for depth 4");
    TsStage<T> s0, s1, s2, s3;
    ts_issue<T, FWD>(s0, L, ldl, n, r0, w, nblocks, 0, nj, out, lane, grp);
    ts_issue<T, FWD>(s1, L, ldl, n, r0, w, nblocks, 1, nj, out, lane, grp);
    ts_issue<T, FWD>(s2, L, ldl, n, r0, w, nblocks, 2, nj, out, lane, grp);
    ts_issue<T, FWD>(s3, L, ldl, n, r0, w, nblocks, 3, nj, out, lane, grp);
    for (int64_t q = 0; q < nj; q += TS_DEPTH) {
        ts_consume<T, FWD>(s0, n, nblocks, q, nj, out, lane, grp_s, err, ok, acc1, acc16);
        ts_issue<T, FWD>(s0, L, ldl, n, r0, w, nblocks, q + 4, nj, out, lane, grp);
        ts_consume<T, FWD>(s1, n, nblocks, q + 1, nj, out, lane, grp_s, err, ok, acc1, acc16);
        ts_issue<T, FWD>(s1, L, ldl, n, r0, w, nblocks, q + 5, nj, out, lane, grp);
        ts_consume<T, FWD>(s2, n, nblocks, q + 2, nj, out, lane, grp_s, err, ok, acc1, acc16);
        ts_issue<T, FWD>(s2, L, ldl, n, r0, w, nblocks, q + 6, nj, out, lane, grp);
        ts_consume<T, FWD>(s3, n, nblocks, q + 3, nj, out, lane, grp_s, err, ok, acc1, acc16);
        ts_issue<T, FWD>(s3, L, ldl, n, r0, w, nblocks, q + 7, nj, out, lane, grp);
    }
    if (!ok) s_fail = 1;  // benign race: every writer stores 1

    if (FWD) {
        red[lane][grp] = acc1;
    } else {
#pragma unroll
        for (int c = 0; c < 16; ++c) red[lane][16 * grp + c] = acc16[c];
    }
    __syncthreads();
    if (!FWD) {
        // column sums over the 64 rows: four 16-row partials per column (all 256 lanes, the
        // 16 LDS reads of a lane in flight together), then added in a fixed order
        double ps = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) ps += red[16 * grp + i][lane];
        redp[grp][lane] = ps;
        __syncthreads();
    }
    if (t < TS_B) {
        double s = 0.0;
        if (FWD) {
            s = ((red[t][0] + red[t][1]) + red[t][2]) + red[t][3];
        } else {
            s = ((redp[0][t] + redp[1][t]) + redp[2][t]) + redp[3][t];
        }
        sv[t] = (t < w ? rhs_v : 0.0) - s;
        if (t < w) sInv[t] = 1.0 / sL[t][t];
    }
    __syncthreads();
    if (t < 64) {
        const int i = t;
        double xi = sv[i];
        if (FWD) {
            for (int c = 0; c < w; ++c) {
                const double xc = __shfl(xi * sInv[c], c, 64);
                if (i == c) xi = xc;
                if (i > c && i < w) xi = xi - sL[i][c] * xc;
            }
        } else {
            for (int c = w - 1; c >= 0; --c) {
                const double xc = __shfl(xi * sInv[c], c, 64);
                if (i == c) xi = xc;
                if (i < c) xi = xi - sL[c][i] * xc;
            }
        }
        if (i < w)
            __hip_atomic_store(&out[r0 + i], s_fail ? __builtin_nan("") : xi, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// dinv block b (64 x 64, column-major, zero above the diagonal and outside the block) :=
// L_bb^-1 in FP64. Lane c computes column c by forward substitution in registers.
template <typename T>
__global__ __launch_bounds__(64) void trsv_diag_inverse_kernel(const T* __restrict__ L,
                                                               int64_t ldl, int64_t n,
                                                               double* __restrict__ dinv) {
    __shared__ double sL[TS_B][TS_B + 1];
    const int64_t b = blockIdx.x;
    const int64_t r0 = b * TS_B;
    const int w = (int)(n - r0 < TS_B ? n - r0 : TS_B);
    const int c = threadIdx.x;
    for (int idx = c; idx < TS_B * TS_B; idx += 64) {
        const int i = idx % TS_B, j = idx / TS_B;
        sL[i][j] = (i < w && j < w && i >= j) ? (double)L[(r0 + i) + (r0 + j) * ldl] : 0.0;
    }
    __syncthreads();
    double x[TS_B];
#pragma unroll
    for (int i = 0; i < TS_B; ++i) {
        double v = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) v = fma(-sL[i][k], x[k], v);
        x[i] = (i < w && c < w && i >= c) ? v / sL[i][i] : 0.0;
    }
    double* D = dinv + b * (TS_B * TS_B);
#pragma unroll
    for (int i = 0; i < TS_B; ++i) D[i + c * TS_B] = x[i];
}

template <typename T, bool FWD>
__global__ __launch_bounds__(256) void trsv_syncfree_kernel(const T* __restrict__ L, int64_t ldl,
                                                            int64_t n, const double* __restrict__ rhs,
                                                            double* out,
                                                            int32_t* __restrict__ sync) {
    __shared__ int s_block;
    const int64_t nblocks = (n + TS_B - 1) / TS_B;
    if (threadIdx.x == 0) {
        const int ticket = atomicAdd(&sync[0], 1);
        s_block = FWD ? ticket : (int)(nblocks - 1 - ticket);
    }
    __syncthreads();
    trsv_block<T, FWD>(L, ldl, n, rhs, out, &sync[1], s_block);
}

size_t trsv_sync_ints(int64_t n) { return (size_t)((n + TS_B - 1) / TS_B + 2); }

size_t trsv_dinv_doubles(int64_t n) { return (size_t)((n + TS_B - 1) / TS_B) * TS_B * TS_B; }

template <typename T>
int trsv_diag_inverse(const T* L, int64_t n, int64_t ldl, double* dinv, hipStream_t s) {
    if (n <= 0) return QMLH_OK;
    hipLaunchKernelGGL(trsv_diag_inverse_kernel<T>, dim3((unsigned)((n + TS_B - 1) / TS_B)),
                       dim3(64), 0, s, L, ldl, n, dinv);
    QMLH_CHECK_LAUNCH("trsv diagonal inverse launch");
    return QMLH_OK;
}

template <typename T>
int trsv_syncfree(const T* L, int64_t n, int64_t ldl, const double* rhs, double* out,
                  int32_t* sync, bool forward, hipStream_t s) {
    if (n <= 0) return QMLH_OK;
    const int64_t nblocks = (n + TS_B - 1) / TS_B;
    hipLaunchKernelGGL(trsv_prepare_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, sync,
                       (int64_t)2, out, n);
    if (forward)
        hipLaunchKernelGGL((trsv_syncfree_kernel<T, true>), dim3((unsigned)nblocks), dim3(256), 0,
                           s, L, ldl, n, rhs, out, sync);
    else
        hipLaunchKernelGGL((trsv_syncfree_kernel<T, false>), dim3((unsigned)nblocks), dim3(256),
                           0, s, L, ldl, n, rhs, out, sync);
    QMLH_CHECK_LAUNCH("trsv launch");
    return QMLH_OK;
}

template int trsv_syncfree<double>(const double*, int64_t, int64_t, const double*, double*,
                                   int32_t*, bool, hipStream_t);
template int trsv_syncfree<float>(const float*, int64_t, int64_t, const double*, double*,
                                  int32_t*, bool, hipStream_t);
template int trsv_diag_inverse<double>(const double*, int64_t, int64_t, double*, hipStream_t);
template int trsv_diag_inverse<float>(const float*, int64_t, int64_t, double*, hipStream_t);

// ------------------------------------------------------------------------------------------
// 128-row blocks for solves that have diagonal-block inverses (the preconditioner's). One
// workgroup of 512 lanes per 128-row block: each 256-lane half accumulates the dependencies of
// its 64 rows (FWD) or 64 columns (BWD) exactly as trsv_block does, then the 128 residual
// values are multiplied by the FP64 inverse of the 128 x 128 diagonal block in one step. The
// dependency chain of the solve has half as many links, and each link costs one diagonal
// product instead of two solves and a hand-over. Published values are read by later blocks in
// 64-column units, as before.
// ------------------------------------------------------------------------------------------
static constexpr int TS_B2 = 2 * TS_B;

// dinv128 block B := L_BB^-1 (128 x 128, column-major) from the 64 x 64 inverses D0 = dinv64
// block 2B and D1 = block 2B + 1 and the off-diagonal block L10:
//     [D0 0; X D1],  X = -D1 (L10 D0),
// all in FP64. Entries outside the matrix are zero. 256 lanes per block.
template <typename T>
__global__ __launch_bounds__(256) void trsv_inverse128_kernel(const T* __restrict__ L,
                                                              int64_t ldl, int64_t n,
                                                              const double* __restrict__ dinv64,
                                                              double* __restrict__ dinv128) {
    __shared__ double sa[TS_B][TS_B + 1], sb[TS_B][TS_B + 1];
    const int64_t B = blockIdx.x;
    const int64_t r0 = B * TS_B2;
    const int w1 = (int)(n - r0 - TS_B < TS_B ? n - r0 - TS_B : TS_B);  // rows of the 2nd half
    const int t = threadIdx.x;
    const int i = t & 63, cg = t >> 6;  // output row, group of 16 output columns
    const double* D0 = dinv64 + (2 * B) * (TS_B * TS_B);
    double* Dout = dinv128 + B * (TS_B2 * TS_B2);
    double xr[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) xr[c] = 0.0;
    if (w1 > 0) {  // uniform
        const double* D1 = dinv64 + (2 * B + 1) * (TS_B * TS_B);
        // sa = L10 (rows r0 + 64 .., columns r0 ..), sb = D0
        for (int idx = t; idx < TS_B * TS_B; idx += 256) {
            const int ii = idx & 63, kk = idx >> 6;
            sa[ii][kk] = ii < w1 ? (double)L[(r0 + TS_B + ii) + (r0 + kk) * ldl] : 0.0;
            sb[ii][kk] = D0[ii + kk * TS_B];
        }
        __syncthreads();
        double tr[16];  // T = L10 D0, row i, columns 16 cg ..
#pragma unroll
        for (int c = 0; c < 16; ++c) tr[c] = 0.0;
        for (int k = 0; k < TS_B; ++k) {
            const double a = sa[i][k];
#pragma unroll
            for (int c = 0; c < 16; ++c) tr[c] = fma(a, sb[k][16 * cg + c], tr[c]);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 16; ++c) sa[i][16 * cg + c] = tr[c];  // sa = T
        for (int idx = t; idx < TS_B * TS_B; idx += 256) {
            const int ii = idx & 63, kk = idx >> 6;
            sb[ii][kk] = D1[ii + kk * TS_B];  // sb = D1
        }
        __syncthreads();
        for (int k = 0; k < TS_B; ++k) {  // X = -D1 T
            const double a = sb[i][k];
#pragma unroll
            for (int c = 0; c < 16; ++c) xr[c] = fma(-a, sa[k][16 * cg + c], xr[c]);
        }
        // D1 into the lower right quarter
        for (int idx = t; idx < TS_B * TS_B; idx += 256) {
            const int ii = idx & 63, kk = idx >> 6;
            Dout[(TS_B + ii) + (TS_B + kk) * TS_B2] = D1[ii + kk * TS_B];
        }
    } else {
        for (int idx = t; idx < TS_B * TS_B; idx += 256) {
            const int ii = idx & 63, kk = idx >> 6;
            Dout[(TS_B + ii) + (TS_B + kk) * TS_B2] = 0.0;
        }
    }
    for (int idx = t; idx < TS_B * TS_B; idx += 256) {
        const int ii = idx & 63, kk = idx >> 6;
        Dout[ii + kk * TS_B2] = D0[ii + kk * TS_B];           // upper left: D0
        Dout[ii + (TS_B + kk) * TS_B2] = 0.0;                 // upper right: 0
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) Dout[(TS_B + i) + (16 * cg + c) * TS_B2] = xr[c];  // lower left
}

template <typename T, bool FWD>
__global__ __launch_bounds__(512) void trsv128_kernel(const T* __restrict__ L, int64_t ldl,
                                                      int64_t n, const double* __restrict__ rhs,
                                                      double* out, int32_t* __restrict__ sync,
                                                      const double* __restrict__ dinv128) {
    __shared__ double red[2][TS_B][TS_B + 1];
    __shared__ double redp[2][4][TS_B];
    __shared__ double sres[TS_B2];
    __shared__ double red2[TS_B2][4];
    __shared__ int s_block, s_fail;
    const int t = threadIdx.x;
    const int h = t >> 8, tt = t & 255;
    const int lane = tt & 63, grp = tt >> 6;
    const int grp_s = __builtin_amdgcn_readfirstlane(grp);
    const int64_t nblocks = (n + TS_B - 1) / TS_B;     // 64-row blocks
    const int64_t nb2 = (n + TS_B2 - 1) / TS_B2;        // 128-row blocks
    int32_t* err = &sync[1];
    if (t == 0) {
        const int ticket = atomicAdd(&sync[0], 1);
        s_block = FWD ? ticket : (int)(nb2 - 1 - ticket);
        s_fail = 0;
    }
    __syncthreads();
    const int64_t B = s_block;
    const int64_t r0 = B * TS_B2;
    const int w = (int)(n - r0 < TS_B2 ? n - r0 : TS_B2);
    const int64_t rh = r0 + TS_B * h;                     // this half's 64 rows / columns
    const int wh = w - TS_B * h < 0 ? 0 : (w - TS_B * h > TS_B ? TS_B : w - TS_B * h);

    // The diagonal product's operands do not depend on other blocks: load them first.
    // Lane (i, part) forms x_i's partial over columns 32 part .. 32 part + 31:
    // FWD x = D s reads D(i, c); BWD x = D^T s reads D(c, i), a contiguous run.
    const int di = t & (TS_B2 - 1), dpart = t >> 7;
    const double* D = dinv128 + B * (TS_B2 * TS_B2);
    double dv[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const int cc = 32 * dpart + c;
        dv[c] = FWD ? D[di + cc * TS_B2] : D[cc + di * TS_B2];
    }

    // this lane's right-hand side value (used after the dependencies): loaded now, off the
    // critical path
    const double rhs_v = wh > 0 ? rhs[rh + cp_clamp(lane, wh)] : 0.0;
    double acc1 = 0.0;
    double acc16[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc16[c] = 0.0;
    bool ok = true;
    // dependencies: the 64-blocks before this block (FWD) or after it (BWD), in solve order
    const int64_t nj = FWD ? 2 * B : nblocks - 2 * B - 2 > 0 ? nblocks - 2 * B - 2 : 0;
    if constexpr (sizeof(T) == 4 && !FWD) {
        // The backward solve keeps 16 partial sums per lane: three stages fit its registers
        // (four spill).
        TsStage<T> s0, s1, s2;
        ts_issue<T, FWD>(s0, L, ldl, n, rh, wh, nblocks, 0, nj, out, lane, grp);
        ts_issue<T, FWD>(s1, L, ldl, n, rh, wh, nblocks, 1, nj, out, lane, grp);
        ts_issue<T, FWD>(s2, L, ldl, n, rh, wh, nblocks, 2, nj, out, lane, grp);
        for (int64_t q = 0; q < nj; q += 3) {
            ts_consume<T, FWD>(s0, n, nblocks, q, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s0, L, ldl, n, rh, wh, nblocks, q + 3, nj, out, lane, grp);
            ts_consume<T, FWD>(s1, n, nblocks, q + 1, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s1, L, ldl, n, rh, wh, nblocks, q + 4, nj, out, lane, grp);
            ts_consume<T, FWD>(s2, n, nblocks, q + 2, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s2, L, ldl, n, rh, wh, nblocks, q + 5, nj, out, lane, grp);
        }
    } else if constexpr (sizeof(T) == 8) {
        // FP64 tiles are twice the bytes: two stages keep as many bytes in flight as four FP32
        // stages, in registers the workgroup's 512 lanes have.
        TsStage<T> s0, s1;
        ts_issue<T, FWD>(s0, L, ldl, n, rh, wh, nblocks, 0, nj, out, lane, grp);
        ts_issue<T, FWD>(s1, L, ldl, n, rh, wh, nblocks, 1, nj, out, lane, grp);
        for (int64_t q = 0; q < nj; q += 2) {
            ts_consume<T, FWD>(s0, n, nblocks, q, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s0, L, ldl, n, rh, wh, nblocks, q + 2, nj, out, lane, grp);
            ts_consume<T, FWD>(s1, n, nblocks, q + 1, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s1, L, ldl, n, rh, wh, nblocks, q + 3, nj, out, lane, grp);
        }
    } else {
        TsStage<T> s0, s1, s2, s3;
        ts_issue<T, FWD>(s0, L, ldl, n, rh, wh, nblocks, 0, nj, out, lane, grp);
        ts_issue<T, FWD>(s1, L, ldl, n, rh, wh, nblocks, 1, nj, out, lane, grp);
        ts_issue<T, FWD>(s2, L, ldl, n, rh, wh, nblocks, 2, nj, out, lane, grp);
        ts_issue<T, FWD>(s3, L, ldl, n, rh, wh, nblocks, 3, nj, out, lane, grp);
        for (int64_t q = 0; q < nj; q += TS_DEPTH) {
            ts_consume<T, FWD>(s0, n, nblocks, q, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s0, L, ldl, n, rh, wh, nblocks, q + 4, nj, out, lane, grp);
            ts_consume<T, FWD>(s1, n, nblocks, q + 1, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s1, L, ldl, n, rh, wh, nblocks, q + 5, nj, out, lane, grp);
            ts_consume<T, FWD>(s2, n, nblocks, q + 2, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s2, L, ldl, n, rh, wh, nblocks, q + 6, nj, out, lane, grp);
            ts_consume<T, FWD>(s3, n, nblocks, q + 3, nj, out, lane, grp_s, err, ok, acc1, acc16);
            ts_issue<T, FWD>(s3, L, ldl, n, rh, wh, nblocks, q + 7, nj, out, lane, grp);
        }
    }
    if (!ok) s_fail = 1;  // benign race: every writer stores 1

    if (FWD) {
        red[h][lane][grp] = acc1;
    } else {
#pragma unroll
        for (int c = 0; c < 16; ++c) red[h][lane][16 * grp + c] = acc16[c];
    }
    __syncthreads();
    if (!FWD) {  // column sums: four 16-row partials per column, added in a fixed order
        double ps = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) ps += red[h][16 * grp + i][lane];
        redp[h][grp][lane] = ps;
        __syncthreads();
    }
    if (tt < TS_B) {  // residual of this half's 64 rows
        double sum = 0.0;
        if (FWD) {
            sum = ((red[h][tt][0] + red[h][tt][1]) + red[h][tt][2]) + red[h][tt][3];
        } else {
            sum = ((redp[h][0][tt] + redp[h][1][tt]) + redp[h][2][tt]) + redp[h][3][tt];
        }
        sres[TS_B * h + tt] = (tt < wh ? rhs_v : 0.0) - sum;
    }
    __syncthreads();
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < 32; ++c) p = fma(dv[c], sres[32 * dpart + c], p);
    red2[di][dpart] = p;
    __syncthreads();
    if (t < w) {
        const double xi = ((red2[t][0] + red2[t][1]) + red2[t][2]) + red2[t][3];
        __hip_atomic_store(&out[r0 + t], s_fail ? __builtin_nan("") : xi, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

size_t trsv_dinv128_doubles(int64_t n) {
    return (size_t)((n + TS_B2 - 1) / TS_B2) * TS_B2 * TS_B2;
}

template <typename T>
int trsv_inverse128(const T* L, int64_t n, int64_t ldl, double* dinv64, double* dinv128,
                    hipStream_t s) {
    if (n <= 0) return QMLH_OK;
    QMLH_TRY(trsv_diag_inverse<T>(L, n, ldl, dinv64, s));
    hipLaunchKernelGGL(trsv_inverse128_kernel<T>, dim3((unsigned)((n + TS_B2 - 1) / TS_B2)),
                       dim3(256), 0, s, L, ldl, n, dinv64, dinv128);
    QMLH_CHECK_LAUNCH("trsv 128 inverse launch");
    return QMLH_OK;
}

template <typename T>
int trsv_syncfree128(const T* L, int64_t n, int64_t ldl, const double* rhs, double* out,
                     int32_t* sync, bool forward, hipStream_t s, const double* dinv128) {
    if (n <= 0) return QMLH_OK;
    const int64_t nb2 = (n + TS_B2 - 1) / TS_B2;
    hipLaunchKernelGGL(trsv_prepare_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, sync,
                       (int64_t)2, out, n);
    if (forward)
        hipLaunchKernelGGL((trsv128_kernel<T, true>), dim3((unsigned)nb2), dim3(512), 0, s, L, ldl,
                           n, rhs, out, sync, dinv128);
    else
        hipLaunchKernelGGL((trsv128_kernel<T, false>), dim3((unsigned)nb2), dim3(512), 0, s, L,
                           ldl, n, rhs, out, sync, dinv128);
    QMLH_CHECK_LAUNCH("trsv 128 launch");
    return QMLH_OK;
}

template int trsv_inverse128<float>(const float*, int64_t, int64_t, double*, double*, hipStream_t);
template int trsv_syncfree128<float>(const float*, int64_t, int64_t, const double*, double*,
                                     int32_t*, bool, hipStream_t, const double*);
template int trsv_inverse128<double>(const double*, int64_t, int64_t, double*, double*,
                                     hipStream_t);
template int trsv_syncfree128<double>(const double*, int64_t, int64_t, const double*, double*,
                                      int32_t*, bool, hipStream_t, const double*);

}  // namespace qmlh
