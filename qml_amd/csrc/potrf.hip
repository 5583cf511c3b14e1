// Blocked right-looking Cholesky (lower, column-major) and the two triangular solves.
//
// Replaces LAPACK dpotrf("U") / dpotrs("U", n, 1) as called at qml/fcho_solve.f90:35,46 (see
// the header for why "U" on the Fortran view is the lower triangle here).
//
// Factorisation, outer block nb (512; 768 with split-FP16 trailing updates), per step k
// (from order 4096 on with look-ahead: the next block column is factored on a second stream
// while the rest of the trailing update runs, see potrf_lookahead; the split-FP16 factor
// works in groups of two block columns, see potrf_lookahead_split):
//   for each 64-wide sub-panel of the nb-wide block column (left to right):
//     1. the 64 x 64 diagonal block is factored in the registers of one wavefront (FP64
//        arithmetic), then every row of the column below is solved against it (256 rows per
//        workgroup);
//     2. an MFMA update of the block column's remaining columns (all rows below, i >= j),
//        whose workgroup for the top-left tile then factors the next diagonal block;
//   3. the trailing update A[k+kb:n, k+kb:n] -= P P^T on the lower tiles only (MFMA).
// The storage type T is double (FP64 MFMA updates) or float (FP32 MFMA updates, the
// mixed-precision factor); diagonal blocks and substitutions always run in FP64.
// A device int32 `info` follows LAPACK: the first non-positive (or NaN) pivot j sets
// info = j + 1, its Schur-complement value is left on the diagonal, and every later kernel of
// the factorisation returns without work.
#include <mutex>
#include <type_traits>
#include <vector>

#include "chol_panel.h"
#include "linalg.h"
#include "mfma_core.h"
#include "timer.h"

namespace qmlh {

static constexpr int PF_NB = 512;
// Block size when the trailing updates run as split-FP16 products: the update of C costs a
// fixed read and write of the FP32 trailing matrix per step, so a deeper update pays off.
static constexpr int PF_NB_SPLIT = 768;
static constexpr int PF_IB = CP_B;
static constexpr int UP_TILE = WG_T;

// Update shapes: FULL -- every entry of the M x N block; SYRK -- the lower triangle of an
// M x M block (lower tiles only); TRAP -- the entries i >= j of an M x N block whose (0, 0)
// sits on the diagonal (tiles entirely above the diagonal do nothing); TRAP_POTF2 -- TRAP, and
// the workgroup of tile (0, 0) then factors the w_next x w_next diagonal block at C (the next
// sub-panel's, which only that workgroup writes), so the next sub-panel needs no launch of its
// own for it.
// UPD_SYRK_UT: the lower triangle of C -= A A^T for an upper-triangular A (row i zero left of
// column i): tile row bi starts its k loop at its first row, bi * UP_TILE, skipping only
// products that are exactly zero (the inverse's -L^-T L^-1 in one launch).
enum { UPD_FULL = 0, UPD_SYRK = 1, UPD_TRAP = 2, UPD_TRAP_POTF2 = 3, UPD_SYRK_UT = 4 };

// ------------------------------------------------------------------------------------------
// C(i, j) -= sum_k A(i, k) B(j, k); A and C column-major, B column-major (LB = COL: B(j, k) at
// B[j + k ldb]) or row-major (LB = ROW: B(j, k) at B[j ldb + k]). The accumulators start from
// C (loaded while the first operand chunks are in flight) and the products are subtracted, so
// C is read once and written once.
// ------------------------------------------------------------------------------------------
template <typename T, int MODE, int LB = LAYOUT_COL>
__global__ __launch_bounds__(256, 2) void gemm_update_kernel(
    const T* __restrict__ A, int64_t lda, const T* __restrict__ B, int64_t ldb, int64_t M,
    int64_t N, int64_t Kd, T* __restrict__ C, int64_t ldc, int32_t* __restrict__ info,
    int64_t tiles_m, int w_next, int64_t col_base_next, PivotRule pr) {
    __shared__ __attribute__((aligned(16))) T lds[WG_LDS_ELEMS];
    constexpr bool MASK = MODE != UPD_FULL;
    if (info && *info) return;
    int64_t bi, bj;
    const int64_t bid = blockIdx.x;
    if (MODE == UPD_SYRK || MODE == UPD_SYRK_UT) {
        int64_t i = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > bid) --i;
        while ((i + 1) * (i + 2) / 2 <= bid) ++i;
        bi = i;
        bj = bid - i * (i + 1) / 2;
    } else {
        bi = bid % tiles_m;
        bj = bid / tiles_m;
        if (MODE >= UPD_TRAP && bj > bi) return;  // tile entirely above the diagonal
    }
    // First MFMA operand: B rows (output columns); second: A rows (output rows), so the 16
    // lanes of one accumulator element run down 16 consecutive rows of column-major C.
    const int64_t k0 = MODE == UPD_SYRK_UT ? bi * UP_TILE : 0;
    Operand<T, LAYOUT_COL> opA{A + k0 * lda, lda, M, Kd - k0};
    Operand<T, LB> opB{B + (LB == LAYOUT_COL ? k0 * ldb : k0), ldb, N, Kd - k0};
    const int w = threadIdx.x >> 6;
    const int64_t col0 = bj * UP_TILE + 64 * (w & 1);
    const int64_t row0 = bi * UP_TILE + 64 * (w >> 1);
    const bool live = row0 < M && col0 < N && !(MASK && col0 > row0 + 63);
    const bool interior = live && row0 + 64 <= M && col0 + 64 <= N && (!MASK || col0 + 63 <= row0);
    WaveTile<T> t;
    wave_tile_load_c<T, MASK>(t, C, ldc, row0, col0, M, N, live, interior);
    wg_tile_mma_lds<T, LB, LAYOUT_COL, true>(t, lds, opB, opA, bj * UP_TILE, bi * UP_TILE, Kd - k0);
    if (live) wave_tile_store_c<T, MASK>(t, C, ldc, row0, col0, M, N, interior);
    if (MODE == UPD_TRAP_POTF2 && bid == 0 && w_next > 0) {
        // every wave of this workgroup got here; the barrier makes the tile's stores visible
        // to the whole workgroup and frees the operand LDS for the factor
        __syncthreads();
        potf2_body_sp<T>(C, ldc, w_next, col_base_next, info, reinterpret_cast<double*>(lds),
                         pr.floor, pr.bshift, pr.shift_acc, pr.bpivot);
    }
}

// C -= A B^T in the given shape. timer_class: the measurement class the launch is counted in
// (by default SYRK for the trailing update shape and GEMM for the others).
template <typename T, int LB = LAYOUT_COL>
static int launch_update(int mode, const T* A, int64_t lda, const T* B, int64_t ldb, int64_t M,
                         int64_t N, int64_t Kd, T* C, int64_t ldc, int32_t* info,
                         hipStream_t s, int timer_class = -1, int w_next = 0,
                         int64_t col_base_next = 0, PivotRule pr = PivotRule()) {
    if (M <= 0 || N <= 0 || Kd <= 0) return QMLH_OK;
    const int64_t tm = ceil_div(M, UP_TILE), tn = ceil_div(N, UP_TILE);
    if (timer_class < 0) timer_class = mode == UPD_SYRK ? QMLH_TIMER_SYRK : QMLH_TIMER_GEMM;
    double flops = 2.0 * (double)M * (double)N * (double)Kd;
    if (mode == UPD_SYRK) flops = (double)M * (double)(M + 1) * (double)Kd;
    if (mode == UPD_SYRK_UT) flops = (double)M * (double)M * (double)M / 3.0;  // M = N = Kd
    if (mode >= UPD_TRAP) flops = ((double)M * (double)N - (double)N * (double)(N - 1) / 2.0) * 2.0 * (double)Kd;
    TimerScope ts(timer_class, flops, s);
    switch (mode) {
        case UPD_SYRK:
            hipLaunchKernelGGL((gemm_update_kernel<T, UPD_SYRK, LB>), dim3((unsigned)(tm * (tm + 1) / 2)),
                               dim3(256), 0, s, A, lda, B, ldb, M, N, Kd, C, ldc, info, tm, 0,
                               (int64_t)0, PivotRule());
            break;
        case UPD_SYRK_UT:
            hipLaunchKernelGGL((gemm_update_kernel<T, UPD_SYRK_UT, LB>), dim3((unsigned)(tm * (tm + 1) / 2)),
                               dim3(256), 0, s, A, lda, B, ldb, M, N, Kd, C, ldc, info, tm, 0,
                               (int64_t)0, PivotRule());
            break;
        case UPD_TRAP:
            hipLaunchKernelGGL((gemm_update_kernel<T, UPD_TRAP, LB>), dim3((unsigned)(tm * tn)),
                               dim3(256), 0, s, A, lda, B, ldb, M, N, Kd, C, ldc, info, tm, 0,
                               (int64_t)0, PivotRule());
            break;
        case UPD_TRAP_POTF2:
            hipLaunchKernelGGL((gemm_update_kernel<T, UPD_TRAP_POTF2, LB>), dim3((unsigned)(tm * tn)),
                               dim3(256), 0, s, A, lda, B, ldb, M, N, Kd, C, ldc, info, tm, w_next,
                               col_base_next, pr);
            break;
        default:
            hipLaunchKernelGGL((gemm_update_kernel<T, UPD_FULL, LB>), dim3((unsigned)(tm * tn)),
                               dim3(256), 0, s, A, lda, B, ldb, M, N, Kd, C, ldc, info, tm, 0,
                               (int64_t)0, PivotRule());
    }
    QMLH_CHECK_LAUNCH("gemm_update launch");
    return QMLH_OK;
}

// 64-wide sub-panel: the diagonal block (one workgroup), then the rows below it.
template <typename T>
__global__ __launch_bounds__(256) void potf2_kernel(T* __restrict__ A, int64_t lda, int w,
                                                    int64_t col_base, int32_t* __restrict__ info,
                                                    PivotRule pr) {
    potf2_body<T>(A, lda, w, col_base, info, pr.floor, pr.bshift, pr.shift_acc, pr.bpivot);
}

template <typename T>
__global__ __launch_bounds__(256) void trsm_rows_kernel(T* __restrict__ A, int64_t lda, int w,
                                                        int64_t m,
                                                        const int32_t* __restrict__ info) {
    trsm_body<T>(A, lda, w, m, info, blockIdx.x);
}

__global__ void clear_info_kernel(int32_t* info) { *info = 0; }

// Factor the block column A[k:rows, k:k+kb] (its trailing updates from earlier columns
// applied): per 64-wide sub-panel the diagonal factor, the rows below, and the update of the
// rest of the block column. rows = n, or n + r when r extra rows below the matrix are solved
// along (they are panel rows like any other).
// The update after each sub-panel also factors the next sub-panel's diagonal block
// (UPD_TRAP_POTF2), so only the first diagonal block of the column has a launch of its own.
template <typename T>
static int factor_block_column(T* A, int64_t lda, int64_t rows, int64_t k, int64_t kb,
                               int32_t* info, hipStream_t s, hipEvent_t rows_wait = nullptr,
                               PivotRule pr = PivotRule()) {
    for (int64_t jj = 0; jj < kb; jj += PF_IB) {
        const int64_t c = k + jj;
        const int w = (int)(kb - jj < PF_IB ? kb - jj : PF_IB);
        T* Acc = A + c + c * lda;
        const int64_t m = rows - c - w;  // rows below the diagonal block
        if (jj == 0) {
            hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(256), 0, s, Acc, lda, w, c, info,
                               pr);
            QMLH_CHECK_LAUNCH("potf2 launch");
        }
        if (jj == 0 && rows_wait)  // the rows' last update may still run on another stream
            QMLH_TRY(check_hip(hipStreamWaitEvent(s, rows_wait, 0), "potrf: wait"));
        if (m > 0) {
            hipLaunchKernelGGL(trsm_rows_kernel<T>, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0,
                               s, Acc, lda, w, m, info);
            QMLH_CHECK_LAUNCH("panel rows launch");
        }
        const int64_t ncols = kb - jj - w;  // rest of this block column
        if (ncols > 0) {
            const int w2 = (int)(ncols < PF_IB ? ncols : PF_IB);
            QMLH_TRY(launch_update<T>(UPD_TRAP_POTF2, Acc + w, lda, Acc + w, lda, m, ncols, w,
                                      Acc + w + w * lda, lda, info, s, -1, w2, c + w, pr));
        }
    }
    return QMLH_OK;
}

// Extra rows A[n:n+r, c0:n] -= A[n:n+r, k:k+kb] A[c0:n, k:k+kb]^T: the trailing update of the
// r rows solved along with the factorisation (FP32/FP64 MFMA in the storage type, also when
// the matrix itself takes split-FP16 updates).
template <typename T>
static int extra_rows_update(T* A, int64_t lda, int64_t n, int64_t r, int64_t k, int64_t kb,
                             int64_t c0, int32_t* info, hipStream_t s) {
    if (r <= 0 || c0 >= n) return QMLH_OK;
    return launch_update<T>(UPD_FULL, A + n + k * lda, lda, A + c0 + k * lda, lda, r, n - c0, kb,
                            A + n + c0 * lda, lda, info, s);
}

// Lower triangle of C (m x m) -= P P^T, P (m x kb) the panel below the block column.
template <typename T>
static int trailing_syrk(T* P, int64_t lda, int64_t m, int64_t kb, T* C, int32_t* info,
                         hipStream_t s, void* split_ws) {
    if (m <= 0) return QMLH_OK;
    if (std::is_same<T, float>::value && split_ws)
        return syrk_f16x3(reinterpret_cast<const float*>(P), lda, m, kb,
                           reinterpret_cast<float*>(C), lda, info, split_ws, s);
    return launch_update<T>(UPD_SYRK, P, lda, P, lda, m, m, kb, C, lda, info, s);
}

// Matrices at least this large are factored with look-ahead.
static constexpr int64_t PF_LOOKAHEAD_MIN = 4096;
// Trailing matrices at least this large keep the extra rows' update on the side stream.
static constexpr int64_t PF_EXT_SIDE_MIN = 20000;

// Look-ahead schedule. Step k's trailing update is split into the next block column's part
// (an MFMA update on the caller's stream) and the rest; the next block column is factored on a
// second stream while the rest of the update runs, so the latency-bound panel work overlaps the
// bandwidth/MFMA-bound trailing update. Events order the two streams; the caller's stream waits
// for the last panel, so on return all work is ordered on `s`.
//
// With r extra rows (A[n:n+r, :]), the next block column's update and panel include them. The
// rest of their trailing update (an FP32 product) touches nothing the trailing update or the
// panel reads or writes: while the trailing matrix is large (the caller's stream is the
// critical path) it runs on the side stream after the panel, later (the panels are the
// critical path) on the caller's stream after the trailing update.
// The look-ahead's side streams and events, created once and reused: a per-device pool, one
// context per factorisation in flight on that device (concurrent callers take different ones),
// returned when the factorisation's launches are enqueued. Streams are non-blocking at the
// default priority.
struct LookaheadCtx {
    int device = -1;
    hipStream_t s2 = nullptr, s3 = nullptr;
    hipEvent_t ev_a = nullptr, ev_p = nullptr, ev_x = nullptr, ev_e = nullptr, ev_b = nullptr;
    hipEvent_t ev_fa = nullptr, ev_r = nullptr;
};

static std::mutex g_la_mu;
static std::vector<LookaheadCtx*> g_la_free;

static int lookahead_acquire(LookaheadCtx** out) {
    int dev = 0;
    QMLH_TRY(check_hip(hipGetDevice(&dev), "potrf: device"));
    {
        std::lock_guard<std::mutex> lock(g_la_mu);
        for (size_t i = 0; i < g_la_free.size(); ++i) {
            if (g_la_free[i]->device == dev) {
                *out = g_la_free[i];
                g_la_free.erase(g_la_free.begin() + (int64_t)i);
                return QMLH_OK;
            }
        }
    }
    LookaheadCtx* c = new LookaheadCtx;
    c->device = dev;
    int rc = check_hip(hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking), "potrf: side stream");
    if (rc == QMLH_OK)
        rc = check_hip(hipStreamCreateWithFlags(&c->s3, hipStreamNonBlocking), "potrf: third stream");
    hipEvent_t* evs[7] = {&c->ev_a, &c->ev_p, &c->ev_x, &c->ev_e, &c->ev_b, &c->ev_fa, &c->ev_r};
    for (int i = 0; i < 7 && rc == QMLH_OK; ++i)
        rc = check_hip(hipEventCreateWithFlags(evs[i], hipEventDisableTiming), "potrf: event");
    if (rc != QMLH_OK) {  // nothing was enqueued on them yet
        if (c->s2) (void)hipStreamDestroy(c->s2);
        if (c->s3) (void)hipStreamDestroy(c->s3);
        for (int i = 0; i < 7; ++i)
            if (*evs[i]) (void)hipEventDestroy(*evs[i]);
        delete c;
        return rc;
    }
    *out = c;
    return QMLH_OK;
}

// Returning a context once its launches are enqueued is safe: the next user's records and
// waits are ordered after this factorisation's on the same streams, and the caller's stream
// already waits for this factorisation's last side-stream work.
static void lookahead_release(LookaheadCtx* c) {
    std::lock_guard<std::mutex> lock(g_la_mu);
    g_la_free.push_back(c);
}

template <typename T>
static int potrf_lookahead(T* A, int64_t n, int64_t lda, int64_t r, int32_t* info,
                           hipStream_t s, int64_t nb, PivotRule pf) {
    LookaheadCtx* ctx = nullptr;
    QMLH_TRY(lookahead_acquire(&ctx));
    hipStream_t s2 = ctx->s2;
    hipEvent_t ev_a = ctx->ev_a, ev_p = ctx->ev_p;
    int rc = factor_block_column<T>(A, lda, n + r, 0, n < nb ? n : nb, info, s, nullptr, pf);
    for (int64_t k = 0; rc == QMLH_OK && k < n; k += nb) {
        const int64_t kb = n - k < nb ? n - k : nb;
        const int64_t m = n - k - kb;
        if (m <= 0) break;
        const int64_t kb2 = m < nb ? m : nb;  // next block column
        T* P = A + (k + kb) + k * lda;
        T* C = P + kb * lda;  // A(k + kb, k + kb)
        rc = launch_update<T>(UPD_TRAP, P, lda, P, lda, m + r, kb2, kb, C, lda, info, s);
        if (rc == QMLH_OK) rc = check_hip(hipEventRecord(ev_a, s), "potrf: record");
        if (rc == QMLH_OK) rc = check_hip(hipStreamWaitEvent(s2, ev_a, 0), "potrf: wait");
        if (rc == QMLH_OK)
            rc = factor_block_column<T>(A, lda, n + r, k + kb, kb2, info, s2, nullptr, pf);
        const bool ext_side = m >= PF_EXT_SIDE_MIN;
        if (rc == QMLH_OK && ext_side)
            rc = extra_rows_update<T>(A, lda, n, r, k, kb, k + kb + kb2, info, s2);
        if (rc == QMLH_OK) rc = check_hip(hipEventRecord(ev_p, s2), "potrf: record");
        if (rc == QMLH_OK && m > kb2)
            rc = trailing_syrk<T>(P + kb2, lda, m - kb2, kb, C + kb2 + kb2 * lda, info, s,
                                  nullptr);
        if (rc == QMLH_OK && !ext_side)
            rc = extra_rows_update<T>(A, lda, n, r, k, kb, k + kb + kb2, info, s);
        if (rc == QMLH_OK) rc = check_hip(hipStreamWaitEvent(s, ev_p, 0), "potrf: wait");
    }
    // `s` already waits for the last panel, so every later use of these streams is ordered
    // after this one.
    lookahead_release(ctx);
    return rc;
}

// Split-FP16 look-ahead in groups of two block columns. The panels of a group are adjacent
// columns, so the group's rows below it form one column-major block of 2 nb columns: it is
// split once, and the next group's columns and the rest of the trailing matrix each take one
// update with 2 nb features, which halves the reads and writes of the FP32 trailing matrix per
// product against one update per block column. Per group:
//   caller's stream: split of the group (all rows below it, region 1) -> update of the next
//     group's first block column -> [event a] -> update of its second block column ->
//     [event b] -> update of the rest of the trailing matrix;
//   side stream, after event a: the next group -- its first block column factored, split
//     (region 0), then after event b the second block column updated from it and factored.
// The caller's stream then waits for the side stream before the next group's split. With r
// extra rows (A[n:n+r, :]) the next group's part of their update (an FP32 product) runs on a
// third stream that the side stream waits for before solving rows, and the rest of their
// update right after it on the third stream (it writes only extra rows of later columns,
// which only later work on that stream reads, and whose solve waits for its next event); for
// r <= 2 nb both are one split-FP16 product (f16x3_gemm_rows). The third stream also takes the
// extra rows' update of the next group's second block column from its first (factor_group_split).
static constexpr int64_t PF_GROUP = 2 * PF_NB_SPLIT;

static size_t split_group_region(int64_t n) { return f16x3_region_bytes(n, PF_NB_SPLIT); }

// Factor the w (<= 2 nb) columns from k, whose updates from earlier columns are applied (to
// the second block column once b_ready has fired, when given). With r extra rows and a third
// stream sx (events ev_fa, ev_r), their update of the second block column from the first runs
// on sx beside the split-FP16 update of the matrix rows (it reads the factored first block
// column and writes only extra rows), and the second block column's factor waits for it.
static int factor_group_split(float* A, int64_t lda, int64_t n, int64_t r, int64_t k, int64_t w,
                              int32_t* info, hipStream_t st, hipEvent_t rows_wait,
                              hipEvent_t b_ready, void* ws, PivotRule pf,
                              hipStream_t sx = nullptr, hipEvent_t ev_fa = nullptr,
                              hipEvent_t ev_r = nullptr) {
    const int64_t wa = w < PF_NB_SPLIT ? w : PF_NB_SPLIT;
    QMLH_TRY(factor_block_column<float>(A, lda, n + r, k, wa, info, st, rows_wait, pf));
    const int64_t wb = w - wa;
    if (wb <= 0) return QMLH_OK;
    const int64_t ma = n - k - wa;  // rows of the matrix below the first block column
    float* Pa = A + (k + wa) + k * lda;
    float* Cb = Pa + wa * lda;  // A(k + wa, k + wa)
    const bool side = r > 0 && sx && ev_fa && ev_r;
    if (side) {
        QMLH_TRY(check_hip(hipEventRecord(ev_fa, st), "potrf: record"));
        QMLH_TRY(check_hip(hipStreamWaitEvent(sx, ev_fa, 0), "potrf: wait"));
        QMLH_TRY(launch_update<float>(UPD_FULL, Pa + ma, lda, Pa, lda, r, wb, wa, Cb + ma, lda,
                                      info, sx));
        QMLH_TRY(check_hip(hipEventRecord(ev_r, sx), "potrf: record"));
    }
    QMLH_TRY(f16x3_split(Pa, lda, ma, wa, ws, st, 0));
    if (b_ready) QMLH_TRY(check_hip(hipStreamWaitEvent(st, b_ready, 0), "potrf: wait"));
    QMLH_TRY(f16x3_update(ws, ma, wa, 0, ma, wb, Cb, lda, info, st, 0));
    if (side) {
        QMLH_TRY(check_hip(hipStreamWaitEvent(st, ev_r, 0), "potrf: wait"));
    } else if (r > 0) {
        QMLH_TRY(launch_update<float>(UPD_FULL, Pa + ma, lda, Pa, lda, r, wb, wa, Cb + ma, lda,
                                      info, st));
    }
    return factor_block_column<float>(A, lda, n + r, k + wa, wb, info, st, nullptr, pf);
}

static int potrf_lookahead_split(float* A, int64_t n, int64_t lda, int64_t r, int32_t* info,
                                 hipStream_t s, void* ws, PivotRule pf) {
    LookaheadCtx* ctx = nullptr;
    QMLH_TRY(lookahead_acquire(&ctx));
    hipStream_t s2 = ctx->s2, s3 = ctx->s3;
    hipEvent_t ev_a = ctx->ev_a, ev_p = ctx->ev_p, ev_x = ctx->ev_x, ev_e = ctx->ev_e;
    hipEvent_t ev_b = ctx->ev_b;
    const size_t region = split_group_region(n);
    const size_t region_x = region + f16x3_region_bytes(n - PF_GROUP, PF_GROUP);
    int64_t k = 0, w = n < PF_GROUP ? n : PF_GROUP;
    int rc = factor_group_split(A, lda, n, r, 0, w, info, s, nullptr, nullptr, ws, pf);
    while (rc == QMLH_OK && k + w < n) {
        const int64_t m = n - k - w;                  // rows below the group
        const int64_t w2 = m < PF_GROUP ? m : PF_GROUP;  // the next group
        float* P = A + (k + w) + k * lda;             // the group's rows below it (m x w)
        float* C = P + w * lda;                       // A(k + w, k + w)
        hipEvent_t rows_wait = nullptr;
        if (rc == QMLH_OK) rc = f16x3_split(P, lda, m, w, ws, s, region);
        if (rc == QMLH_OK && r > 0) {
            // the extra rows' update by this group, on the third stream: with r <= 2 nb one
            // split-FP16 product over all later columns (the extra rows split with their own
            // scale into the tail of the group's region, which a group split never reaches:
            // it has at most n - 2 nb rows); else FP32 products, the next group's columns
            // first. The next group's rows solve waits for the event after them.
            rc = check_hip(hipEventRecord(ev_x, s), "potrf: record");
            if (rc == QMLH_OK) rc = check_hip(hipStreamWaitEvent(s3, ev_x, 0), "potrf: wait");
            if (r <= PF_GROUP) {
                if (rc == QMLH_OK)
                    rc = f16x3_gemm_rows(P + m, lda, r, w, ws, region_x, region, m, C + m, lda,
                                         info, s3);
                if (rc == QMLH_OK) rc = check_hip(hipEventRecord(ev_e, s3), "potrf: record");
            } else {
                if (rc == QMLH_OK)
                    rc = launch_update<float>(UPD_FULL, P + m, lda, P, lda, r, w2, w, C + m, lda,
                                              info, s3);
                if (rc == QMLH_OK) rc = check_hip(hipEventRecord(ev_e, s3), "potrf: record");
                if (rc == QMLH_OK)
                    rc = extra_rows_update<float>(A, lda, n, r, k, w, k + w + w2, info, s3);
            }
            rows_wait = ev_e;
        }
        const int64_t wa2 = w2 < PF_NB_SPLIT ? w2 : PF_NB_SPLIT;  // next group's first block column
        if (rc == QMLH_OK)
            rc = f16x3_update(ws, m, w, 0, m, wa2, C, lda, info, s, region, true);
        if (rc == QMLH_OK) rc = check_hip(hipEventRecord(ev_a, s), "potrf: record");
        if (rc == QMLH_OK) rc = check_hip(hipStreamWaitEvent(s2, ev_a, 0), "potrf: wait");
        if (rc == QMLH_OK && w2 > wa2) {
            rc = f16x3_update(ws, m, w, wa2, m - wa2, w2 - wa2, C + wa2 + wa2 * lda, lda, info, s,
                              region, true);
            if (rc == QMLH_OK) rc = check_hip(hipEventRecord(ev_b, s), "potrf: record");
        }
        if (rc == QMLH_OK)
            rc = factor_group_split(A, lda, n, r, k + w, w2, info, s2, rows_wait,
                                    w2 > wa2 ? ev_b : nullptr, ws, pf, s3, ctx->ev_fa,
                                    ctx->ev_r);
        if (rc == QMLH_OK) rc = check_hip(hipEventRecord(ev_p, s2), "potrf: record");
        if (rc == QMLH_OK && m > w2)
            rc = f16x3_update(ws, m, w, w2, m - w2, 0, C + w2 + w2 * lda, lda, info, s, region);
        if (rc == QMLH_OK) rc = check_hip(hipStreamWaitEvent(s, ev_p, 0), "potrf: wait");
        k += w;
        w = w2;
    }
    // `s` already waits for the last group, whose rows waited for the third stream's last
    // work (the last group's "rest" is empty), so every later use of these streams is ordered
    // after this one.
    lookahead_release(ctx);
    return rc;
}

template <typename T>
int potrf_impl(T* A, int64_t n, int64_t lda, int32_t* info, hipStream_t s, void* split_ws,
               int64_t r, PivotRule pr) {
    if (n < 0 || r < 0) return set_error(QMLH_ERR_ARG, "potrf: n < 0 or r < 0");
    if (lda < (n + r > 1 ? n + r : 1)) return set_error(QMLH_ERR_ARG, "potrf: lda < n + r");
    if (!info) return set_error(QMLH_ERR_ARG, "potrf: info is null");
    hipLaunchKernelGGL(clear_info_kernel, dim3(1), dim3(1), 0, s, info);
    QMLH_CHECK_LAUNCH("potrf: clear info");
    if (n == 0) return QMLH_OK;
    if (!A) return set_error(QMLH_ERR_ARG, "potrf: A is null");
    TimerScope ts(QMLH_TIMER_POTRF, (double)n * (double)n * (double)n / 3.0, s);
    const bool split16 = std::is_same<T, float>::value && split_ws;
    const int64_t nb = split16 ? PF_NB_SPLIT : PF_NB;
    if (split16) QMLH_TRY(f16x3_prepare(reinterpret_cast<const float*>(A), n, lda, split_ws, s));
    if (n >= PF_LOOKAHEAD_MIN) {
        if constexpr (std::is_same<T, float>::value)
            if (split16)
                return potrf_lookahead_split(A, n, lda, r, info, s, split_ws, pr);
        return potrf_lookahead<T>(A, n, lda, r, info, s, nb, pr);
    }
    for (int64_t k = 0; k < n; k += nb) {
        const int64_t kb = n - k < nb ? n - k : nb;
        QMLH_TRY(factor_block_column<T>(A, lda, n + r, k, kb, info, s, nullptr, pr));
        const int64_t m = n - k - kb;
        T* P = A + (k + kb) + k * lda;
        QMLH_TRY(trailing_syrk<T>(P, lda, m, kb, P + kb * lda, info, s, split16 ? split_ws : nullptr));
        QMLH_TRY(extra_rows_update<T>(A, lda, n, r, k, kb, k + kb, info, s));
    }
    return QMLH_OK;
}

// ------------------------------------------------------------------------------------------
// Triangular solves, 64-row steps, one fused kernel per step. Every workgroup of a step loads
// the 64 x 64 diagonal block into LDS and solves it redundantly (one wavefront, FP64), then
// applies its slice of the update; the solved block goes to a separate output vector so no
// workgroup reads what another one writes within the step.
// ------------------------------------------------------------------------------------------

// Load L(k0 + i, k0 + j), i, j < w (lower part) into LDS and the reciprocal pivots; returns
// after a barrier.
template <typename T>
__device__ __forceinline__ void load_diag_block(double (*sL)[PF_IB + 1], double* sInv,
                                                const T* __restrict__ L, int64_t ldl, int64_t k0,
                                                int w) {
    stage_lower64<false, PF_IB + 1>(sL, L + k0 + k0 * ldl, ldl, w);
    __syncthreads();
    if (threadIdx.x < w) sInv[threadIdx.x] = 1.0 / sL[threadIdx.x][threadIdx.x];
    __syncthreads();
}

// Forward step: z[k0 .. k0+w) = L_kk^-1 r[k0 .. k0+w); r[i] -= L(i, k0..k0+w) z_k for the
// rows i in [k0 + w, n) this workgroup owns. FULL: w == 64 (loops fully unrolled).
template <typename T, bool FULL>
__global__ __launch_bounds__(256) void trsv_fwd_step_kernel(const T* __restrict__ L, int64_t ldl,
                                                            int64_t n, int64_t k0, int w_,
                                                            double* __restrict__ r,
                                                            double* __restrict__ z) {
    __shared__ double sL[PF_IB][PF_IB + 1];
    __shared__ double sInv[PF_IB];
    __shared__ double sz[PF_IB];
    const int w = FULL ? PF_IB : w_;
    load_diag_block(sL, sInv, L, ldl, k0, w);
    if (threadIdx.x < 64) {
        const int i = threadIdx.x;
        double xi = (i < w) ? r[k0 + i] : 0.0;
        for (int c = 0; c < w; ++c) {
            const double xc = __shfl(xi * sInv[c], c, 64);
            if (i == c) xi = xc;
            if (i > c && i < w) xi = xi - sL[i][c] * xc;
        }
        if (i < w) {
            sz[i] = xi;
            if (blockIdx.x == 0) z[k0 + i] = xi;
        }
    }
    __syncthreads();
    const int64_t row = k0 + w + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const T* lr = L + row + k0 * ldl;
    double s = 0.0;
    if (FULL) {
        double v[PF_IB];
#pragma unroll
        for (int c = 0; c < PF_IB; ++c) v[c] = (double)lr[c * ldl];
#pragma unroll
        for (int c = 0; c < PF_IB; ++c) s = fma(v[c], sz[c], s);
    } else {
        for (int c = 0; c < w; ++c) s = fma((double)lr[c * ldl], sz[c], s);
    }
    r[row] = r[row] - s;
}

// Backward step: x[k0 .. k0+w) = L_kk^-T z[k0 .. k0+w); z[i] -= sum_c L(k0 + c, i) x_c for
// the rows i < k0 this workgroup owns (column i of L, rows k0 .. k0+w, is contiguous).
template <typename T, bool FULL>
__global__ __launch_bounds__(256) void trsv_bwd_step_kernel(const T* __restrict__ L, int64_t ldl,
                                                            int64_t k0, int w_,
                                                            double* __restrict__ z,
                                                            double* __restrict__ x) {
    __shared__ double sL[PF_IB][PF_IB + 1];
    __shared__ double sInv[PF_IB];
    __shared__ double sx[PF_IB];
    const int w = FULL ? PF_IB : w_;
    load_diag_block(sL, sInv, L, ldl, k0, w);
    if (threadIdx.x < 64) {
        const int i = threadIdx.x;
        double xi = (i < w) ? z[k0 + i] : 0.0;
        for (int c = w - 1; c >= 0; --c) {
            const double xc = __shfl(xi * sInv[c], c, 64);
            if (i == c) xi = xc;
            if (i < c) xi = xi - sL[c][i] * xc;
        }
        if (i < w) {
            sx[i] = xi;
            if (blockIdx.x == 0) x[k0 + i] = xi;
        }
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k0) return;
    const T* lc = L + k0 + i * ldl;
    double s = 0.0;
    if (FULL) {
        double v[PF_IB];
#pragma unroll
        for (int c = 0; c < PF_IB; ++c) v[c] = (double)lc[c];
#pragma unroll
        for (int c = 0; c < PF_IB; ++c) s = fma(v[c], sx[c], s);
    } else {
        for (int c = 0; c < w; ++c) s = fma((double)lc[c], sx[c], s);
    }
    z[i] = z[i] - s;
}

// Below this order the per-step kernels are used (few steps, no spin-waiting).

template <typename T>
int trsv_lower_fwd(const T* L, int64_t n, int64_t lda, double* r, double* z, int32_t* sync,
                   hipStream_t s) {
    if (n >= TRSV_SYNCFREE_MIN && sync) return trsv_syncfree<T>(L, n, lda, r, z, sync, true, s);
    for (int64_t k = 0; k < n; k += PF_IB) {
        const int w = (int)(n - k < PF_IB ? n - k : PF_IB);
        const int64_t rest = n - k - w;
        const unsigned grid = (unsigned)(rest > 0 ? ceil_div(rest, 256) : 1);
        if (w == PF_IB)
            hipLaunchKernelGGL((trsv_fwd_step_kernel<T, true>), dim3(grid), dim3(256), 0, s, L,
                               lda, n, k, w, r, z);
        else
            hipLaunchKernelGGL((trsv_fwd_step_kernel<T, false>), dim3(grid), dim3(256), 0, s, L,
                               lda, n, k, w, r, z);
        QMLH_CHECK_LAUNCH("trsv forward launch");
    }
    return QMLH_OK;
}

template <typename T>
int trsv_lower_bwd(const T* L, int64_t n, int64_t lda, double* z, double* x, int32_t* sync,
                   hipStream_t s) {
    if (n >= TRSV_SYNCFREE_MIN && sync) return trsv_syncfree<T>(L, n, lda, z, x, sync, false, s);
    const int64_t last = ((n - 1) / PF_IB) * PF_IB;
    for (int64_t k = last; k >= 0; k -= PF_IB) {
        const int w = (int)(n - k < PF_IB ? n - k : PF_IB);
        const unsigned grid = (unsigned)(k > 0 ? ceil_div(k, 256) : 1);
        if (w == PF_IB)
            hipLaunchKernelGGL((trsv_bwd_step_kernel<T, true>), dim3(grid), dim3(256), 0, s, L,
                               lda, k, w, z, x);
        else
            hipLaunchKernelGGL((trsv_bwd_step_kernel<T, false>), dim3(grid), dim3(256), 0, s, L,
                               lda, k, w, z, x);
        QMLH_CHECK_LAUNCH("trsv backward launch");
    }
    return QMLH_OK;
}

// x := (L L^T)^-1 x with L lower, storage T; `tmp` is an n-vector of scratch.
template <typename T>
int potrs_impl(const T* L, int64_t n, int64_t lda, double* x, double* tmp, int32_t* sync,
               hipStream_t s) {
    if (n < 0) return set_error(QMLH_ERR_ARG, "potrs: n < 0");
    if (lda < (n > 1 ? n : 1)) return set_error(QMLH_ERR_ARG, "potrs: lda < n");
    if (n == 0) return QMLH_OK;
    if (!L || !x || !tmp) return set_error(QMLH_ERR_ARG, "potrs: null pointer");
    TimerScope ts(QMLH_TIMER_POTRS, 2.0 * (double)n * (double)n, s);
    // Forward L z = y: x holds y and is consumed as scratch, z lands in tmp; backward L^T x = z:
    // tmp is consumed, the solution lands in x.
    QMLH_TRY(trsv_lower_fwd<T>(L, n, lda, x, tmp, sync, s));
    return trsv_lower_bwd<T>(L, n, lda, tmp, x, sync, s);
}

// ------------------------------------------------------------------------------------------
// Multiple right-hand sides: V := L^-1 V (V n x nrhs column-major, FP32 or FP64), right-looking in
// 256-row blocks. Within a block, each 64-row step multiplies by the FP64 inverse of the 64 x 64
// diagonal block (trsv_diag_inverse) and updates the rest of the block; then one FP32 MFMA
// product updates every row below the 256-row block, so V streams through HBM once per 256
// rows instead of once per 64.
// ------------------------------------------------------------------------------------------
// V[0:w, j0 .. j0+63] := D V[0:w, j0 ..] with D the 64 x 64 FP64 inverse (zero outside w).
template <typename T>
__global__ __launch_bounds__(256) void dinv_apply_kernel(const double* __restrict__ D,
                                                         T* __restrict__ V, int64_t ldv,
                                                         int w, int64_t nrhs) {
    __shared__ double sD[PF_IB][PF_IB + 1];  // sD[i][c] = D(i, c)
    __shared__ double sV[PF_IB][PF_IB + 1];  // sV[c][j] = V(c, j0 + j)
    const int t = threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.x * PF_IB;
    double dv[PF_IB * PF_IB / 256], vv[PF_IB * PF_IB / 256];
#pragma unroll
    for (int k = 0; k < PF_IB * PF_IB / 256; ++k) {
        const int idx = t + 256 * k;
        const int r = idx & (PF_IB - 1), c = idx / PF_IB;  // consecutive lanes: consecutive rows
        dv[k] = D[idx];
        vv[k] = (r < w && j0 + c < nrhs) ? (double)V[r + (j0 + c) * ldv] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PF_IB * PF_IB / 256; ++k) {
        const int idx = t + 256 * k;
        const int r = idx & (PF_IB - 1), c = idx / PF_IB;
        sD[r][c] = dv[k];
        sV[r][c] = vv[k];
    }
    __syncthreads();
    const int i = t & (PF_IB - 1), q = t >> 6;  // row i, columns 16 q ..
    double o[16];
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) o[jj] = 0.0;
    for (int c = 0; c < PF_IB; ++c) {
        const double dic = sD[i][c];
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) o[jj] = fma(dic, sV[c][16 * q + jj], o[jj]);
    }
    if (i < w) {
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int64_t j = j0 + 16 * q + jj;
            if (j < nrhs) V[i + j * ldv] = (T)o[jj];
        }
    }
}

template <typename T>
int trsm_lower_fwd_multi(const T* L, int64_t n, int64_t ldl, T* V, int64_t ldv, int64_t nrhs,
                         const double* dinv, hipStream_t s, bool lower_rhs) {
    if (n <= 0 || nrhs <= 0) return QMLH_OK;
    constexpr int64_t NBM = 256;
    // With a lower-triangular right-hand side (column j zero above row j, e.g. the identity),
    // rows r.. only involve columns j < r + width: the others are still zero there.
    auto cols = [&](int64_t row_end) { return lower_rhs && row_end < nrhs ? row_end : nrhs; };
    for (int64_t k = 0; k < n; k += NBM) {
        const int64_t kb = n - k < NBM ? n - k : NBM;
        for (int64_t jj = 0; jj < kb; jj += PF_IB) {
            const int64_t c = k + jj;
            const int w = (int)(kb - jj < PF_IB ? kb - jj : PF_IB);
            const int64_t nc = cols(c + w);
            hipLaunchKernelGGL(dinv_apply_kernel<T>, dim3((unsigned)ceil_div(nc, PF_IB)), dim3(256),
                               0, s, dinv + (c / PF_IB) * (PF_IB * PF_IB), V + c, ldv, w, nc);
            QMLH_CHECK_LAUNCH("trsm multi diagonal launch");
            // rest of the 256-row block: V[c+w : k+kb] -= L[c+w : k+kb, c : c+w] V[c : c+w];
            // B(j, kk) = V[(c + kk) + j ldv] (row layout).
            const int64_t rest_in = kb - jj - w;
            QMLH_TRY((launch_update<T, LAYOUT_ROW>(UPD_FULL, L + (c + w) + c * ldl, ldl, V + c,
                                                       ldv, rest_in, nc, w, V + (c + w), ldv,
                                                       nullptr, s)));
        }
        const int64_t rest = n - k - kb;
        QMLH_TRY((launch_update<T, LAYOUT_ROW>(UPD_FULL, L + (k + kb) + k * ldl, ldl, V + k,
                                                   ldv, rest, cols(k + kb), kb, V + (k + kb), ldv,
                                                   nullptr, s)));
    }
    return QMLH_OK;
}

// ------------------------------------------------------------------------------------------
// V := L^-1 V with an FP32 factor L and an FP64 right-hand side, in FP64 arithmetic: the same
// 256-row blocked scheme, with each 256-wide block column of L (from its diagonal block down)
// first widened to FP64 in a panel buffer, so every product is an FP64 MFMA update. The
// result is L^-1 V for the stored (FP32-rounded) L to FP64 accuracy, which a preconditioner
// built from that L needs when V's directions reach L's smallest pivots.
// ------------------------------------------------------------------------------------------
static constexpr int64_t TRSM64_NB = 256;

// P(i, j) = (double)L(i, j), i < rows, j < cols (column-major, grid-stride).
__global__ __launch_bounds__(256) void widen_block_kernel(const float* __restrict__ L,
                                                          int64_t ldl, int64_t rows,
                                                          int64_t cols, double* __restrict__ P,
                                                          int64_t ldp) {
    const int64_t total = rows * cols;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % rows, j = idx / rows;
        P[i + j * ldp] = (double)L[i + j * ldl];
    }
}

size_t trsm_f32l_panel_doubles(int64_t n) { return (size_t)(n > 0 ? n : 1) * TRSM64_NB; }

int trsm_lower_fwd_f32l(const float* L, int64_t n, int64_t ldl, double* V, int64_t ldv,
                        int64_t nrhs, const double* dinv, double* panel, hipStream_t s) {
    if (n <= 0 || nrhs <= 0) return QMLH_OK;
    const int64_t ldp = n;
    for (int64_t k = 0; k < n; k += TRSM64_NB) {
        const int64_t kb = n - k < TRSM64_NB ? n - k : TRSM64_NB;
        const int64_t rows = n - k;
        const int64_t total = rows * kb;
        const unsigned g = (unsigned)(ceil_div(total, (int64_t)256) < 8192 ? ceil_div(total, (int64_t)256) : 8192);
        hipLaunchKernelGGL(widen_block_kernel, dim3(g), dim3(256), 0, s, L + k + k * ldl, ldl, rows,
                           kb, panel, ldp);
        QMLH_CHECK_LAUNCH("trsm f32l widen launch");
        for (int64_t jj = 0; jj < kb; jj += PF_IB) {
            const int64_t c = k + jj;
            const int w = (int)(kb - jj < PF_IB ? kb - jj : PF_IB);
            hipLaunchKernelGGL(dinv_apply_kernel<double>, dim3((unsigned)ceil_div(nrhs, (int64_t)PF_IB)),
                               dim3(256), 0, s, dinv + (c / PF_IB) * (PF_IB * PF_IB), V + c, ldv, w,
                               nrhs);
            QMLH_CHECK_LAUNCH("trsm f32l diagonal launch");
            const int64_t rest_in = kb - jj - w;
            QMLH_TRY((launch_update<double, LAYOUT_ROW>(UPD_FULL, panel + (jj + w) + jj * ldp, ldp,
                                                        V + c, ldv, rest_in, nrhs, w, V + (c + w),
                                                        ldv, nullptr, s)));
        }
        QMLH_TRY((launch_update<double, LAYOUT_ROW>(UPD_FULL, panel + kb, ldp, V + k, ldv,
                                                    n - k - kb, nrhs, kb, V + (k + kb), ldv,
                                                    nullptr, s)));
    }
    return QMLH_OK;
}

// out(j, i) = in(i, j) for the R x Cc column-major `in`: 64 x 64 tiles through LDS (padded
// rows, so the transposed reads hit distinct banks); one-dimensional grid of tiles.
__global__ __launch_bounds__(256) void transpose_f32_kernel(const float* __restrict__ in,
                                                            int64_t ldi, int64_t R, int64_t Cc,
                                                            float* __restrict__ out, int64_t ldo,
                                                            int64_t tiles_r) {
    __shared__ float t[64][65];
    const int64_t r0 = (blockIdx.x % tiles_r) * 64, c0 = (blockIdx.x / tiles_r) * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int q = ty; q < 64; q += 4) {
        const int64_t i = r0 + tx, j = c0 + q;
        if (i < R && j < Cc) t[q][tx] = in[i + j * ldi];
    }
    __syncthreads();
#pragma unroll
    for (int q = ty; q < 64; q += 4) {
        const int64_t i = r0 + q, j = c0 + tx;
        if (i < R && j < Cc) out[j + i * ldo] = t[tx][q];
    }
}

static int transpose_f32(const float* in, int64_t ldi, int64_t R, int64_t Cc, float* out,
                         int64_t ldo, hipStream_t s) {
    if (R <= 0 || Cc <= 0) return QMLH_OK;
    const int64_t tr = ceil_div(R, 64), tc = ceil_div(Cc, 64);
    hipLaunchKernelGGL(transpose_f32_kernel, dim3((unsigned)(tr * tc)), dim3(256), 0, s, in, ldi,
                       R, Cc, out, ldo, tr);
    QMLH_CHECK_LAUNCH("transpose launch");
    return QMLH_OK;
}

template int trsm_lower_fwd_multi<float>(const float*, int64_t, int64_t, float*, int64_t, int64_t,
                                         const double*, hipStream_t, bool);

// ------------------------------------------------------------------------------------------
// Inverse of an SPD matrix from its Cholesky factor (the drop-in for dpotrf + dpotri):
//   W := L^-1 by the multi-right-hand-side forward solve of the identity (the columns' entries
//   above the diagonal stay exact zeros), A := W^T (transpose), W := -(A A^T) on the lower
//   tiles (FP64 MFMA rank-n update from zero), then A(i, j) = A(j, i) = -W(i, j), i >= j.
// ------------------------------------------------------------------------------------------
// W := diag * I (n x n): the identity, or zero.
__global__ __launch_bounds__(256) void identity_f64_kernel(double* __restrict__ W, int64_t ldw,
                                                           int64_t n, double diag) {
    const int64_t total = n * n;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % n, j = idx / n;
        W[i + j * ldw] = i == j ? diag : 0.0;
    }
}

// out(j, i) = in(i, j), 64 x 64 tiles through LDS (tiles_r tiles down the rows of `in`).
__global__ __launch_bounds__(256) void transpose_f64_kernel(const double* __restrict__ in,
                                                            int64_t ldi, int64_t n,
                                                            double* __restrict__ out, int64_t ldo,
                                                            int64_t tiles_r) {
    __shared__ double t[64][65];
    const int64_t r0 = (blockIdx.x % tiles_r) * 64, c0 = (blockIdx.x / tiles_r) * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int q = ty; q < 64; q += 4) {
        const int64_t i = r0 + tx, j = c0 + q;
        if (i < n && j < n) t[q][tx] = in[i + j * ldi];
    }
    __syncthreads();
    for (int q = ty; q < 64; q += 4) {
        const int64_t i = r0 + q, j = c0 + tx;
        if (i < n && j < n) out[j + i * ldo] = t[tx][q];
    }
}

// A(i, j) = A(j, i) = -W(i, j) for i >= j, tile by tile: tile (bi, bj), bi >= bj, reads W's
// lower tile and writes A's tile (bi, bj) and its mirror (bj, bi) through LDS.
__global__ __launch_bounds__(256) void negate_symmetrize_kernel(const double* __restrict__ W,
                                                                int64_t ldw, int64_t n,
                                                                double* __restrict__ A,
                                                                int64_t lda) {
    __shared__ double t[64][65];
    const int64_t bid = blockIdx.x;
    int64_t bj = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while (bj * (bj + 1) / 2 > bid) --bj;
    while ((bj + 1) * (bj + 2) / 2 <= bid) ++bj;
    const int64_t bi = bid - bj * (bj + 1) / 2;  // bi <= bj: tile rows bj, columns bi of W
    const int64_t r0 = bj * 64, c0 = bi * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int q = ty; q < 64; q += 4) {
        const int64_t i = r0 + tx, j = c0 + q;
        double v = 0.0;
        if (i < n && j < n) {
            // the diagonal tile's upper half mirrors its lower half
            v = (i >= j) ? -W[i + j * ldw] : -W[j + i * ldw];
            A[i + j * lda] = v;
        }
        t[q][tx] = v;
    }
    __syncthreads();
    if (bi == bj) return;
    for (int q = ty; q < 64; q += 4) {
        // mirror: A(c0 + tx, r0 + q) = value at (r0 + q, c0 + tx) = t[tx][q]
        const int64_t i = c0 + tx, j = r0 + q;
        if (i < n && j < n) A[i + j * lda] = t[tx][q];
    }
}

template int potrf_impl<double>(double*, int64_t, int64_t, int32_t*, hipStream_t, void*,
                                int64_t, PivotRule);
template int potrf_impl<float>(float*, int64_t, int64_t, int32_t*, hipStream_t, void*, int64_t,
                               PivotRule);

__global__ void set_f64_kernel(double* p, double v) { *p = v; }
template int potrs_impl<double>(const double*, int64_t, int64_t, double*, double*, int32_t*,
                                hipStream_t);
template int potrs_impl<float>(const float*, int64_t, int64_t, double*, double*, int32_t*,
                               hipStream_t);
template int trsv_lower_fwd<float>(const float*, int64_t, int64_t, double*, double*, int32_t*,
                                   hipStream_t);
template int trsv_lower_bwd<float>(const float*, int64_t, int64_t, double*, double*, int32_t*,
                                   hipStream_t);
template int trsv_lower_fwd<double>(const double*, int64_t, int64_t, double*, double*,
                                    int32_t*, hipStream_t);
template int trsv_lower_bwd<double>(const double*, int64_t, int64_t, double*, double*,
                                    int32_t*, hipStream_t);

}  // namespace qmlh

extern "C" int qmlh_potrf(double* A, int64_t n, int64_t lda, int32_t* info,
                          qmlh_stream_t stream) {
    return qmlh::potrf_impl<double>(A, n, lda, info, (hipStream_t)stream);
}

extern "C" int qmlh_spotrf(float* A, int64_t n, int64_t lda, int32_t* info,
                           qmlh_stream_t stream) {
    return qmlh::potrf_impl<float>(A, n, lda, info, (hipStream_t)stream);
}

// The split of one block column (region 0); from the look-ahead order on also the split of a
// group of two (region 1).
extern "C" size_t qmlh_spotrf_f16x3_workspace(int64_t n) {
    using namespace qmlh;
    if (n <= 0) return 0;
    const size_t base = syrk_f16_workspace(n, PF_NB_SPLIT);
    return n >= PF_LOOKAHEAD_MIN ? base + f16x3_region_bytes(n, PF_GROUP) : base;
}

extern "C" int qmlh_spotrf_f16x3(float* A, int64_t n, int64_t lda, int32_t* info, void* ws,
                                  size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n > 0 && (!ws || ws_bytes < qmlh_spotrf_f16x3_workspace(n)))
        return set_error(QMLH_ERR_WORKSPACE, "spotrf_f16x3: workspace too small");
    if (reinterpret_cast<uintptr_t>(ws) & 15)
        return set_error(QMLH_ERR_ARG, "spotrf_f16x3: workspace must be 16-byte aligned");
    return potrf_impl<float>(A, n, lda, info, (hipStream_t)stream, ws);
}

static int spotrf_lowrank_impl(float* A, int64_t n, int64_t lda, float* V, int64_t ldv,
                               int64_t r, int32_t f16x3, double pivot_floor,
                               double breakdown_pivot, double breakdown_shift,
                               double* shift_dev, int32_t* info,
                               void* ws, size_t ws_bytes, hipStream_t s) {
    using namespace qmlh;
    if (n < 0 || r < 0 || lda < n + r || (r > 0 && ldv < n))
        return set_error(QMLH_ERR_ARG, "spotrf_lowrank: bad shape (lda >= n + r, ldv >= n)");
    if (n > 0 && (!A || !info || (r > 0 && !V)))
        return set_error(QMLH_ERR_ARG, "spotrf_lowrank: null pointer");
    if (f16x3 && n > 0) {
        if (!ws || ws_bytes < qmlh_spotrf_f16x3_workspace(n))
            return set_error(QMLH_ERR_WORKSPACE, "spotrf_lowrank: workspace too small");
        if (reinterpret_cast<uintptr_t>(ws) & 15)
            return set_error(QMLH_ERR_ARG, "spotrf_lowrank: workspace must be 16-byte aligned");
    }
    if (!(pivot_floor >= 0.0)) return set_error(QMLH_ERR_ARG, "spotrf_lowrank: pivot_floor < 0");
    if (!(breakdown_shift >= 0.0) || !(breakdown_pivot >= 0.0) ||
        (breakdown_shift > 0.0 && !shift_dev))
        return set_error(QMLH_ERR_ARG,
                         "spotrf_lowrank: breakdown_shift or breakdown_pivot < 0, or no shift output");
    if (n > 0) QMLH_TRY(transpose_f32(V, ldv, n, r, A + n, lda, s));
    PivotRule pr;
    pr.floor = pivot_floor;
    if (breakdown_shift > 0.0) {
        pr.bshift = breakdown_shift;
        pr.shift_acc = shift_dev;
        pr.bpivot = breakdown_pivot;
        hipLaunchKernelGGL(set_f64_kernel, dim3(1), dim3(1), 0, s, shift_dev, 0.0);
        QMLH_CHECK_LAUNCH("spotrf_lowrank: clear shift");
    }
    QMLH_TRY(potrf_impl<float>(A, n, lda, info, s, f16x3 ? ws : nullptr, r, pr));
    if (n > 0) QMLH_TRY(transpose_f32(A + n, lda, r, n, V, ldv, s));
    return QMLH_OK;
}

extern "C" int qmlh_spotrf_lowrank(float* A, int64_t n, int64_t lda, float* V, int64_t ldv,
                                   int64_t r, int32_t f16x3, double pivot_floor, int32_t* info,
                                   void* ws, size_t ws_bytes, qmlh_stream_t stream) {
    return spotrf_lowrank_impl(A, n, lda, V, ldv, r, f16x3, pivot_floor, 0.0, 0.0, nullptr, info,
                               ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int qmlh_spotrf_lowrank_shift(float* A, int64_t n, int64_t lda, float* V,
                                         int64_t ldv, int64_t r, int32_t f16x3,
                                         double pivot_floor, double breakdown_pivot,
                                         double breakdown_shift, double* shift_dev,
                                         int32_t* info, void* ws, size_t ws_bytes,
                                         qmlh_stream_t stream) {
    return spotrf_lowrank_impl(A, n, lda, V, ldv, r, f16x3, pivot_floor, breakdown_pivot,
                               breakdown_shift, shift_dev, info, ws, ws_bytes,
                               (hipStream_t)stream);
}

// Workspace: an n-vector, from order TRSV_SYNCFREE_MIN the FP64 inverses of the factor's 64-
// and 128-row diagonal blocks, the solves' sync words.
extern "C" size_t qmlh_potrs_workspace(int64_t n) {
    using namespace qmlh;
    if (n <= 0) return 0;
    const size_t inv = n >= TRSV_SYNCFREE_MIN ? trsv_dinv_doubles(n) + trsv_dinv128_doubles(n) : 0;
    return ((size_t)n + inv) * sizeof(double) + trsv_sync_ints(n) * sizeof(int32_t);
}

extern "C" int qmlh_potrs(const double* L, int64_t n, int64_t lda, double* x, void* ws,
                          size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n > 0 && ws_bytes < qmlh_potrs_workspace(n))
        return set_error(QMLH_ERR_WORKSPACE, "potrs: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    double* tmp = static_cast<double*>(ws);
    if (n >= TRSV_SYNCFREE_MIN) {
        if (lda < n) return set_error(QMLH_ERR_ARG, "potrs: lda < n");
        if (!L || !x) return set_error(QMLH_ERR_ARG, "potrs: null pointer");
        // Both solves in one launch each, in 128-row blocks whose diagonal step is a product with
        // the FP64 inverse of the block (half the dependency chain of 64-row substitution).
        double* dinv = tmp + n;
        double* dinv128 = dinv + trsv_dinv_doubles(n);
        int32_t* sync = reinterpret_cast<int32_t*>(dinv128 + trsv_dinv128_doubles(n));
        TimerScope ts(QMLH_TIMER_POTRS, 2.0 * (double)n * (double)n, s);
        QMLH_TRY(trsv_inverse128<double>(L, n, lda, dinv, dinv128, s));
        QMLH_TRY(trsv_syncfree128<double>(L, n, lda, x, tmp, sync, true, s, dinv128));
        return trsv_syncfree128<double>(L, n, lda, tmp, x, sync, false, s, dinv128);
    }
    int32_t* sync = n > 0 ? reinterpret_cast<int32_t*>(tmp + n) : nullptr;
    return potrs_impl<double>(L, n, lda, x, tmp, sync, s);
}

extern "C" size_t qmlh_cho_invert_workspace(int64_t n) {
    if (n <= 0) return 0;
    return (size_t)(n * n + (int64_t)qmlh::trsv_dinv_doubles(n)) * sizeof(double);
}

extern "C" int qmlh_cho_invert(double* A, int64_t n, int64_t lda, int32_t* info, void* ws,
                               size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || lda < (n > 1 ? n : 1)) return set_error(QMLH_ERR_ARG, "cho_invert: n / lda");
    if (n == 0) return QMLH_OK;
    if (!A || !info) return set_error(QMLH_ERR_ARG, "cho_invert: null pointer");
    if (!ws || ws_bytes < qmlh_cho_invert_workspace(n))
        return set_error(QMLH_ERR_WORKSPACE, "cho_invert: workspace too small");
    if (reinterpret_cast<uintptr_t>(ws) & 15)
        return set_error(QMLH_ERR_ARG, "cho_invert: workspace must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    double* W = static_cast<double*>(ws);
    double* dinv = W + n * n;
    const int64_t ldw = n;
    QMLH_TRY(potrf_impl<double>(A, n, lda, info, s));
    const int64_t grid = ceil_div(n * n, 256) < 65536 ? ceil_div(n * n, 256) : 65536;
    hipLaunchKernelGGL(identity_f64_kernel, dim3((unsigned)grid), dim3(256), 0, s, W, ldw, n, 1.0);
    QMLH_CHECK_LAUNCH("cho_invert identity launch");
    QMLH_TRY(trsv_diag_inverse<double>(A, n, lda, dinv, s));
    QMLH_TRY(trsm_lower_fwd_multi<double>(A, n, lda, W, ldw, n, dinv, s, true));  // W = L^-1
    const int64_t tr = ceil_div(n, 64);
    hipLaunchKernelGGL(transpose_f64_kernel, dim3((unsigned)(tr * tr)), dim3(256), 0, s, W, ldw,
                       n, A, lda, tr);  // A = L^-T
    QMLH_CHECK_LAUNCH("cho_invert transpose launch");
    hipLaunchKernelGGL(identity_f64_kernel, dim3((unsigned)grid), dim3(256), 0, s, W, ldw, n, 0.0);
    QMLH_CHECK_LAUNCH("cho_invert clear launch");
    // W := 0 - A A^T = -(L^-T L^-1) on the lower triangle (FP64 MFMA) in one launch over
    // every lower tile: A = L^-T is upper triangular, so tile row bi only needs the products
    // over k >= its first row.
    QMLH_TRY((launch_update<double>(UPD_SYRK_UT, A, lda, A, lda, n, n, n, W, ldw, nullptr, s)));
    hipLaunchKernelGGL(negate_symmetrize_kernel, dim3((unsigned)(tr * (tr + 1) / 2)), dim3(256),
                       0, s, W, ldw, n, A, lda);
    QMLH_CHECK_LAUNCH("cho_invert symmetrize launch");
    return QMLH_OK;
}

extern "C" size_t qmlh_trsm_f32l_workspace(int64_t n) {
    if (n <= 0) return 0;
    return (qmlh::trsv_dinv_doubles(n) + qmlh::trsm_f32l_panel_doubles(n)) * sizeof(double);
}

extern "C" int qmlh_trsm_f32l(const float* L, int64_t n, int64_t ldl, double* V, int64_t ldv,
                              int64_t nrhs, void* ws, size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || nrhs < 0 || ldl < (n > 1 ? n : 1) || ldv < (n > 1 ? n : 1))
        return set_error(QMLH_ERR_ARG, "trsm_f32l: bad shape");
    if (n == 0 || nrhs == 0) return QMLH_OK;
    if (!L || !V || !ws) return set_error(QMLH_ERR_ARG, "trsm_f32l: null pointer");
    if (ws_bytes < qmlh_trsm_f32l_workspace(n))
        return set_error(QMLH_ERR_WORKSPACE, "trsm_f32l: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    double* dinv = static_cast<double*>(ws);
    double* panel = dinv + trsv_dinv_doubles(n);
    QMLH_TRY(trsv_diag_inverse<float>(L, n, ldl, dinv, s));
    TimerScope ts(QMLH_TIMER_GEMM, (double)n * (double)n * (double)nrhs, s);
    return trsm_lower_fwd_f32l(L, n, ldl, V, ldv, nrhs, dinv, panel, s);
}

// Y (n x r, FP64) := the hybrid Woodbury term: columns k .. r-1 widened from the factorisation's
// FP32 solve Y32 = L^-1 V, columns 0 .. k-1 solved again in FP64 arithmetic from the FP64
// right-hand sides Y holds there on entry (the leading columns of V, where u = V(:, 0) is the
// direction of the kernel's largest eigenvalues).
namespace qmlh {
__global__ __launch_bounds__(256) void widen_columns_kernel(const float* __restrict__ Y32,
                                                            int64_t ldy32, double* __restrict__ Y,
                                                            int64_t ldy, int64_t n, int64_t k0,
                                                            int64_t r) {
    const int64_t total = n * (r - k0);
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * 256) {
        const int64_t i = idx % n, j = k0 + idx / n;
        Y[i + j * ldy] = (double)Y32[i + j * ldy32];
    }
}
}  // namespace qmlh

// Workspace: the right-hand side copy (n doubles), the FP64 inverses of the factor's 64- and
// 128-row diagonal blocks (from order TRSV_SYNCFREE_MIN, for the 128-row single-launch solve),
// the solve's sync words.
extern "C" size_t qmlh_lowrank_hybrid_workspace(int64_t n) {
    using namespace qmlh;
    if (n <= 0) return 0;
    const size_t inv = n >= TRSV_SYNCFREE_MIN ? trsv_dinv_doubles(n) + trsv_dinv128_doubles(n) : 0;
    return ((size_t)n + inv) * sizeof(double) + trsv_sync_ints(n) * sizeof(int32_t);
}

extern "C" int qmlh_lowrank_hybrid(const float* L, int64_t n, int64_t ldl, const float* Y32,
                                   int64_t ldy32, double* Y, int64_t ldy, int64_t r, int64_t k,
                                   void* ws, size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (n < 0 || r < 0 || k < 0 || k > r || ldl < (n > 1 ? n : 1) || ldy < (n > 1 ? n : 1) ||
        (r > k && ldy32 < (n > 1 ? n : 1)))
        return set_error(QMLH_ERR_ARG, "lowrank_hybrid: bad shape");
    if (n == 0 || r == 0) return QMLH_OK;
    if (!L || !Y || (r > k && !Y32) || (k > 0 && !ws))
        return set_error(QMLH_ERR_ARG, "lowrank_hybrid: null pointer");
    if (k > 0 && ws_bytes < qmlh_lowrank_hybrid_workspace(n))
        return set_error(QMLH_ERR_WORKSPACE, "lowrank_hybrid: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    if (r > k) {
        const int64_t total = n * (r - k);
        const unsigned g = (unsigned)(ceil_div(total, (int64_t)256) < 8192 ? ceil_div(total, (int64_t)256) : 8192);
        hipLaunchKernelGGL(widen_columns_kernel, dim3(g), dim3(256), 0, s, Y32, ldy32, Y, ldy, n, k,
                           r);
        QMLH_CHECK_LAUNCH("lowrank_hybrid widen launch");
    }
    if (k == 0) return QMLH_OK;
    double* tmp = static_cast<double*>(ws);
    const bool big = n >= TRSV_SYNCFREE_MIN;
    double* dinv = tmp + n;
    double* dinv128 = dinv + (big ? trsv_dinv_doubles(n) : 0);
    int32_t* sync = reinterpret_cast<int32_t*>(dinv128 + (big ? trsv_dinv128_doubles(n) : 0));
    // Large orders: the single-launch solve in 128-row blocks, each block's diagonal step one
    // product with the FP64 inverse of its diagonal block (the preconditioner's own L^-1).
    if (big) QMLH_TRY(trsv_inverse128<float>(L, n, ldl, dinv, dinv128, s));
    for (int64_t j = 0; j < k; ++j) {
        QMLH_TRY(check_hip(hipMemcpyAsync(tmp, Y + j * ldy, (size_t)n * sizeof(double),
                                          hipMemcpyDeviceToDevice, s),
                           "lowrank_hybrid: copy"));
        if (big)
            QMLH_TRY(trsv_syncfree128<float>(L, n, ldl, tmp, Y + j * ldy, sync, true, s, dinv128));
        else
            QMLH_TRY(trsv_lower_fwd<float>(L, n, ldl, tmp, Y + j * ldy, sync, s));
    }
    return QMLH_OK;
}
