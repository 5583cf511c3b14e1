// Batched KRR fits of many uneven problems (learning curves, (sigma, lambda) grids) as one
// sequence of batched launches, optionally recorded into one hipGraph.
//
// Every problem p trains on rows x_row0 .. x_row0 + n_p - 1 of a shared representation pool
// X (row-major FP64, as the reference's kernels take it), with its own kernel kind, sigma and
// lambda, and solves (K_p + lambda_p I) alpha_p = y_p -- the plain FP64 path per problem
// (qml/kernels.py:33-87 kernel, qml/math.py:61-100 cho_solve), batched over problems:
//
//   build    one launch per kernel kind over the lower 128 x 128 tiles of every problem:
//            K_p(i, j), i >= j, plus lambda_p on the diagonal (FP64 MFMA cross term for the
//            Gaussian, FP64 VALU L1 tile for the Laplacian);
//   factor   right-looking blocked Cholesky, all problems in lock step over column blocks of
//            width nb (64, 128, 256 or 512; tuned). Per 64-wide sub-panel one launch factors the
//            diagonal blocks (one workgroup per problem, FP64 in one wavefront's registers),
//            one solves the panel rows below them (256 rows per workgroup), and an MFMA launch
//            updates the rest of the nb-wide panel; per block column one MFMA launch updates
//            the lower trailing tiles. With branches > 1 the problems are split into that many
//            groups of about equal n^3 work, each stepping in lock step on a stream of its own
//            (forked from and joined back to the caller's stream), so that the small,
//            latency-bound panel launches of one group run beside the updates of another;
//   solve    forward and backward substitutions of all problems in two single launches:
//            workgroups take 64-row blocks from an atomic ticket in level order (block l of
//            every problem that has one, then block l + 1, ...), wait (bounded) for the blocks
//            of their own problem they depend on, and publish with release flags.
//
// Each launch reads its jobs (one per active problem: pointers, sizes, first tile) from a
// table the planner uploads once; a workgroup finds its job by binary search over the jobs'
// first-tile indices. Problems are ordered largest first, so the big factorisations start at
// once and small ones fill the gaps. The launch sequence only depends on the problem sizes, nb
// and the branch count, so it is recorded once and replayed as one graph launch.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "chol_panel.h"
#include "mfma_core.h"
#include "pair_tile.h"

namespace qmlh {

static constexpr int BT = WG_T;   // update / build tile (128)
static constexpr int BW = CP_B;   // sub-panel width
static constexpr int BROWS = 256; // panel rows per workgroup
static constexpr int BS_B = 64;   // solve block
static constexpr unsigned BS_SPIN_LIMIT = 1u << 24;

struct BBuildJob {
    double* K;
    int64_t ld, n, x_row0, tiles_m, tile_begin;
    double scale, lambda;
};

struct BPanelJob {
    double* A;  // top-left of the w x w diagonal block; the panel rows follow below it
    int64_t ld, m, col_base, block_begin;
    int32_t* info;
    int32_t w, pad;
};

struct BUpdateJob {  // C(i, j) -= sum_k A(i, k) A(j, k), i >= j, column-major, ld
    const double* A;
    double* C;
    int64_t ld, M, N, Kd, tiles_m, tile_begin;
    int32_t* info;
    int32_t lower_tiles, pad;
};

struct BSolveProb {
    const double* L;
    const double* rhs;
    double* out;
    int32_t* flags;
    int64_t ld, n;
};

// Index of the last job whose first tile is <= bid (jobs sorted by first tile).
template <typename J>
__device__ __forceinline__ int find_job(const J* __restrict__ jobs, int njobs, int64_t bid,
                                        int64_t J::*first) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].*first <= bid) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void lower_tile(int64_t bid, int64_t& bi, int64_t& bj) {
    int64_t i = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > bid) --i;
    while ((i + 1) * (i + 2) / 2 <= bid) ++i;
    bi = i;
    bj = bid - i * (i + 1) / 2;
}

__global__ void bclear_kernel(int32_t* __restrict__ info, int64_t nprob,
                              int32_t* __restrict__ sync, int64_t nsync) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nprob) info[i] = 0;
    if (i < nsync) sync[i] = 0;
}

// ---- build ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void bbuild_gauss_kernel(const BBuildJob* __restrict__ jobs,
                                                              int njobs,
                                                              const double* __restrict__ X,
                                                              int64_t ldx, int64_t d,
                                                              const double* __restrict__ norms) {
    __shared__ __attribute__((aligned(16))) double lds[WG_LDS_ELEMS];
    const int jid = find_job(jobs, njobs, (int64_t)blockIdx.x, &BBuildJob::tile_begin);
    const BBuildJob jb = jobs[jid];
    int64_t bi, bj;
    lower_tile((int64_t)blockIdx.x - jb.tile_begin, bi, bj);
    const double* Xp = X + jb.x_row0 * ldx;
    const double* np = norms + jb.x_row0;
    Operand<double, LAYOUT_ROW> op{Xp, ldx, jb.n, d};
    WaveTile<double> t;
    wave_tile_zero(t);
    wg_tile_mma_lds(t, lds, op, op, bj * BT, bi * BT, d);  // first operand: columns j

    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t col0 = bj * BT + 64 * (w & 1);
    const int64_t row0 = bi * BT + 64 * (w >> 1);
    if (row0 >= jb.n || col0 >= jb.n || col0 > row0 + 63) return;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t i = row0 + 16 * b + (lane & 15);
            if (i >= jb.n) continue;
            const double ni = np[i];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t j = col0 + tile_row<double>(a, e);
                if (j > i) continue;
                double d2 = ni + np[j] - 2.0 * t.acc[a][b][e];
                d2 = d2 > 0.0 ? d2 : 0.0;
                if (i == j) d2 = 0.0;
                double v = exp(jb.scale * d2);
                if (i == j) v += jb.lambda;
                jb.K[i + j * jb.ld] = v;
            }
        }
    }
}

__global__ __launch_bounds__(PT_THREADS, 2) void bbuild_laplace_kernel(
    const BBuildJob* __restrict__ jobs, int njobs, const double* __restrict__ X, int64_t ldx,
    int64_t d) {
    __shared__ __attribute__((aligned(16))) double sA[PT_BK][PT_LDS];
    __shared__ __attribute__((aligned(16))) double sB[PT_BK][PT_LDS];
    const int jid = find_job(jobs, njobs, (int64_t)blockIdx.x, &BBuildJob::tile_begin);
    const BBuildJob jb = jobs[jid];
    int64_t bi, bj;
    lower_tile((int64_t)blockIdx.x - jb.tile_begin, bi, bj);
    const double* Xp = X + jb.x_row0 * ldx;
    const int64_t row0 = bi * PT_BM, col0 = bj * PT_BN;
    double acc[8][8];
    pair_tile_accumulate<OP_L1>(acc, sA, sB, Xp, ldx, jb.n, Xp, ldx, jb.n, d, row0, col0);
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t i = row0 + 2 * tx + 32 * (r / 2) + (r % 2);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t j = col0 + 2 * ty + 32 * (c / 2) + (c % 2);
            if (i < jb.n && j <= i) {
                double v = exp(jb.scale * acc[r][c]);
                if (i == j) v += jb.lambda;
                jb.K[i + j * jb.ld] = v;
            }
        }
    }
}

// ---- factor --------------------------------------------------------------------------------
// Per 64-wide sub-panel step over every active problem: one launch factors each job's diagonal
// block (one workgroup per job, chol_panel.h) and sets that problem's info; the next solves the
// panel rows below (256 rows per workgroup) against the finished factor.
__global__ __launch_bounds__(256) void bpotf2_kernel(const BPanelJob* __restrict__ jobs) {
    const BPanelJob jb = jobs[blockIdx.x];
    potf2_body<double>(jb.A, jb.ld, jb.w, jb.col_base, jb.info);
}

__global__ __launch_bounds__(256) void btrsm_kernel(const BPanelJob* __restrict__ jobs,
                                                    int njobs) {
    const int jid = find_job(jobs, njobs, (int64_t)blockIdx.x, &BPanelJob::block_begin);
    const BPanelJob jb = jobs[jid];
    trsm_body<double>(jb.A, jb.ld, jb.w, jb.m, jb.info, (int64_t)blockIdx.x - jb.block_begin);
}

__global__ __launch_bounds__(256, 2) void bupdate_kernel(const BUpdateJob* __restrict__ jobs,
                                                         int njobs) {
    __shared__ __attribute__((aligned(16))) double lds[WG_LDS_ELEMS];
    const int jid = find_job(jobs, njobs, (int64_t)blockIdx.x, &BUpdateJob::tile_begin);
    const BUpdateJob jb = jobs[jid];
    if (*jb.info) return;
    const int64_t tid = (int64_t)blockIdx.x - jb.tile_begin;
    int64_t bi, bj;
    if (jb.lower_tiles) {
        lower_tile(tid, bi, bj);
    } else {
        bi = tid % jb.tiles_m;
        bj = tid / jb.tiles_m;
    }
    if (bj * BT > bi * BT + BT - 1) return;  // tile entirely above the diagonal
    Operand<double, LAYOUT_COL> opA{jb.A, jb.ld, jb.M, jb.Kd};
    Operand<double, LAYOUT_COL> opB{jb.A, jb.ld, jb.N, jb.Kd};
    const int w = threadIdx.x >> 6;
    const int64_t col0 = bj * BT + 64 * (w & 1);
    const int64_t row0 = bi * BT + 64 * (w >> 1);
    const bool live = row0 < jb.M && col0 < jb.N && col0 <= row0 + 63;
    const bool interior = live && row0 + 64 <= jb.M && col0 + 64 <= jb.N && col0 + 63 <= row0;
    WaveTile<double> t;
    wave_tile_load_c<double, true>(t, jb.C, jb.ld, row0, col0, jb.M, jb.N, live, interior);
    wg_tile_mma_lds<double, LAYOUT_COL, LAYOUT_COL, true>(t, lds, opB, opA, bj * BT, bi * BT,
                                                          jb.Kd);
    if (!live) return;
    wave_tile_store_c<double, true>(t, jb.C, jb.ld, row0, col0, jb.M, jb.N, interior);
}

// ---- solve ---------------------------------------------------------------------------------
__device__ __forceinline__ bool bwait_flag(int32_t* flag, int32_t* err) {
    unsigned spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++spins > BS_SPIN_LIMIT) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// One 64-row block of one system (forward z = L^-1 r or backward x = L^-T z); the same scheme
// as the single-system solve: the diagonal block is staged first, the contributions of the
// finished blocks accumulate in a fixed order, one wavefront substitutes in FP64.
template <bool FWD>
__device__ __forceinline__ void bsolve_block(const BSolveProb& P, int32_t* err, int64_t b) {
    __shared__ double sL[BS_B][BS_B + 1];
    __shared__ double sInv[BS_B];
    __shared__ double red[BS_B][BS_B + 1];
    __shared__ double sv[BS_B];
    __shared__ int s_fail;
    const int t = threadIdx.x;
    const int64_t n = P.n, ld = P.ld;
    const int64_t nblocks = (n + BS_B - 1) / BS_B;
    if (t == 0) s_fail = 0;
    const int64_t r0 = b * BS_B;
    const int w = (int)(n - r0 < BS_B ? n - r0 : BS_B);
    const int lane = t & 63, grp = t >> 6;
    for (int idx = t; idx < BS_B * BS_B; idx += 256) {
        const int i = idx % BS_B, jj = idx / BS_B;
        if (i < w && jj < w && i >= jj) sL[i][jj] = P.L[(r0 + i) + (r0 + jj) * ld];
    }
    __syncthreads();
    double acc1 = 0.0;
    double acc16[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc16[c] = 0.0;
    bool ok = true;
    const int64_t nj = FWD ? b : nblocks - 1 - b;
    for (int64_t q = 0; q < nj; ++q) {
        const int64_t j = FWD ? q : nblocks - 1 - q;
        const int64_t c0 = j * BS_B;
        const int wj = (int)(n - c0 < BS_B ? n - c0 : BS_B);
        double lv[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int cc = 16 * grp + c;
            if (FWD)
                lv[c] = (lane < w && cc < wj) ? P.L[(r0 + lane) + (c0 + cc) * ld] : 0.0;
            else
                lv[c] = (lane < wj && cc < w) ? P.L[(c0 + lane) + (r0 + cc) * ld] : 0.0;
        }
        if (ok) ok = bwait_flag(&P.flags[j], err);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (FWD) {
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int cc = 16 * grp + c;
                const double zc = cc < wj ? __hip_atomic_load(&P.out[c0 + cc], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
                                          : 0.0;
                acc1 = fma(lv[c], zc, acc1);
            }
        } else {
            const double xj = lane < wj ? __hip_atomic_load(&P.out[c0 + lane], __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT)
                                        : 0.0;
#pragma unroll
            for (int c = 0; c < 16; ++c) acc16[c] = fma(lv[c], xj, acc16[c]);
        }
    }
    if (!ok) s_fail = 1;
    if (FWD) {
        red[lane][grp] = acc1;
    } else {
#pragma unroll
        for (int c = 0; c < 16; ++c) red[lane][16 * grp + c] = acc16[c];
    }
    __syncthreads();
    if (t < BS_B) {
        double s = 0.0;
        if (FWD) {
            s = ((red[t][0] + red[t][1]) + red[t][2]) + red[t][3];
        } else {
            for (int i = 0; i < BS_B; ++i) s += red[i][t];
        }
        sv[t] = (t < w ? P.rhs[r0 + t] : 0.0) - s;
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
            __hip_atomic_store(&P.out[r0 + i], s_fail ? __builtin_nan("") : xi, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (t == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&P.flags[b], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// sync: [0] ticket, [1] error. level_pre[l] = number of (problem, block) pairs in levels < l;
// level l holds block l (forward) / block nblocks - 1 - l (backward) of problems 0 .. count-1
// (problems are sorted by block count, descending).
template <bool FWD>
__global__ __launch_bounds__(256) void bsolve_kernel(const BSolveProb* __restrict__ probs,
                                                     const int64_t* __restrict__ level_pre,
                                                     int nlevels, int32_t* __restrict__ sync) {
    __shared__ int64_t s_ticket;
    if (threadIdx.x == 0) s_ticket = atomicAdd(&sync[0], 1);
    __syncthreads();
    const int64_t ticket = s_ticket;
    int lo = 0, hi = nlevels - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (level_pre[mid] <= ticket) lo = mid; else hi = mid - 1;
    }
    const int p = (int)(ticket - level_pre[lo]);
    const BSolveProb P = probs[p];
    const int64_t nblocks = (P.n + BS_B - 1) / BS_B;
    const int64_t b = FWD ? lo : nblocks - 1 - lo;
    bsolve_block<FWD>(P, &sync[1], b);
}

// ---- planner -------------------------------------------------------------------------------
enum { BL_CLEAR, BL_BUILD_GAUSS, BL_BUILD_LAPLACE, BL_POTF2, BL_TRSM, BL_UPDATE, BL_SOLVE_FWD,
       BL_SOLVE_BWD };

struct BLaunch {
    int kind;
    int64_t grid;
    size_t table_off;  // byte offset of the job table in the plan tables
    int njobs;
    int branch;  // factor launches: the problem group (stream) it belongs to; -1: main stream
};

static constexpr int MAX_BRANCHES = 4;

struct HostProblem {
    int64_t n, x_row0, y_off, a_off, ld, k_off, z_off, flag_off;
    double sigma, lambda;
    int32_t kind;
    int index;
};

struct PlanLayout {
    std::vector<HostProblem> probs;  // sorted by n, descending
    std::vector<BLaunch> launches;
    std::vector<unsigned char> tables;
    size_t level_off = 0;  // byte offset of the solve level prefix in the tables
    int branches = 1;      // problem groups factored on streams of their own
    size_t k_elems = 0, z_elems = 0, sync_ints = 0;
    size_t ws_k = 0, ws_z = 0, ws_sync = 0, ws_tables = 0, ws_total = 0;
};

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

template <typename J>
static size_t push_table(std::vector<unsigned char>& tab, const std::vector<J>& v) {
    const size_t off = align_up(tab.size(), 64);
    tab.resize(off + v.size() * sizeof(J));
    if (!v.empty()) std::memcpy(tab.data() + off, v.data(), v.size() * sizeof(J));
    return off;
}

// Device base addresses the job tables point into (all zero when only sizing the plan: the
// tables are then built from integer addresses and never uploaded).
struct Bases {
    uintptr_t K, z, sync, info, y, alpha;
};

template <typename T>
static T* at(uintptr_t base, int64_t elems) {
    return reinterpret_cast<T*>(base + (uintptr_t)elems * sizeof(T));
}

static int build_plan(const qmlh_problem* P, int32_t nprob, int32_t nb, int32_t branches,
                      const Bases* B, PlanLayout& L) {
    L = PlanLayout();
    for (int i = 0; i < nprob; ++i) {
        HostProblem h;
        h.n = P[i].n;
        h.x_row0 = P[i].x_row0;
        h.y_off = P[i].y_offset;
        h.a_off = P[i].alpha_offset;
        h.sigma = P[i].sigma;
        h.lambda = P[i].lambda;
        h.kind = P[i].kind;
        h.index = i;
        L.probs.push_back(h);
    }
    std::stable_sort(L.probs.begin(), L.probs.end(),
                     [](const HostProblem& a, const HostProblem& b) { return a.n > b.n; });
    size_t koff = 0, zoff = 0, foff = 4;  // sync: [0..1] forward ticket/err, [2..3] backward
    for (auto& h : L.probs) {
        h.ld = (int64_t)align_up((size_t)h.n, 16);
        h.k_off = (int64_t)koff;
        koff = align_up(koff + (size_t)h.ld * (size_t)h.n, 64);
        h.z_off = (int64_t)zoff;
        zoff = align_up(zoff + (size_t)h.n, 64);
        h.flag_off = (int64_t)foff;
        foff += 2 * (size_t)((h.n + BS_B - 1) / BS_B);  // forward flags then backward flags
    }
    L.k_elems = koff;
    L.z_elems = zoff;
    L.sync_ints = foff;
    const Bases b0{0, 0, 0, 0, 0, 0};
    const Bases& bs = B ? *B : b0;
    auto Kp = [&](const HostProblem& h, int64_t r, int64_t c) -> double* {
        return at<double>(bs.K, h.k_off + r + c * h.ld);
    };
    auto infop = [&](const HostProblem& h) -> int32_t* { return at<int32_t>(bs.info, h.index); };

    // clear
    L.launches.push_back({BL_CLEAR, (int64_t)ceil_div((int64_t)std::max<size_t>(nprob, foff), 256),
                          0, 0, -1});
    // build, one launch per kind present
    for (int kind : {QMLH_PAIR_GAUSSIAN, QMLH_PAIR_LAPLACIAN}) {
        std::vector<BBuildJob> jobs;
        int64_t tiles = 0;
        for (const auto& h : L.probs) {
            if (h.kind != kind) continue;
            const int64_t tm = ceil_div(h.n, (int64_t)BT);
            BBuildJob j;
            j.K = Kp(h, 0, 0);
            j.ld = h.ld;
            j.n = h.n;
            j.x_row0 = h.x_row0;
            j.tiles_m = tm;
            j.tile_begin = tiles;
            j.scale = kind == QMLH_PAIR_GAUSSIAN ? -0.5 / (h.sigma * h.sigma) : -1.0 / h.sigma;
            j.lambda = h.lambda;
            tiles += tm * (tm + 1) / 2;
            jobs.push_back(j);
        }
        if (jobs.empty()) continue;
        L.launches.push_back({kind == QMLH_PAIR_GAUSSIAN ? BL_BUILD_GAUSS : BL_BUILD_LAPLACE,
                              tiles, push_table(L.tables, jobs), (int)jobs.size(), -1});
    }
    // factor: the lock-step launch sequence of one group of problems (indices into L.probs,
    // largest first)
    auto factor_group = [&](const std::vector<int>& idx, int branch, std::vector<BLaunch>& out) {
        const int64_t nmax_g = idx.empty() ? 0 : L.probs[idx[0]].n;
        for (int64_t k = 0; k < nmax_g; k += nb) {
            for (int64_t jj = 0; jj < nb && k + jj < nmax_g; jj += BW) {
                std::vector<BPanelJob> pj;
                std::vector<BUpdateJob> uj;
                int64_t blocks = 0, tiles = 0;
                for (int i : idx) {
                    const HostProblem& h = L.probs[i];
                    const int64_t c = k + jj;
                    if (h.n <= c) continue;
                    const int64_t kb = std::min<int64_t>(nb, h.n - k);
                    if (jj >= kb) continue;
                    const int w = (int)std::min<int64_t>(BW, kb - jj);
                    BPanelJob p;
                    p.A = Kp(h, c, c);
                    p.ld = h.ld;
                    p.m = h.n - c - w;
                    p.col_base = c;
                    p.block_begin = blocks;
                    p.info = infop(h);
                    p.w = w;
                    p.pad = 0;
                    blocks += ceil_div(p.m, (int64_t)BROWS);
                    pj.push_back(p);
                    const int64_t ncols = kb - jj - w;  // rest of this nb-wide panel
                    if (ncols > 0) {
                        BUpdateJob u;
                        u.A = Kp(h, c + w, c);
                        u.C = Kp(h, c + w, c + w);
                        u.ld = h.ld;
                        u.M = h.n - c - w;
                        u.N = ncols;
                        u.Kd = w;
                        u.tiles_m = ceil_div(u.M, (int64_t)BT);
                        u.tile_begin = tiles;
                        u.info = infop(h);
                        u.lower_tiles = 0;
                        u.pad = 0;
                        tiles += u.tiles_m * ceil_div(u.N, (int64_t)BT);
                        uj.push_back(u);
                    }
                }
                if (!pj.empty()) {
                    const size_t off = push_table(L.tables, pj);
                    out.push_back({BL_POTF2, (int64_t)pj.size(), off, (int)pj.size(), branch});
                    // jobs with m = 0 own no row blocks: find_job skips them (same first block)
                    if (blocks > 0) out.push_back({BL_TRSM, blocks, off, (int)pj.size(), branch});
                }
                if (!uj.empty())
                    out.push_back({BL_UPDATE, tiles, push_table(L.tables, uj), (int)uj.size(),
                                   branch});
            }
            std::vector<BUpdateJob> tj;
            int64_t tiles = 0;
            for (int i : idx) {
                const HostProblem& h = L.probs[i];
                if (h.n <= k) continue;
                const int64_t kb = std::min<int64_t>(nb, h.n - k);
                const int64_t m = h.n - k - kb;
                if (m <= 0) continue;
                BUpdateJob u;
                u.A = Kp(h, k + kb, k);
                u.C = Kp(h, k + kb, k + kb);
                u.ld = h.ld;
                u.M = m;
                u.N = m;
                u.Kd = kb;
                u.tiles_m = ceil_div(m, (int64_t)BT);
                u.tile_begin = tiles;
                u.info = infop(h);
                u.lower_tiles = 1;
                u.pad = 0;
                tiles += u.tiles_m * (u.tiles_m + 1) / 2;
                tj.push_back(u);
            }
            if (!tj.empty())
                out.push_back({BL_UPDATE, tiles, push_table(L.tables, tj), (int)tj.size(), branch});
        }
    };
    // Problem groups of about equal Cholesky work (n^3): each problem, largest first, joins the
    // least loaded group. Each group's sequence runs on a stream of its own, so the
    // latency-bound panel launches of one group overlap the updates of another.
    const int G = std::max(1, std::min<int>(branches, (int)L.probs.size()));
    L.branches = G;
    std::vector<std::vector<int>> groups(G);
    std::vector<double> load(G, 0.0);
    for (int i = 0; i < (int)L.probs.size(); ++i) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        groups[g].push_back(i);
        const double n = (double)L.probs[i].n;
        load[g] += n * n * n;
    }
    std::vector<std::vector<BLaunch>> seq(G);
    for (int g = 0; g < G; ++g) factor_group(groups[g], g, seq[g]);
    // interleaved, so that the host enqueues the groups evenly
    for (size_t i = 0;; ++i) {
        bool any = false;
        for (int g = 0; g < G; ++g) {
            if (i < seq[g].size()) {
                L.launches.push_back(seq[g][i]);
                any = true;
            }
        }
        if (!any) break;
    }
    const int64_t nmax = L.probs.empty() ? 0 : L.probs[0].n;
    // solves: problem table (forward, backward) and the level prefix
    {
        std::vector<BSolveProb> fw, bw;
        for (const auto& h : L.probs) {
            BSolveProb f;
            f.L = Kp(h, 0, 0);
            f.ld = h.ld;
            f.n = h.n;
            f.rhs = at<const double>(bs.y, h.y_off);
            f.out = at<double>(bs.z, h.z_off);
            f.flags = at<int32_t>(bs.sync, h.flag_off);
            fw.push_back(f);
            BSolveProb g = f;
            g.rhs = f.out;
            g.out = at<double>(bs.alpha, h.a_off);
            g.flags = f.flags + (h.n + BS_B - 1) / BS_B;
            bw.push_back(g);
        }
        const int64_t nlev = (nmax + BS_B - 1) / BS_B;
        std::vector<int64_t> pre;
        int64_t acc = 0;
        for (int64_t l = 0; l < nlev; ++l) {
            pre.push_back(acc);
            for (const auto& h : L.probs)
                if ((h.n + BS_B - 1) / BS_B > l) ++acc;
        }
        pre.push_back(acc);
        if (nlev > 0) {
            L.level_off = push_table(L.tables, pre);
            const size_t fw_off = push_table(L.tables, fw);
            const size_t bw_off = push_table(L.tables, bw);
            // For the solves, njobs carries the number of levels.
            L.launches.push_back({BL_SOLVE_FWD, acc, fw_off, (int)nlev, -1});
            L.launches.push_back({BL_SOLVE_BWD, acc, bw_off, (int)nlev, -1});
        }
    }
    L.ws_k = 0;
    L.ws_z = align_up(L.ws_k + L.k_elems * sizeof(double), 256);
    L.ws_sync = align_up(L.ws_z + L.z_elems * sizeof(double), 256);
    L.ws_tables = align_up(L.ws_sync + L.sync_ints * sizeof(int32_t), 256);
    L.ws_total = align_up(L.ws_tables + L.tables.size(), 256);
    return QMLH_OK;
}

}  // namespace qmlh

struct qmlh_batched_plan {
    qmlh::PlanLayout layout;
    unsigned char* ws = nullptr;
    const double* X = nullptr;
    int64_t ldx = 0, d = 0;
    const double* norms = nullptr;
    int32_t* info = nullptr;
    int32_t nprob = 0;
    // streams of problem groups 1 .. branches - 1 (group 0 runs on the caller's stream), the
    // fork event and one join event per side stream; all at the default priority
    int nside = 0;
    hipStream_t side[qmlh::MAX_BRANCHES - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[qmlh::MAX_BRANCHES - 1] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void release_streams(qmlh_batched_plan* p) {
    for (int i = 0; i < qmlh::MAX_BRANCHES - 1; ++i) {
        if (p->join[i]) (void)hipEventDestroy(p->join[i]);
        if (p->side[i]) (void)hipStreamDestroy(p->side[i]);
        p->join[i] = nullptr;
        p->side[i] = nullptr;
    }
    if (p->fork) (void)hipEventDestroy(p->fork);
    p->fork = nullptr;
    p->nside = 0;
}

static int validate(const qmlh_problem* probs, int32_t nprob, int32_t nb, int32_t branches) {
    using namespace qmlh;
    if (nprob < 0) return set_error(QMLH_ERR_ARG, "batched: nprob < 0");
    if (nprob > 0 && !probs) return set_error(QMLH_ERR_ARG, "batched: problems are null");
    if (nb != 64 && nb != 128 && nb != 256 && nb != 512)
        return set_error(QMLH_ERR_ARG, "batched: nb must be 64, 128, 256 or 512 (got %d)", nb);
    if (branches < 1 || branches > MAX_BRANCHES)
        return set_error(QMLH_ERR_ARG, "batched: branches must be 1 .. %d (got %d)", MAX_BRANCHES,
                         branches);
    std::vector<std::pair<int64_t, int64_t>> out;
    for (int i = 0; i < nprob; ++i) {
        const qmlh_problem& p = probs[i];
        if (p.n < 1) return set_error(QMLH_ERR_ARG, "batched: problem %d has n < 1", i);
        if (p.x_row0 < 0 || p.y_offset < 0 || p.alpha_offset < 0)
            return set_error(QMLH_ERR_ARG, "batched: problem %d has a negative offset", i);
        if (p.kind != QMLH_PAIR_GAUSSIAN && p.kind != QMLH_PAIR_LAPLACIAN)
            return set_error(QMLH_ERR_ARG, "batched: problem %d kind must be Gaussian or Laplacian", i);
        if (!(p.sigma > 0.0)) return set_error(QMLH_ERR_ARG, "batched: problem %d sigma <= 0", i);
        out.push_back({p.alpha_offset, p.alpha_offset + p.n});
    }
    std::sort(out.begin(), out.end());
    for (size_t i = 1; i < out.size(); ++i)
        if (out[i].first < out[i - 1].second)
            return set_error(QMLH_ERR_ARG, "batched: alpha ranges of two problems overlap");
    return QMLH_OK;
}

extern "C" size_t qmlh_krr_batched_workspace(const qmlh_problem* probs, int32_t nprob,
                                             int32_t nb, int32_t branches) {
    using namespace qmlh;
    if (validate(probs, nprob, nb, branches) != QMLH_OK) return 0;
    PlanLayout L;
    build_plan(probs, nprob, nb, branches, nullptr, L);
    return L.ws_total;
}

extern "C" int qmlh_krr_batched_create(const qmlh_problem* probs, int32_t nprob, int32_t nb,
                                       int32_t branches, const double* X, int64_t x_rows,
                                       int64_t ldx, int64_t d, const double* norms,
                                       const double* y, int64_t y_len, double* alpha,
                                       int64_t alpha_len, int32_t* info, void* ws,
                                       size_t ws_bytes, qmlh_batched_plan** plan) {
    using namespace qmlh;
    if (!plan) return set_error(QMLH_ERR_ARG, "batched: plan is null");
    *plan = nullptr;
    QMLH_TRY(validate(probs, nprob, nb, branches));
    if (d < 0 || ldx < d) return set_error(QMLH_ERR_ARG, "batched: ldx < d");
    for (int i = 0; i < nprob; ++i) {
        const qmlh_problem& q = probs[i];
        if (q.x_row0 + q.n > x_rows)
            return set_error(QMLH_ERR_ARG, "batched: problem %d reads past the pool rows", i);
        if (q.y_offset + q.n > y_len)
            return set_error(QMLH_ERR_ARG, "batched: problem %d reads past the labels", i);
        if (q.alpha_offset + q.n > alpha_len)
            return set_error(QMLH_ERR_ARG, "batched: problem %d writes past alpha", i);
    }
    if (nprob > 0 && (!X || !y || !alpha || !info || !ws))
        return set_error(QMLH_ERR_ARG, "batched: null pointer");
    bool gauss = false;
    for (int i = 0; i < nprob; ++i) gauss = gauss || probs[i].kind == QMLH_PAIR_GAUSSIAN;
    if (gauss && !norms) return set_error(QMLH_ERR_ARG, "batched: Gaussian problems need norms");
    if ((reinterpret_cast<uintptr_t>(ws) & 255) != 0)
        return set_error(QMLH_ERR_ARG, "batched: workspace must be 256-byte aligned");
    PlanLayout probe;
    build_plan(probs, nprob, nb, branches, nullptr, probe);
    if (ws_bytes < probe.ws_total)
        return set_error(QMLH_ERR_ARG, "batched: workspace too small (%zu < %zu bytes)", ws_bytes,
                         probe.ws_total);
    qmlh_batched_plan* p = new (std::nothrow) qmlh_batched_plan();
    if (!p) return set_error(QMLH_ERR_HIP, "batched: out of host memory");
    unsigned char* w = static_cast<unsigned char*>(ws);
    const uintptr_t wb = reinterpret_cast<uintptr_t>(w);
    Bases b{wb + probe.ws_k, wb + probe.ws_z, wb + probe.ws_sync,
            reinterpret_cast<uintptr_t>(info), reinterpret_cast<uintptr_t>(y),
            reinterpret_cast<uintptr_t>(alpha)};
    build_plan(probs, nprob, nb, branches, &b, p->layout);
    p->ws = w;
    p->X = X;
    p->ldx = ldx;
    p->d = d;
    p->norms = norms;
    p->info = info;
    p->nprob = nprob;
    int rc = QMLH_OK;
    if (p->layout.branches > 1) {
        rc = check_hip(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming), "batched: event");
        for (int i = 0; rc == QMLH_OK && i < p->layout.branches - 1; ++i) {
            rc = check_hip(hipStreamCreateWithFlags(&p->side[i], hipStreamNonBlocking),
                           "batched: branch stream");
            if (rc == QMLH_OK)
                rc = check_hip(hipEventCreateWithFlags(&p->join[i], hipEventDisableTiming),
                               "batched: event");
            if (rc == QMLH_OK) p->nside = i + 1;
        }
    }
    if (rc == QMLH_OK && !p->layout.tables.empty())
        rc = check_hip(hipMemcpy(w + p->layout.ws_tables, p->layout.tables.data(),
                                 p->layout.tables.size(), hipMemcpyHostToDevice),
                       "batched: table upload");
    if (rc != QMLH_OK) {
        release_streams(p);
        delete p;
        return rc;
    }
    *plan = p;
    return QMLH_OK;
}

static int launch_one(const qmlh_batched_plan* plan, const qmlh::BLaunch& l, hipStream_t s) {
    using namespace qmlh;
    const PlanLayout& L = plan->layout;
    unsigned char* tab = plan->ws + L.ws_tables;
    int32_t* sync = reinterpret_cast<int32_t*>(plan->ws + L.ws_sync);
    const int64_t* levels = reinterpret_cast<const int64_t*>(tab + L.level_off);
    const dim3 grid((unsigned)l.grid), block(256);
    switch (l.kind) {
        case BL_CLEAR:
            hipLaunchKernelGGL(bclear_kernel, grid, block, 0, s, plan->info, (int64_t)plan->nprob,
                               sync, (int64_t)L.sync_ints);
            break;
        case BL_BUILD_GAUSS:
            hipLaunchKernelGGL(bbuild_gauss_kernel, grid, block, 0, s,
                               reinterpret_cast<const BBuildJob*>(tab + l.table_off), l.njobs,
                               plan->X, plan->ldx, plan->d, plan->norms);
            break;
        case BL_BUILD_LAPLACE:
            hipLaunchKernelGGL(bbuild_laplace_kernel, grid, block, 0, s,
                               reinterpret_cast<const BBuildJob*>(tab + l.table_off), l.njobs,
                               plan->X, plan->ldx, plan->d);
            break;
        case BL_POTF2:
            hipLaunchKernelGGL(bpotf2_kernel, grid, block, 0, s,
                               reinterpret_cast<const BPanelJob*>(tab + l.table_off));
            break;
        case BL_TRSM:
            hipLaunchKernelGGL(btrsm_kernel, grid, block, 0, s,
                               reinterpret_cast<const BPanelJob*>(tab + l.table_off), l.njobs);
            break;
        case BL_UPDATE:
            hipLaunchKernelGGL(bupdate_kernel, grid, block, 0, s,
                               reinterpret_cast<const BUpdateJob*>(tab + l.table_off), l.njobs);
            break;
        case BL_SOLVE_FWD:
            hipLaunchKernelGGL(bsolve_kernel<true>, grid, block, 0, s,
                               reinterpret_cast<const BSolveProb*>(tab + l.table_off), levels,
                               l.njobs, sync);
            break;
        case BL_SOLVE_BWD:
            hipLaunchKernelGGL(bsolve_kernel<false>, grid, block, 0, s,
                               reinterpret_cast<const BSolveProb*>(tab + l.table_off), levels,
                               l.njobs, sync + 2);
            break;
    }
    return check_hip(hipGetLastError(), "batched launch");
}

extern "C" int qmlh_krr_batched_run(qmlh_batched_plan* plan, qmlh_stream_t stream) {
    using namespace qmlh;
    if (!plan) return set_error(QMLH_ERR_ARG, "batched: plan is null");
    hipStream_t s = (hipStream_t)stream;
    // Factor launches of group g > 0 go to side stream g - 1, after a fork from the caller's
    // stream; the first main-stream launch after them (the solves) waits for every side stream
    // (a join), so the whole fit stays ordered on the caller's stream, also under capture.
    bool forked = false;
    auto fork = [&]() -> int {
        QMLH_TRY(check_hip(hipEventRecord(plan->fork, s), "batched: fork"));
        for (int i = 0; i < plan->nside; ++i)
            QMLH_TRY(check_hip(hipStreamWaitEvent(plan->side[i], plan->fork, 0), "batched: fork"));
        forked = true;
        return QMLH_OK;
    };
    auto join = [&]() -> int {
        int rc = QMLH_OK;
        for (int i = 0; i < plan->nside; ++i) {
            int r = check_hip(hipEventRecord(plan->join[i], plan->side[i]), "batched: join");
            if (r == QMLH_OK) r = check_hip(hipStreamWaitEvent(s, plan->join[i], 0), "batched: join");
            if (rc == QMLH_OK) rc = r;
        }
        forked = false;
        return rc;
    };
    int rc = QMLH_OK;
    for (const BLaunch& l : plan->layout.launches) {
        if (plan->nside > 0 && l.branch >= 0 && !forked) rc = fork();
        if (rc == QMLH_OK && forked && l.branch < 0) rc = join();
        if (rc != QMLH_OK) break;
        hipStream_t ls = l.branch > 0 ? plan->side[l.branch - 1] : s;
        rc = launch_one(plan, l, ls);
        if (rc != QMLH_OK) break;
    }
    if (forked) {  // also after a failed launch: a capture must not end with unjoined streams
        const int r = join();
        if (rc == QMLH_OK) rc = r;
    }
    return rc;
}

extern "C" int qmlh_krr_batched_capture(qmlh_batched_plan* plan, qmlh_stream_t stream) {
    using namespace qmlh;
    if (!plan) return set_error(QMLH_ERR_ARG, "batched: plan is null");
    if (!stream) return set_error(QMLH_ERR_ARG, "batched: capture needs a non-default stream");
    hipStream_t s = (hipStream_t)stream;
    if (plan->exec) {
        (void)hipGraphExecDestroy(plan->exec);
        plan->exec = nullptr;
    }
    if (plan->graph) {
        (void)hipGraphDestroy(plan->graph);
        plan->graph = nullptr;
    }
    QMLH_TRY(check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
                       "batched: begin capture"));
    const int rc = qmlh_krr_batched_run(plan, stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != QMLH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    QMLH_TRY(check_hip(e, "batched: end capture"));
    plan->graph = g;
    QMLH_TRY(check_hip(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0),
                       "batched: instantiate graph"));
    return QMLH_OK;
}

extern "C" int qmlh_krr_batched_launch(qmlh_batched_plan* plan, qmlh_stream_t stream) {
    using namespace qmlh;
    if (!plan || !plan->exec) return set_error(QMLH_ERR_ARG, "batched: no captured graph");
    return check_hip(hipGraphLaunch(plan->exec, (hipStream_t)stream), "batched: graph launch");
}

extern "C" int64_t qmlh_krr_batched_launches(const qmlh_batched_plan* plan) {
    return plan ? (int64_t)plan->layout.launches.size() : -1;
}

extern "C" int qmlh_krr_batched_factor(const qmlh_batched_plan* plan, int32_t p,
                                       int64_t* offset, int64_t* ld) {
    using namespace qmlh;
    if (!plan || p < 0 || p >= plan->nprob || !offset || !ld)
        return set_error(QMLH_ERR_ARG, "batched: bad factor query");
    for (const auto& h : plan->layout.probs) {
        if (h.index == p) {
            *offset = (int64_t)(plan->layout.ws_k / sizeof(double)) + h.k_off;
            *ld = h.ld;
            return QMLH_OK;
        }
    }
    return set_error(QMLH_ERR_ARG, "batched: problem not found");
}

extern "C" void qmlh_krr_batched_destroy(qmlh_batched_plan* plan) {
    if (!plan) return;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    release_streams(plan);
    delete plan;
}
