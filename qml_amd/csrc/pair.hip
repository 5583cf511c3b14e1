// Pairwise kernels and distances on the FP64 VALU: Laplacian / Manhattan (sum |a-b|) and
// Gaussian / L2 (sum (a-b)^2) with direct differences, as qml/fkernels.f90:340-400 and
// qml/fdistance.f90:23-80 compute them.
//
// Tiling: a 256-lane workgroup owns a 128 x 128 output tile; each lane owns an 8 x 8 register
// tile (rows 2*tx + 32*r + {0,1}, columns 2*ty + 32*c + {0,1}, tx = t % 16, ty = t / 16).
// Features stream through LDS in chunks of 16, stored k-major so a lane fetches its two
// adjacent rows (or columns) for one feature with a single 16-byte LDS read. Each lane keeps
// the feature sum of every entry in feature order k = 0, 1, ..., d-1, so an entry equals the
// reference's serial sum. Per pair-feature the L1 form costs two FP64 VALU ops (v_add_f64 for
// a - b, v_add_f64 with the |.| source modifier), the squared form one sub and one FMA.
#include "common.h"
#include "timer.h"

namespace qmlh {

static constexpr int PT_BM = 128;
static constexpr int PT_BN = 128;
static constexpr int PT_BK = 16;
static constexpr int PT_LDS = PT_BM + 2;  // padded k-major row (doubles), keeps 16 B alignment
static constexpr int PT_THREADS = 256;

enum { OP_L1 = 0, OP_SQ = 1 };
enum { EPI_IDENTITY = 0, EPI_SQRT = 1, EPI_EXP = 2 };

template <int OP>
__device__ __forceinline__ double pair_term(double acc, double a, double b) {
    if (OP == OP_L1) {
        return acc + fabs(a - b);
    } else {
        double t = a - b;
        return fma(t, t, acc);
    }
}

template <int EPI>
__device__ __forceinline__ double pair_epilogue(double acc, double scale) {
    if (EPI == EPI_EXP) return exp(scale * acc);
    if (EPI == EPI_SQRT) return sqrt(acc);
    return acc;
}

// Load a 128-row x 16-feature chunk of X (row-major, ld) into LDS as s[k][row].
// Rows >= n and features >= d read as zero; zero features add exactly zero to either sum.
__device__ __forceinline__ void pair_load_chunk(double (*s)[PT_LDS], const double* __restrict__ X,
                                                int64_t ld, int64_t n, int64_t d, int64_t row0,
                                                int64_t k0, int t) {
    const int k = t % PT_BK;
    const int rbase = t / PT_BK;  // 0..15
    const int64_t kk = k0 + k;
#pragma unroll
    for (int q = 0; q < PT_BM / 16; ++q) {
        int r = rbase + 16 * q;
        int64_t row = row0 + r;
        double v = 0.0;
        if (row < n && kk < d) v = X[row * ld + kk];
        s[k][r] = v;
    }
}

template <int OP, int EPI>
__global__ __launch_bounds__(PT_THREADS, 2) void pair_tile_kernel(
    const double* __restrict__ A, int64_t lda, int64_t na, const double* __restrict__ B,
    int64_t ldb, int64_t nb, int64_t d, double scale, double* __restrict__ K, int64_t ldk,
    int symmetric, int64_t tiles_m) {
    __shared__ __attribute__((aligned(16))) double sA[PT_BK][PT_LDS];
    __shared__ __attribute__((aligned(16))) double sB[PT_BK][PT_LDS];

    int64_t bi, bj;
    const int64_t bid = blockIdx.x;
    if (symmetric) {
        // Upper-triangular tile enumeration: bid -> (bi <= bj).
        int64_t j = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
        while (j * (j + 1) / 2 > bid) --j;
        while ((j + 1) * (j + 2) / 2 <= bid) ++j;
        bj = j;
        bi = bid - j * (j + 1) / 2;
    } else {
        bi = bid % tiles_m;
        bj = bid / tiles_m;
    }
    const int64_t row0 = bi * PT_BM;
    const int64_t col0 = bj * PT_BN;

    const int t = threadIdx.x;
    const int tx = t % 16;
    const int ty = t / 16;

    double acc[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = 0.0;

    for (int64_t k0 = 0; k0 < d; k0 += PT_BK) {
        pair_load_chunk(sA, A, lda, na, d, row0, k0, t);
        pair_load_chunk(sB, B, ldb, nb, d, col0, k0, t);
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < PT_BK; ++k) {
            double a[8], b[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double2 va = *reinterpret_cast<const double2*>(&sA[k][2 * tx + 32 * r]);
                a[2 * r] = va.x;
                a[2 * r + 1] = va.y;
                const double2 vb = *reinterpret_cast<const double2*>(&sB[k][2 * ty + 32 * r]);
                b[2 * r] = vb.x;
                b[2 * r + 1] = vb.y;
            }
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[r][c] = pair_term<OP>(acc[r][c], a[r], b[c]);
        }
        __syncthreads();
    }

    const bool mirror = symmetric && (bi != bj);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t i = row0 + 2 * tx + 32 * (r / 2) + (r % 2);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t j = col0 + 2 * ty + 32 * (c / 2) + (c % 2);
            if (i < na && j < nb) {
                double v = pair_epilogue<EPI>(acc[r][c], scale);
                K[i + j * ldk] = v;
                if (mirror) K[j + i * ldk] = v;
            }
        }
    }
}

template <int OP, int EPI>
static int launch_pair(const double* A, int64_t na, int64_t lda, const double* B, int64_t nb,
                       int64_t ldb, int64_t d, double scale, double* K, int64_t ldk,
                       bool symmetric, hipStream_t s) {
    int64_t tm = ceil_div(na, PT_BM), tn = ceil_div(nb, PT_BN);
    int64_t ntiles = symmetric ? tm * (tm + 1) / 2 : tm * tn;
    if (ntiles > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "pairwise: grid too large");
    double work = (double)d * (symmetric ? (double)na * (double)(na + 1) / 2.0
                                         : (double)na * (double)nb);
    TimerScope ts(QMLH_TIMER_PAIR, work, s);
    hipLaunchKernelGGL((pair_tile_kernel<OP, EPI>), dim3((unsigned)ntiles), dim3(PT_THREADS), 0,
                       s, A, lda, na, B, ldb, nb, d, scale, K, ldk, symmetric ? 1 : 0, tm);
    QMLH_CHECK_LAUNCH("pairwise launch");
    return QMLH_OK;
}

}  // namespace qmlh

extern "C" int qmlh_pairwise(int32_t kind, const double* A, int64_t na, int64_t lda,
                             const double* B, int64_t nb, int64_t ldb, int64_t d, double sigma,
                             double* K, int64_t ldk, int32_t flags, qmlh_stream_t stream) {
    using namespace qmlh;
    if (na < 0 || nb < 0 || d < 0) return set_error(QMLH_ERR_ARG, "pairwise: negative size");
    if (lda < d || ldb < d) return set_error(QMLH_ERR_ARG, "pairwise: ld < d");
    if (ldk < na) return set_error(QMLH_ERR_ARG, "pairwise: ldk < na");
    bool sym = (flags & QMLH_SYMMETRIC) != 0;
    if (sym && (na != nb)) return set_error(QMLH_ERR_ARG, "pairwise: symmetric needs na == nb");
    if (na == 0 || nb == 0) return QMLH_OK;
    if (!A || !B || !K) return set_error(QMLH_ERR_ARG, "pairwise: null pointer");
    hipStream_t s = (hipStream_t)stream;
    switch (kind) {
        case QMLH_PAIR_LAPLACIAN:
            return launch_pair<OP_L1, EPI_EXP>(A, na, lda, B, nb, ldb, d, -1.0 / sigma, K, ldk,
                                               sym, s);
        case QMLH_PAIR_GAUSSIAN:
            return launch_pair<OP_SQ, EPI_EXP>(A, na, lda, B, nb, ldb, d,
                                               -0.5 / (sigma * sigma), K, ldk, sym, s);
        case QMLH_PAIR_MANHATTAN:
            return launch_pair<OP_L1, EPI_IDENTITY>(A, na, lda, B, nb, ldb, d, 0.0, K, ldk, sym,
                                                    s);
        case QMLH_PAIR_L2:
            return launch_pair<OP_SQ, EPI_SQRT>(A, na, lda, B, nb, ldb, d, 0.0, K, ldk, sym, s);
        default:
            return set_error(QMLH_ERR_ARG, "pairwise: unknown kind %d", kind);
    }
}
