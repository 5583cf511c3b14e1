// Pairwise kernels and distances on the FP64 VALU: Laplacian / Manhattan (sum |a-b|) and
// Gaussian / L2 (sum (a-b)^2) with direct differences, as qml/fkernels.f90:340-400 and
// qml/fdistance.f90:23-80 compute them; the Sargan and Matern kernels
// (qml/fkernels.f90:427-531) as epilogues of the same L1 / squared sums; p-distances
// (qml/fdistance.f90:83-155) with |a-b|^p per feature. The 128 x 128 tile (pair_tile.h) keeps each entry's
// feature sum in feature order, so an entry equals the reference's serial sum. Per
// pair-feature the L1 form costs two FP64 VALU ops (v_add_f64 for a - b, v_add_f64 with the
// |.| source modifier), the squared form one sub and one FMA.
#include "pair_tile.h"
#include "timer.h"

namespace qmlh {

template <int OP, int EPI>
__global__ __launch_bounds__(PT_THREADS, 2) void pair_tile_kernel(
    const double* __restrict__ A, int64_t lda, int64_t na, const double* __restrict__ B,
    int64_t ldb, int64_t nb, int64_t d, PairEpi ep, double p, double* __restrict__ K,
    int64_t ldk, int symmetric, int64_t tiles_m, int lower) {
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
        if (lower) {  // the transposed tile, on or below the diagonal, written as is
            bi = j;
            bj = bid - j * (j + 1) / 2;
        }
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
    pair_tile_accumulate<OP>(acc, sA, sB, A, lda, na, B, ldb, nb, d, row0, col0, p);

    const bool mirror = symmetric && !lower && (bi != bj);
    const bool diag_lower = symmetric && lower && bi == bj;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t i = row0 + 2 * tx + 32 * (r / 2) + (r % 2);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t j = col0 + 2 * ty + 32 * (c / 2) + (c % 2);
            if (i < na && j < nb && !(diag_lower && i < j)) {
                double v = pair_epilogue<EPI>(acc[r][c], ep);
                K[i + j * ldk] = v;
                if (mirror) K[j + i * ldk] = v;
            }
        }
    }
}

template <int OP, int EPI>
static int launch_pair(const double* A, int64_t na, int64_t lda, const double* B, int64_t nb,
                       int64_t ldb, int64_t d, const PairEpi& ep, double* K, int64_t ldk,
                       bool symmetric, hipStream_t s, double p = 0.0, bool lower = false) {
    int64_t tm = ceil_div(na, PT_BM), tn = ceil_div(nb, PT_BN);
    int64_t ntiles = symmetric ? tm * (tm + 1) / 2 : tm * tn;
    if (ntiles > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "pairwise: grid too large");
    double work = (double)d * (symmetric ? (double)na * (double)(na + 1) / 2.0
                                         : (double)na * (double)nb);
    TimerScope ts(QMLH_TIMER_PAIR, work, s);
    hipLaunchKernelGGL((pair_tile_kernel<OP, EPI>), dim3((unsigned)ntiles), dim3(PT_THREADS), 0,
                       s, A, lda, na, B, ldb, nb, d, ep, p, K, ldk, symmetric ? 1 : 0, tm,
                       symmetric && lower ? 1 : 0);
    QMLH_CHECK_LAUNCH("pairwise launch");
    return QMLH_OK;
}

static PairEpi pair_epi(double s0, double s1 = 0.0) {
    PairEpi e{};
    e.s0 = s0;
    e.s1 = s1;
    return e;
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
    const bool lower = (flags & QMLH_LOWER) != 0;
    if (sym && (na != nb)) return set_error(QMLH_ERR_ARG, "pairwise: symmetric needs na == nb");
    if (lower && !sym) return set_error(QMLH_ERR_ARG, "pairwise: QMLH_LOWER needs QMLH_SYMMETRIC");
    if (na == 0 || nb == 0) return QMLH_OK;
    if (!A || !B || !K) return set_error(QMLH_ERR_ARG, "pairwise: null pointer");
    hipStream_t s = (hipStream_t)stream;
    switch (kind) {
        case QMLH_PAIR_LAPLACIAN:
            return launch_pair<OP_L1, EPI_EXP>(A, na, lda, B, nb, ldb, d, pair_epi(-1.0 / sigma),
                                               K, ldk, sym, s, 0.0, lower);
        case QMLH_PAIR_GAUSSIAN:
            return launch_pair<OP_SQ, EPI_EXP>(A, na, lda, B, nb, ldb, d,
                                               pair_epi(-0.5 / (sigma * sigma)), K, ldk, sym, s,
                                               0.0, lower);
        case QMLH_PAIR_MANHATTAN:
            return launch_pair<OP_L1, EPI_IDENTITY>(A, na, lda, B, nb, ldb, d, pair_epi(0.0), K,
                                                    ldk, sym, s, 0.0, lower);
        case QMLH_PAIR_L2:
            return launch_pair<OP_SQ, EPI_SQRT>(A, na, lda, B, nb, ldb, d, pair_epi(0.0), K, ldk,
                                                sym, s, 0.0, lower);
        default:
            return set_error(QMLH_ERR_ARG, "pairwise: unknown kind %d", kind);
    }
}

extern "C" int qmlh_pairwise_ext(int32_t kind, const double* A, int64_t na, int64_t lda,
                                 const double* B, int64_t nb, int64_t ldb, int64_t d,
                                 const double* params_h, int32_t nparams, double* K, int64_t ldk,
                                 int32_t flags, qmlh_stream_t stream) {
    using namespace qmlh;
    if (na < 0 || nb < 0 || d < 0) return set_error(QMLH_ERR_ARG, "pairwise_ext: negative size");
    if (lda < d || ldb < d) return set_error(QMLH_ERR_ARG, "pairwise_ext: ld < d");
    if (ldk < na) return set_error(QMLH_ERR_ARG, "pairwise_ext: ldk < na");
    const bool sym = (flags & QMLH_SYMMETRIC) != 0;
    const bool lower = (flags & QMLH_LOWER) != 0;
    if (sym && (na != nb)) return set_error(QMLH_ERR_ARG, "pairwise_ext: symmetric needs na == nb");
    if (lower && !sym)
        return set_error(QMLH_ERR_ARG, "pairwise_ext: QMLH_LOWER needs QMLH_SYMMETRIC");
    if (nparams < 1 || !params_h) return set_error(QMLH_ERR_ARG, "pairwise_ext: parameters");
    if (na == 0 || nb == 0) return QMLH_OK;
    if (!A || !B || !K) return set_error(QMLH_ERR_ARG, "pairwise_ext: null pointer");
    hipStream_t s = (hipStream_t)stream;
    const double sigma = params_h[0];
    switch (kind) {
        case QMLH_PAIR_SARGAN: {
            const int ng = nparams - 1;
            if (ng > PAIR_MAX_GAMMAS)
                return set_error(QMLH_ERR_ARG, "pairwise_ext: at most %d gammas", PAIR_MAX_GAMMAS);
            PairEpi e = pair_epi(-1.0 / sigma);
            e.ng = ng;
            for (int m = 0; m < ng; ++m) e.g[m] = params_h[1 + m];
            return launch_pair<OP_L1, EPI_SARGAN>(A, na, lda, B, nb, ldb, d, e, K, ldk, sym, s,
                                                  0.0, lower);
        }
        case QMLH_PAIR_MATERN_L2: {
            if (nparams < 2) return set_error(QMLH_ERR_ARG, "pairwise_ext: Matern needs the order");
            const int order = (int)params_h[1];
            if (order < 0 || order > 2 || (double)order != params_h[1])
                return set_error(QMLH_ERR_ARG, "pairwise_ext: Matern order %g", params_h[1]);
            // qml/fkernels.f90:440-477: inv = -1/sigma, -sqrt(3)/sigma, -sqrt(5)/sigma.
            PairEpi e = pair_epi(order == 0 ? -1.0 / sigma
                                 : order == 1 ? -sqrt(3.0) / sigma : -sqrt(5.0) / sigma,
                                 order == 2 ? 5.0 / (3.0 * sigma * sigma) : 0.0);
            e.order = order;
            return launch_pair<OP_SQ, EPI_MATERN>(A, na, lda, B, nb, ldb, d, e, K, ldk, sym, s,
                                                  0.0, lower);
        }
        case QMLH_PAIR_P_INT: {
            // (sum |a-b|^p)^(1/p), 1/p = 1 / dble(p) (qml/fdistance.f90:139). p = 1 and p = 3, 4
            // use exact products; other integer powers use pow, within a few ulp of the
            // reference's repeated squaring.
            const double p = params_h[0];
            if (!(p >= 1.0) || p != (double)(int)p || p > 1e6)
                return set_error(QMLH_ERR_ARG, "pairwise_ext: integer p %g", p);
            const int ip = (int)p;
            const PairEpi e = pair_epi(1.0 / (double)ip);
            if (ip == 1)
                return launch_pair<OP_L1, EPI_IDENTITY>(A, na, lda, B, nb, ldb, d, e, K, ldk,
                                                        sym, s, 0.0, lower);
            if (ip == 3)
                return launch_pair<OP_POW3, EPI_ROOT>(A, na, lda, B, nb, ldb, d, e, K, ldk, sym,
                                                      s, 0.0, lower);
            if (ip == 4)
                return launch_pair<OP_POW4, EPI_ROOT>(A, na, lda, B, nb, ldb, d, e, K, ldk, sym,
                                                      s, 0.0, lower);
            return launch_pair<OP_POWR, EPI_ROOT>(A, na, lda, B, nb, ldb, d, e, K, ldk, sym, s, p,
                                                  lower);
        }
        case QMLH_PAIR_P_REAL: {
            const double p = params_h[0];
            if (!(p > 0.0)) return set_error(QMLH_ERR_ARG, "pairwise_ext: p %g <= 0", p);
            return launch_pair<OP_POWR, EPI_ROOT>(A, na, lda, B, nb, ldb, d, pair_epi(1.0 / p), K,
                                                  ldk, sym, s, p, lower);
        }
        default:
            return set_error(QMLH_ERR_ARG, "pairwise_ext: unknown kind %d", kind);
    }
}
