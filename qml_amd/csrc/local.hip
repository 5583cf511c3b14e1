// Local (atom-summed) kernels between molecules:
//   K_s(a, b) = sum over atoms i of molecule a, j of molecule b of exp(inv_s * D(x_i, y_j))
// the quantity of fget_local_kernels_gaussian / fget_local_kernels_laplacian
// (qml/fkernels.f90:23-196) and of fget_vector_kernels_* (qml/fkernels.f90:198-337) behind the
// atomic-kernel wrappers, for several sigmas from one distance tile.
//
// Molecules are grouped into panels of consecutive molecules holding at most LK_TILE atoms and
// LK_MOLS molecules (a molecule larger than LK_TILE atoms is a panel of its own), so that every
// molecule pair belongs to exactly one panel pair and each output entry is written once, by
// one workgroup, without atomics. A workgroup computes the atom-pair distance tile of its panel
// pair (128 x 128 atoms: FP64 MFMA expanded form for the Gaussian, the FP64 VALU direct tile
// for the Laplacian and the direct Gaussian), applies exp for each sigma, and reduces the tile
// to molecule-pair sums with two small MFMA products against 0/1 membership matrices:
//   T = V * Sb^T (atom rows x panel molecules of B),  S = Sa * T (molecules of A x of B),
// where Sa(m, i) = [atom i of the tile belongs to the panel's molecule m]. Padding atoms (past
// the panel's end) have no molecule, so ragged panels need no masks. Products with 0 and 1 are
// exact; the sums run in the matrix cores' fixed order, so results are deterministic.
// The four waves' partial sums meet in LDS in wave order (one barrier per wave), then the
// workgroup writes its molecule pairs. For a symmetric call (the same molecules on both sides)
// only panel pairs P <= Q run and every value is mirrored, so K_s is exactly symmetric.
#include "mfma_core.h"
#include "pair_tile.h"
#include "timer.h"

#include <vector>

namespace qmlh {

static constexpr int LK_TILE = 128;
static constexpr int LK_MOLS = QMLH_LOCAL_PANEL_MOLS;  // molecules per panel (two blocks of 16)
static constexpr int LK_MB = LK_MOLS / 16;              // molecule blocks per panel side

// Workgroups per launch; larger panel-pair counts are covered by a grid-stride loop.
static constexpr int64_t LK_MAX_GRID = 1 << 20;

enum { LK_GAUSS_MFMA = 0, LK_GAUSS_VALU = 1, LK_LAPLACE = 2 };

struct LkSide {
    const double* X;
    int64_t ld;
    const double* norms;     // squared row norms (MFMA Gaussian only)
    const int64_t* offsets;  // (nmol + 1) atom offsets
    const int64_t* panels;   // (npanels + 1) first molecule of each panel
    int64_t atoms;           // total atoms (rows of X)
};

struct LkSigmas {
    double inv[4];  // exp argument scale per sigma of this launch
    int n;          // sigmas in this launch (<= NS)
    int64_t first;  // index of inv[0] in the caller's list
};

// Panel-local molecule index of tile atom `t` (row0 + t), -1 past the panel's last atom.
__device__ __forceinline__ int lk_local_mol(const int64_t* __restrict__ off, int64_t m0,
                                            int nm, int64_t atom, int64_t atom_end) {
    if (atom >= atom_end) return -1;
    int lo = 0;
    for (int m = 1; m < nm; ++m)
        if (off[m0 + m] <= atom) lo = m;
    return lo;
}

// Reduce this wave's values to molecule-pair sums S (added to S).
// Lane layout shared by both tile kinds: value val(r, c) belongs to tile row rowof(r, lane & 15)
// and tile column colof(c, lane >> 4); rmol[r][e] is the panel molecule of tile row
// rowof(r, (lane >> 4) + 4e), cmol[c] that of tile column colof(c, lane >> 4) (-1: padding).
// Stage 1, per row set r: T(m, n) = sum_c val(r, c) [cmol = 16 nb + n] (16 rows x 16
// molecules); stage 2: S(ma, n) += sum_e [rmol[r][e] = 16 mb + ma] T((lane >> 4) + 4e, n).
template <int NR, int NC, typename Val>
__device__ __forceinline__ void lk_reduce(Val val, const int (&rmol)[NR][4], const int (&cmol)[NC],
                                          int nblk_a, int nblk_b,
                                          double4_t (&S)[LK_MB][LK_MB]) {
    typedef MfmaTraits<double> M;
    const int me = threadIdx.x & 15;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double vr[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) vr[c] = val(r, c);
#pragma unroll
        for (int nb = 0; nb < LK_MB; ++nb) {
            if (nb >= nblk_b) break;
            double4_t T = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const double ind = cmol[c] == 16 * nb + me ? 1.0 : 0.0;
                T = M::mma(vr[c], ind, T);
            }
#pragma unroll
            for (int mb = 0; mb < LK_MB; ++mb) {
                if (mb >= nblk_a) break;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const double ind = rmol[r][e] == 16 * mb + me ? 1.0 : 0.0;
                    S[mb][nb] = M::mma(ind, T[e], S[mb][nb]);
                }
            }
        }
    }
}

// acc[ma][n] += S in wave order (one barrier per wave); lane element e holds
// S(16 mb + (lane >> 4) + 4e, 16 nb + (lane & 15)).
__device__ __forceinline__ void lk_accumulate(double* __restrict__ acc,
                                              const double4_t (&S)[LK_MB][LK_MB], int nblk_a,
                                              int nblk_b) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    for (int w2 = 0; w2 < 4; ++w2) {
        if (w == w2) {
#pragma unroll
            for (int mb = 0; mb < LK_MB; ++mb)
#pragma unroll
                for (int nb = 0; nb < LK_MB; ++nb) {
                    if (mb >= nblk_a || nb >= nblk_b) continue;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int ma = 16 * mb + (lane >> 4) + 4 * e;
                        const int n = 16 * nb + (lane & 15);
                        acc[ma * LK_MOLS + n] += S[mb][nb][e];
                    }
                }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void lk_zero(double4_t (&S)[LK_MB][LK_MB]) {
#pragma unroll
    for (int x = 0; x < LK_MB; ++x)
#pragma unroll
        for (int y = 0; y < LK_MB; ++y) S[x][y] = double4_t{0.0, 0.0, 0.0, 0.0};
}

template <int KIND, int NS>
__global__ __launch_bounds__(256, 1) void local_kernels_kernel(LkSide A, LkSide B, int64_t d,
                                                               LkSigmas sg, double* __restrict__ K,
                                                               int64_t ldk, int64_t sstride,
                                                               int symmetric, int64_t tiles_a,
                                                               int64_t ntiles) {
    __shared__ __attribute__((aligned(16))) double lds[KIND == LK_GAUSS_MFMA ? WG_LDS_ELEMS
                                                                              : 2 * PT_BK * PT_LDS];
    __shared__ double acc[NS][LK_MOLS * LK_MOLS];
    __shared__ int smolA[LK_TILE], smolB[LK_TILE];

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;

    for (int64_t bid = blockIdx.x; bid < ntiles; bid += gridDim.x) {
        int64_t bi, bj;
        if (symmetric) {
            int64_t j = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
            while (j * (j + 1) / 2 > bid) --j;
            while ((j + 1) * (j + 2) / 2 <= bid) ++j;
            bj = j;
            bi = bid - j * (j + 1) / 2;
        } else {
            bi = bid % tiles_a;
            bj = bid / tiles_a;
        }
        const int64_t ma0 = A.panels[bi], mb0 = B.panels[bj];
        const int nma = (int)(A.panels[bi + 1] - ma0), nmb = (int)(B.panels[bj + 1] - mb0);
        const int64_t ra0 = A.offsets[ma0], ra1 = A.offsets[ma0 + nma];
        const int64_t rb0 = B.offsets[mb0], rb1 = B.offsets[mb0 + nmb];
        const int nblk_a = (nma + 15) / 16, nblk_b = (nmb + 15) / 16;
        // Panels come from qmlh_local_panels; one with more molecules than the LDS sums hold
        // is skipped rather than overrunning them.
        if (nma > LK_MOLS || nmb > LK_MOLS || nma < 0 || nmb < 0) continue;

        for (int q = t; q < NS * LK_MOLS * LK_MOLS; q += 256) (&acc[0][0])[q] = 0.0;

        for (int64_t r0 = ra0; r0 < ra1; r0 += LK_TILE) {
            for (int64_t c0 = rb0; c0 < rb1; c0 += LK_TILE) {
                if (t < LK_TILE) smolA[t] = lk_local_mol(A.offsets, ma0, nma, r0 + t, ra1);
                else smolB[t - LK_TILE] = lk_local_mol(B.offsets, mb0, nmb, c0 + t - LK_TILE, rb1);
                __syncthreads();
                if (KIND == LK_GAUSS_MFMA) {
                    Operand<double, LAYOUT_ROW> opA{A.X, A.ld, A.atoms, d};
                    Operand<double, LAYOUT_ROW> opB{B.X, B.ld, B.atoms, d};
                    WaveTile<double> tile;
                    wave_tile_zero(tile);
                    wg_tile_mma_lds(tile, lds, opB, opA, c0, r0, d);
                    // Lane value (a, b, e): A row r0 + 64 (w >> 1) + 16 b + (lane & 15),
                    // B column c0 + 64 (w & 1) + 16 a + (lane >> 4) + 4 e.
                    const int rbase = 64 * (w >> 1), cbase = 64 * (w & 1);
                    int rmol[4][4], cmol[16];
                    double ni[4], nj[16];
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int64_t i = r0 + rbase + 16 * b + (lane & 15);
                        ni[b] = A.norms[i < A.atoms ? i : A.atoms - 1];
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            rmol[b][e] = smolA[rbase + 16 * b + (lane >> 4) + 4 * e];
                    }
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int cl = cbase + 16 * a + (lane >> 4) + 4 * e;
                            const int64_t j = c0 + cl;
                            nj[4 * a + e] = B.norms[j < B.atoms ? j : B.atoms - 1];
                            cmol[4 * a + e] = smolB[cl];
                        }
                    double d2[4][16];
#pragma unroll
                    for (int b = 0; b < 4; ++b)
#pragma unroll
                        for (int a = 0; a < 4; ++a)
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                const double x = ni[b] + nj[4 * a + e] - 2.0 * tile.acc[a][b][e];
                                d2[b][4 * a + e] = x > 0.0 ? x : 0.0;
                            }
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        if (s >= sg.n) break;
                        const double inv = sg.inv[s];
                        double4_t S[LK_MB][LK_MB];
                        lk_zero(S);
                        lk_reduce<4, 16>([&](int r, int c) { return exp(inv * d2[r][c]); }, rmol,
                                         cmol, nblk_a, nblk_b, S);
                        lk_accumulate(acc[s], S, nblk_a, nblk_b);
                    }
                } else {
                    double (*sA)[PT_LDS] = reinterpret_cast<double (*)[PT_LDS]>(lds);
                    double (*sB)[PT_LDS] = reinterpret_cast<double (*)[PT_LDS]>(lds + PT_BK * PT_LDS);
                    double dist[8][8];
                    // Rows r0 + 2 tx + 32 (r / 2) + r % 2, columns c0 + 2 ty + 32 (c / 2) + c % 2
                    // with tx = lane & 15, ty = 4 w + (lane >> 4).
                    if (KIND == LK_LAPLACE)
                        pair_tile_accumulate<OP_L1>(dist, sA, sB, A.X, A.ld, A.atoms, B.X, B.ld,
                                                    B.atoms, d, r0, c0);
                    else
                        pair_tile_accumulate<OP_SQ>(dist, sA, sB, A.X, A.ld, A.atoms, B.X, B.ld,
                                                    B.atoms, d, r0, c0);
                    int rmol[8][4], cmol[8];
#pragma unroll
                    for (int r = 0; r < 8; ++r)
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            rmol[r][e] = smolA[2 * ((lane >> 4) + 4 * e) + 32 * (r / 2) + r % 2];
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        cmol[c] = smolB[2 * (4 * w + (lane >> 4)) + 32 * (c / 2) + c % 2];
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        if (s >= sg.n) break;
                        const double inv = sg.inv[s];
                        double4_t S[LK_MB][LK_MB];
                        lk_zero(S);
                        lk_reduce<8, 8>([&](int r, int c) { return exp(dist[r][c] * inv); }, rmol,
                                        cmol, nblk_a, nblk_b, S);
                        lk_accumulate(acc[s], S, nblk_a, nblk_b);
                    }
                }
            }
        }
        __syncthreads();

        // Write the panel pair's molecule sums; mirror for a symmetric call.
        const bool diag = symmetric && bi == bj;
        for (int s = 0; s < NS; ++s) {
            if (s >= sg.n) break;
            double* Ks = K + (sg.first + s) * sstride;
            for (int q = t; q < LK_MOLS * LK_MOLS; q += 256) {
                const int ma = q % LK_MOLS, mb = q / LK_MOLS;
                if (ma >= nma || mb >= nmb) continue;
                if (diag && ma > mb) continue;
                const double val = acc[s][ma * LK_MOLS + mb];
                const int64_t ga = ma0 + ma, gb = mb0 + mb;
                Ks[ga + gb * ldk] = val;
                if (symmetric && ga != gb) Ks[gb + ga * ldk] = val;
            }
        }
        __syncthreads();
    }
}

template <int KIND, int NS>
static void lk_launch(int64_t ntiles, hipStream_t s, const LkSide& A, const LkSide& B,
                      int64_t d, const LkSigmas& sg, double* K, int64_t ldk, int64_t sstride,
                      int symmetric, int64_t tiles_a) {
    const int64_t grid = ntiles < LK_MAX_GRID ? ntiles : LK_MAX_GRID;
    hipLaunchKernelGGL((local_kernels_kernel<KIND, NS>), dim3((unsigned)grid), dim3(256), 0, s, A,
                       B, d, sg, K, ldk, sstride, symmetric, tiles_a, ntiles);
}

}  // namespace qmlh

extern "C" int64_t qmlh_local_panels(const int64_t* offsets_h, int64_t nmol, int64_t* panels_h) {
    using namespace qmlh;
    if (nmol < 0 || !offsets_h || !panels_h)
        return set_error(QMLH_ERR_ARG, "local_panels: bad arguments");
    int64_t np = 0;
    panels_h[0] = 0;
    int64_t start = 0;
    while (start < nmol) {
        int64_t m = start;
        const int64_t a0 = offsets_h[start];
        if (offsets_h[start + 1] < a0) return set_error(QMLH_ERR_ARG, "local_panels: offsets decrease");
        // Take molecules while the panel stays within LK_TILE atoms and LK_MOLS molecules; a
        // first molecule larger than LK_TILE atoms is a panel alone.
        ++m;
        if (offsets_h[m] - a0 <= LK_TILE) {
            while (m < nmol && m - start < LK_MOLS && offsets_h[m + 1] - a0 <= LK_TILE) {
                if (offsets_h[m + 1] < offsets_h[m])
                    return set_error(QMLH_ERR_ARG, "local_panels: offsets decrease");
                ++m;
            }
        }
        panels_h[++np] = m;
        start = m;
    }
    return np;
}

extern "C" int qmlh_local_kernels(int32_t kind, const double* A, int64_t atomsA, int64_t lda,
                                  const double* normsA, const int64_t* offsA,
                                  const int64_t* panelsA, int64_t npA, const double* B,
                                  int64_t atomsB, int64_t ldb, const double* normsB,
                                  const int64_t* offsB, const int64_t* panelsB, int64_t npB,
                                  int64_t d, const double* sigmas_h, int32_t nsig, double* K,
                                  int64_t ldk, int64_t sstride, int32_t flags,
                                  qmlh_stream_t stream) {
    using namespace qmlh;
    if (kind != QMLH_PAIR_GAUSSIAN && kind != QMLH_PAIR_LAPLACIAN)
        return set_error(QMLH_ERR_ARG, "local_kernels: kind %d", kind);
    if (npA < 0 || npB < 0 || atomsA < 0 || atomsB < 0 || d < 0 || nsig < 0)
        return set_error(QMLH_ERR_ARG, "local_kernels: negative size");
    if (lda < d || ldb < d) return set_error(QMLH_ERR_ARG, "local_kernels: ld < d");
    if (npA == 0 || npB == 0 || nsig == 0) return QMLH_OK;
    if (!offsA || !offsB || !panelsA || !panelsB || !K || !sigmas_h)
        return set_error(QMLH_ERR_ARG, "local_kernels: null pointer");
    if ((atomsA > 0 && !A) || (atomsB > 0 && !B))
        return set_error(QMLH_ERR_ARG, "local_kernels: null representation");
    const bool symmetric = (flags & QMLH_SYMMETRIC) != 0;
    if (symmetric && (A != B || offsA != offsB || panelsA != panelsB || npA != npB))
        return set_error(QMLH_ERR_ARG, "local_kernels: symmetric call needs identical sides");
    const bool mfma = kind == QMLH_PAIR_GAUSSIAN && normsA && normsB;
    if (kind == QMLH_PAIR_GAUSSIAN && (normsA != nullptr) != (normsB != nullptr))
        return set_error(QMLH_ERR_ARG, "local_kernels: give both norms or neither");
    for (int32_t s = 0; s < nsig; ++s)
        if (!(sigmas_h[s] > 0.0)) return set_error(QMLH_ERR_ARG, "local_kernels: sigma <= 0");
    LkSide sa{A, lda, normsA, offsA, panelsA, atomsA};
    LkSide sb{B, ldb, normsB, offsB, panelsB, atomsB};
    const int64_t grid = symmetric ? npA * (npA + 1) / 2 : npA * npB;
    hipStream_t s = (hipStream_t)stream;
    // Work: atom pairs x features (the pairs of a symmetric call counted once).
    const double pairs = symmetric ? 0.5 * (double)atomsA * (double)(atomsA + 1)
                                   : (double)atomsA * (double)atomsB;
    TimerScope ts(QMLH_TIMER_LOCAL, pairs * (double)d, s);
    for (int32_t s0 = 0; s0 < nsig; s0 += 4) {
        LkSigmas sg;
        sg.n = nsig - s0 < 4 ? (int)(nsig - s0) : 4;
        sg.first = s0;
        for (int q = 0; q < 4; ++q) {
            const double sig = sigmas_h[s0 + (q < sg.n ? q : 0)];
            sg.inv[q] = kind == QMLH_PAIR_GAUSSIAN ? -0.5 / (sig * sig) : -1.0 / sig;
        }
        const int sym = symmetric ? 1 : 0;
        if (mfma) {
            if (sg.n == 1) lk_launch<LK_GAUSS_MFMA, 1>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
            else lk_launch<LK_GAUSS_MFMA, 4>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
        } else if (kind == QMLH_PAIR_GAUSSIAN) {
            if (sg.n == 1) lk_launch<LK_GAUSS_VALU, 1>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
            else lk_launch<LK_GAUSS_VALU, 4>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
        } else {
            if (sg.n == 1) lk_launch<LK_LAPLACE, 1>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
            else lk_launch<LK_LAPLACE, 4>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
        }
        QMLH_CHECK_LAUNCH("local_kernels launch");
    }
    return QMLH_OK;
}
