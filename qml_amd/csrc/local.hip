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
static constexpr int LK_MOLS = QMLH_LOCAL_PANEL_MOLS;  // molecules per panel: one MFMA block
static constexpr int LK_NS = 4;                         // sigmas per launch
static_assert(LK_MOLS == 16, "one 16-wide MFMA block of molecules per panel side");

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
    int n;          // sigmas in this launch (<= LK_NS)
    int64_t first;  // index of inv[0] in the caller's list
};

__device__ __forceinline__ uint32_t lk_byte(uint32_t p, int i) { return (p >> (8 * i)) & 0xffu; }

// sg.inv[s] without a dynamic index into the by-value argument (which would go to scratch).
__device__ __forceinline__ double lk_inv(const LkSigmas& sg, int s) {
    return s == 0 ? sg.inv[0] : s == 1 ? sg.inv[1] : s == 2 ? sg.inv[2] : sg.inv[3];
}

// Panel-local molecule index of `atom` from the panel's atom offsets staged in LDS (soff[m] =
// first atom of the panel's molecule m), -1 past the panel's last atom.
__device__ __forceinline__ int lk_local_mol(const int* __restrict__ soff, int nm, int atom,
                                            int atom_end) {
    if (atom >= atom_end) return -1;
    int lo = 0;
    for (int m = 1; m < nm; ++m)
        if (soff[m] <= atom) lo = m;
    return lo;
}

// Add this wave's molecule-pair sums of one sigma to S.
// Lane layout shared by both tile kinds: value val(r, c) belongs to tile row rowof(r, lane & 15)
// and tile column colof(c, lane >> 4); byte e of rpack[r] is the panel molecule of tile row
// rowof(r, (lane >> 4) + 4e), byte c % 4 of cpack[c / 4] that of tile column colof(c, lane >> 4)
// (0xff: padding, matching no molecule).
// Stage 1, per row set r: T(m, n) = sum_c val(r, c) [cmol[c] = n] (16 rows x 16 molecules);
// stage 2: S(ma, n) += sum_e [rmol[r][e] = ma] T((lane >> 4) + 4e, n).
template <int NR, int NC, typename Val>
__device__ __forceinline__ void lk_reduce(Val val, const uint32_t (&rpack)[NR],
                                          const uint32_t (&cpack)[NC / 4], double4_t& S) {
    typedef MfmaTraits<double> M;
    const uint32_t me = threadIdx.x & 15;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double4_t T = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < NC; ++c)
            T = M::mma(val(r, c), lk_byte(cpack[c / 4], c % 4) == me ? 1.0 : 0.0, T);
#pragma unroll
        for (int e = 0; e < 4; ++e) S = M::mma(lk_byte(rpack[r], e) == me ? 1.0 : 0.0, T[e], S);
    }
}

// acc[ma][n] += S in wave order (one barrier per wave); lane element e holds
// S((lane >> 4) + 4e, lane & 15).
__device__ __forceinline__ void lk_accumulate(double* __restrict__ acc, const double4_t& S) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    for (int w2 = 0; w2 < 4; ++w2) {
        if (w == w2) {
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[((lane >> 4) + 4 * e) * LK_MOLS + (lane & 15)] += S[e];
        }
        __syncthreads();
    }
}

template <int KIND>
__global__ __launch_bounds__(256, 2) void local_kernels_kernel(LkSide A, LkSide B, int64_t d,
                                                               LkSigmas sg, double* __restrict__ K,
                                                               int64_t ldk, int64_t sstride,
                                                               int symmetric, int tiles_a,
                                                               int ntiles) {
    __shared__ __attribute__((aligned(16))) double lds[KIND == LK_GAUSS_MFMA ? WG_LDS_ELEMS
                                                                              : 2 * PT_BK * PT_LDS];
    __shared__ double acc[LK_NS][LK_MOLS * LK_MOLS];
    __shared__ int smolA[LK_TILE], smolB[LK_TILE];
    __shared__ int soffA[LK_MOLS], soffB[LK_MOLS];

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;

    // Atom, molecule and panel indices fit in 32 bits (checked by the host); fewer 64-bit
    // values keep the scalar registers from spilling.
    for (int bid = blockIdx.x; bid < ntiles; bid += gridDim.x) {
        int bi, bj;
        if (symmetric) {
            int64_t j = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
            while (j * (j + 1) / 2 > bid) --j;
            while ((j + 1) * (j + 2) / 2 <= bid) ++j;
            bj = (int)j;
            bi = (int)(bid - j * (j + 1) / 2);
        } else {
            bi = bid % tiles_a;
            bj = bid / tiles_a;
        }
        const int ma0 = (int)A.panels[bi], mb0 = (int)B.panels[bj];
        const int nma = (int)(A.panels[bi + 1] - ma0), nmb = (int)(B.panels[bj + 1] - mb0);
        const int ra0 = (int)A.offsets[ma0], ra1 = (int)A.offsets[ma0 + nma];
        const int rb0 = (int)B.offsets[mb0], rb1 = (int)B.offsets[mb0 + nmb];
        // Panels come from qmlh_local_panels; one with more molecules than the LDS sums hold
        // is skipped rather than overrunning them.
        if (nma > LK_MOLS || nmb > LK_MOLS || nma < 0 || nmb < 0) continue;

        for (int q = t; q < LK_NS * LK_MOLS * LK_MOLS; q += 256) (&acc[0][0])[q] = 0.0;
        if (t < nma) soffA[t] = (int)A.offsets[ma0 + t];
        else if (t >= 64 && t - 64 < nmb) soffB[t - 64] = (int)B.offsets[mb0 + t - 64];
        __syncthreads();

        for (int r0 = ra0; r0 < ra1; r0 += LK_TILE) {
            for (int c0 = rb0; c0 < rb1; c0 += LK_TILE) {
                if (t < LK_TILE) smolA[t] = lk_local_mol(soffA, nma, r0 + t, ra1);
                else smolB[t - LK_TILE] = lk_local_mol(soffB, nmb, c0 + t - LK_TILE, rb1);
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
                    uint32_t rpack[4], cpack[4];
                    double ni[4], nj[16];
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int i = r0 + rbase + 16 * b + (lane & 15);
                        ni[b] = A.norms[i < A.atoms ? i : A.atoms - 1];
                        uint32_t p = 0;
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            p |= (uint32_t)(smolA[rbase + 16 * b + (lane >> 4) + 4 * e] & 0xff) << (8 * e);
                        rpack[b] = p;
                    }
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        uint32_t p = 0;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int cl = cbase + 16 * a + (lane >> 4) + 4 * e;
                            const int j = c0 + cl;
                            nj[4 * a + e] = B.norms[j < B.atoms ? j : B.atoms - 1];
                            p |= (uint32_t)(smolB[cl] & 0xff) << (8 * e);
                        }
                        cpack[a] = p;
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
#pragma unroll 1
                    for (int s = 0; s < sg.n; ++s) {
                        const double inv = lk_inv(sg, s);
                        double4_t S = {0.0, 0.0, 0.0, 0.0};
                        lk_reduce<4, 16>([&](int r, int c) { return exp(inv * d2[r][c]); },
                                         rpack, cpack, S);
                        lk_accumulate(acc[s], S);
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
                    uint32_t rpack[8], cpack[2];
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        uint32_t p = 0;
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            p |= (uint32_t)(smolA[2 * ((lane >> 4) + 4 * e) + 32 * (r / 2) + r % 2] & 0xff)
                                 << (8 * e);
                        rpack[r] = p;
                    }
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        uint32_t p = 0;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int c = 4 * q + u;
                            p |= (uint32_t)(smolB[2 * (4 * w + (lane >> 4)) + 32 * (c / 2) + c % 2] & 0xff)
                                 << (8 * u);
                        }
                        cpack[q] = p;
                    }
#pragma unroll 1
                    for (int s = 0; s < sg.n; ++s) {
                        const double inv = lk_inv(sg, s);
                        double4_t S = {0.0, 0.0, 0.0, 0.0};
                        lk_reduce<8, 8>([&](int r, int c) { return exp(dist[r][c] * inv); },
                                        rpack, cpack, S);
                        lk_accumulate(acc[s], S);
                    }
                }
            }
        }
        __syncthreads();

        // Write the panel pair's molecule sums; mirror for a symmetric call.
        const bool diag = symmetric && bi == bj;
        for (int s = 0; s < sg.n; ++s) {
            double* Ks = K + (sg.first + s) * sstride;
            for (int q = t; q < LK_MOLS * LK_MOLS; q += 256) {
                const int ma = q % LK_MOLS, mb = q / LK_MOLS;
                if (ma >= nma || mb >= nmb) continue;
                if (diag && ma > mb) continue;
                const double val = acc[s][ma * LK_MOLS + mb];
                const int64_t ga = (int64_t)ma0 + ma, gb = (int64_t)mb0 + mb;
                Ks[ga + gb * ldk] = val;
                if (symmetric && ga != gb) Ks[gb + ga * ldk] = val;
            }
        }
        __syncthreads();
    }
}

template <int KIND>
static void lk_launch(int64_t ntiles, hipStream_t s, const LkSide& A, const LkSide& B,
                      int64_t d, const LkSigmas& sg, double* K, int64_t ldk, int64_t sstride,
                      int symmetric, int64_t tiles_a) {
    const int64_t grid = ntiles < LK_MAX_GRID ? ntiles : LK_MAX_GRID;
    hipLaunchKernelGGL((local_kernels_kernel<KIND>), dim3((unsigned)grid), dim3(256), 0, s, A,
                       B, d, sg, K, ldk, sstride, symmetric, (int)tiles_a, (int)ntiles);
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
    if (grid > 0x7fffffffLL || atomsA > 0x7fffffffLL - LK_TILE || atomsB > 0x7fffffffLL - LK_TILE)
        return set_error(QMLH_ERR_ARG, "local_kernels: more than 2^31 atoms or panel pairs");
    hipStream_t s = (hipStream_t)stream;
    // Work: atom pairs x features (the pairs of a symmetric call counted once).
    const double pairs = symmetric ? 0.5 * (double)atomsA * (double)(atomsA + 1)
                                   : (double)atomsA * (double)atomsB;
    TimerScope ts(QMLH_TIMER_LOCAL, pairs * (double)d, s);
    for (int32_t s0 = 0; s0 < nsig; s0 += LK_NS) {
        LkSigmas sg;
        sg.n = nsig - s0 < LK_NS ? (int)(nsig - s0) : LK_NS;
        sg.first = s0;
        for (int q = 0; q < LK_NS; ++q) {
            const double sig = sigmas_h[s0 + (q < sg.n ? q : 0)];
            sg.inv[q] = kind == QMLH_PAIR_GAUSSIAN ? -0.5 / (sig * sig) : -1.0 / sig;
        }
        const int sym = symmetric ? 1 : 0;
        if (mfma) lk_launch<LK_GAUSS_MFMA>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
        else if (kind == QMLH_PAIR_GAUSSIAN)
            lk_launch<LK_GAUSS_VALU>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
        else lk_launch<LK_LAPLACE>(grid, s, sa, sb, d, sg, K, ldk, sstride, sym, npA);
        QMLH_CHECK_LAUNCH("local_kernels launch");
    }
    return QMLH_OK;
}
