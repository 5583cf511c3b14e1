// Split-FP16 trailing update for the mixed-precision factor: C -= P P^T (lower tiles). The
// panel is scaled by a power of two, 2^e (chosen from the largest diagonal entry of the matrix
// so that every entry lies below 2^14, see sb_scale_exp), and split as Ph = fp16(2^e P),
// Pl = fp16(2^e P - Ph); then
//     2^2e P P^T ~ Ph Ph^T + Ph Pl^T + Pl Ph^T
// on v_mfma_f32_16x16x32_f16 with FP32 accumulation (each FP16 x FP16 product is exact in
// FP32; Ph + Pl carries 22 bits of each entry, and the dropped Pl Pl^T term is ~2^-22
// relative). C is scaled by 2^2e on load and by 2^-2e on store (exact). FP16 has the BF16
// MFMA rate and three more mantissa bits per half than BF16, so the split reaches almost
// FP32 accuracy (a 2^-16 split-BF16 version needed 11-13 refinement iterations where this
// needs 8); its narrower exponent range is what the scale is for. This is the dense O(n^3)
// part of the mixed factorisation; the panel, the diagonal blocks and the substitutions stay
// in FP32 storage / FP64 arithmetic.
//
// Panel layout. The split kernel writes Ph and Pl in MFMA fragment order: element (row, k)
// of the (m_pad x kb) panel at ((c * 4 + q) * m_pad + row) * 8 + e with k = 32 c + 8 q + e,
// so that one 32-deep chunk of 128 rows is four contiguous 2 KB runs, read with 16-byte loads
// and copied unchanged into LDS, where lane (r, g) of a wave finds its 8 k-values of row r
// (k = 8 g .. 8 g + 7) as one 16-byte read.
//
// Tiles: 256 lanes own a 128 x 128 lower tile of C (2 x 2 waves of 64 x 64); chunks of 32
// features are double buffered through LDS (4 panels x 8 KB per buffer), copied there by
// gfx950's direct global-to-LDS loads (no staging registers: 143 VGPRs, leaving room on each
// CU for the look-ahead panel's kernels on the side stream). Trailing updates of order 4096
// and above use 256 x 256 tiles instead (syrk_f16x3_wide_kernel).
//
// Tile order. Workgroups are handed to the 8 XCDs round robin (workgroup b to XCD b % 8), and
// each XCD has its own L2. The tiles are grouped in 8 x 8 super-blocks, and the workgroup
// index is remapped so that the 64 workgroups an XCD holds at a time are the 64 tiles of one
// super-block: they share 8 row blocks and 8 column blocks of the panel, which that XCD's L2
// then fetches once instead of once per tile row.
#include "common.h"
#include "timer.h"

namespace qmlh {

typedef _Float16 sb_elem;
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
// 16-byte staging registers. A native vector type: arrays of the HIP uint4 struct are copied
// with memcpy and end up in scratch memory instead of registers.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

static constexpr int SB_T = 128;          // tile rows
static constexpr int SB_KC = 32;          // features per chunk
static constexpr int SB_PANEL = SB_T * SB_KC;  // FP16 elements per panel per chunk (8 KB)

__device__ __forceinline__ int64_t split_index(int64_t row, int64_t k, int64_t m_pad) {
    const int64_t c = k >> 5, q = (k >> 3) & 3, e = k & 7;
    return ((c * 4 + q) * m_pad + row) * 8 + e;
}

// Panel scale. Every entry of a Cholesky factor obeys |L(i, j)| <= sqrt(A(i, i)), so with
// dmax the largest diagonal entry of the matrix being factored, the panels scaled by
// 2^e = 2^(14 - (ilogb(sqrt(dmax)) + 1)) lie below 2^14 in magnitude: inside FP16's range with
// room, and far above its subnormals for every entry that matters. hdr[0] holds dmax's bits.
static constexpr size_t SB_HDR_BYTES = 256;

__device__ __forceinline__ int sb_scale_exp(const unsigned* hdr) {
    const float dmax = __uint_as_float(hdr[0]);
    if (!(dmax > 0.0f) || !(dmax < 3.0e38f)) return 0;
    const int e = 14 - (ilogbf(sqrtf(dmax)) + 1);
    return e < -40 ? -40 : (e > 40 ? 40 : e);
}

// hdr[1] holds the bits of max |X| over a block X split with its own scale (the extra rows of
// a factorisation, whose entries the diagonal does not bound): 2^e with e = 14 - (ilogb(max) + 1)
// puts every entry below 2^14.
__device__ __forceinline__ int sb_scale_exp_abs(const unsigned* hdr) {
    const float amax = __uint_as_float(hdr[1]);
    if (!(amax > 0.0f) || !(amax < 3.0e38f)) return 0;
    const int e = 14 - (ilogbf(amax) + 1);
    return e < -40 ? -40 : (e > 40 ? 40 : e);
}

__global__ void sb_hdr_clear_kernel(unsigned* hdr) { hdr[threadIdx.x] = 0u; }

__global__ void sb_word_clear_kernel(unsigned* hdr, int idx) { hdr[idx] = 0u; }

// hdr[1] = bits of max |X(i, j)| over the finite entries of the r x w block (blockIdx.y = j).
// Workgroup (x, y): rows 256 x .., columns SB_MAXABS_COLS y ..; one atomic per workgroup (an
// atomic per wave and column serialised thousands of updates of the one word).
static constexpr int SB_MAXABS_COLS = 16;
__global__ __launch_bounds__(256) void sb_maxabs_kernel(const float* __restrict__ X, int64_t ldx,
                                                        int64_t r, int64_t kb,
                                                        unsigned* __restrict__ hdr) {
    __shared__ unsigned wmax[4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.y * SB_MAXABS_COLS;
    unsigned m = 0u;
    if (i < r) {
#pragma unroll 4
        for (int c = 0; c < SB_MAXABS_COLS; ++c) {
            if (c0 + c >= kb) break;
            const float v = fabsf(X[i + (c0 + c) * ldx]);
            const unsigned u = v < 3.0e38f ? __float_as_uint(v) : 0u;
            m = u > m ? u : m;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned w = wmax[0];
        for (int q = 1; q < 4; ++q) w = wmax[q] > w ? wmax[q] : w;
        if (w) atomicMax(hdr + 1, w);
    }
}

// hdr[0] = bits of max_i A(i, i) over the positive diagonal entries (positive floats order
// like their bit patterns).
__global__ __launch_bounds__(256) void sb_diag_max_kernel(const float* __restrict__ A,
                                                          int64_t n, int64_t lda,
                                                          unsigned* __restrict__ hdr) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const float v = i < n ? A[i + i * lda] : 0.0f;
    const unsigned u = (v > 0.0f && v < 3.0e38f) ? __float_as_uint(v) : 0u;
    unsigned m = u;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m) atomicMax(hdr, m);
}

// Ph, Pl <- split of P (column-major FP32, m x kb); rows m .. m_pad-1 and features
// kb .. kb_pad-1 are zero. Scale from the diagonal (hdr[0]) or, abs_scale, from max |P|
// (hdr[1]). One lane per (row, group of 8 features): 8 strided loads (lanes
// run down the rows, so each is coalesced) and one 16-byte store into each of Ph and Pl.
__global__ __launch_bounds__(256) void split_panel_kernel(const float* __restrict__ P,
                                                          int64_t ldp, int64_t m, int64_t m_pad,
                                                          int64_t kb, int64_t kb_pad,
                                                          sb_elem* __restrict__ Ph,
                                                          sb_elem* __restrict__ Pl,
                                                          const unsigned* __restrict__ hdr,
                                                          int abs_scale) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.y * 8;
    if (row >= m_pad || k0 >= kb_pad) return;
    const float sc = ldexpf(1.0f, abs_scale ? sb_scale_exp_abs(hdr) : sb_scale_exp(hdr));
    // loads from clamped indices (no per-load branch), zero selected afterwards
    const int64_t rr = row < m ? row : m - 1;
    float raw[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) raw[e] = P[rr + (k0 + e < kb ? k0 + e : kb - 1) * ldp];
    f16x8_t h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float v = (row < m && k0 + e < kb) ? raw[e] * sc : 0.0f;
        const sb_elem hv = (sb_elem)v;
        h[e] = hv;
        l[e] = (sb_elem)(v - (float)hv);
    }
    const int64_t idx = split_index(row, k0, m_pad);  // k0 % 8 == 0: 8 contiguous elements
    *reinterpret_cast<f16x8_t*>(Ph + idx) = h;
    *reinterpret_cast<f16x8_t*>(Pl + idx) = l;
}

// Copy one 32-deep chunk of 128 rows (4 runs of 128 x 16 B) straight from global memory into
// an LDS panel (gfx950's 16-byte global_load_lds: lane l of a wave writes 16 bytes at the
// wave's LDS base + 16 l, so each wave's 64 lanes fill one contiguous 1 KB run: element
// (q, row) of the panel at (q * 128 + row) * 8). No registers are staged; the data is in LDS
// once vmcnt reaches zero.
__device__ __forceinline__ void sb_load_lds(sb_elem* __restrict__ panel,
                                            const sb_elem* __restrict__ src, int64_t m_pad,
                                            int64_t row0, int64_t c) {
    const int t = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = t + 256 * h;  // 0 .. 511
        const int q = p >> 7, row = p & 127;
        const sb_elem* g = src + ((c * 4 + q) * m_pad + row0 + row) * 8;
        __builtin_amdgcn_global_load_lds(
            (const void*)g,
            (__attribute__((address_space(3))) void*)(panel + (256 * h + 64 * wave) * 8), 16, 0, 0);
    }
}

__device__ __forceinline__ f16x8_t sb_frag(const sb_elem* __restrict__ panel, int row, int g) {
    return *reinterpret_cast<const f16x8_t*>(panel + (g * SB_T + row) * 8);
}

// Flip the sign of 8 FP16 values (exact).
__device__ __forceinline__ f16x8_t sb_neg(f16x8_t v) {
    const u32x4_t u = __builtin_bit_cast(u32x4_t, v) ^ 0x80008000u;
    return __builtin_bit_cast(f16x8_t, u);
}

// C(i, j) -= sum_k P(i, k) P(j, k) for i >= j. The accumulators start from C -- its loads are
// issued together with the first panel chunk, so their latency hides behind the prologue --
// and the products are subtracted by negating the second operand's fragments; the epilogue
// only stores.
//
// Measured alternatives (mixed fit at 50,000 molecules, trailing updates per fit): staging the
// chunks through registers 157-161 ms (the DMA copy: 150 ms); v_mfma_f32_32x32x16_f16 with
// the same wave tile 173 ms; one LDS buffer and three workgroups per CU 125 ms for the updates
// alone, but the fit slows down (no registers left for the look-ahead panel), also when the
// look-ahead is dropped for the large steps; persistent grids with slots kept free for the
// panel 182-195 ms.
static constexpr int SB_SUPER = 8;  // super-block side in tiles
static constexpr int64_t SW_MIN = 4096;  // trailing updates of this order and above: wide tiles
static constexpr int SB_XCDS = 8;

// Launch size for t x t tiles: 64 workgroups per lower super-block, padded to a multiple of
// 8 super-blocks so that the remap below is a bijection.
__host__ __device__ inline int64_t sb_super_count(int64_t t) {
    const int64_t T = (t + SB_SUPER - 1) / SB_SUPER;
    return T * (T + 1) / 2;
}

// Workgroup v -> tile (bi, bj); false for padding and tiles above the diagonal.
// ncols == 0: v -> (XCD x = v % 8, its l-th workgroup) -> super-block 8 (l / 64) + x, tile
// l % 64, so the workgroups an XCD holds at a time share 8 row and 8 column blocks of the panel
// (one L2 per XCD). ncols > 0: row order over the next block column's tiles.
__device__ __forceinline__ bool sb_tile(int64_t v, int64_t t_tiles, int64_t ncols, int64_t& bi,
                                        int64_t& bj) {
    if (ncols > 0) {
        const int64_t tc = (ncols + SB_T - 1) / SB_T;
        bi = v / tc;
        bj = v % tc;
    } else {
        const int64_t x = v % SB_XCDS, l = v / SB_XCDS;
        const int64_t sbi = (l / (SB_SUPER * SB_SUPER)) * SB_XCDS + x;
        const int64_t lt = l % (SB_SUPER * SB_SUPER);
        if (sbi >= sb_super_count(t_tiles)) return false;
        int64_t I = (int64_t)((sqrt(8.0 * (double)sbi + 1.0) - 1.0) * 0.5);
        while (I * (I + 1) / 2 > sbi) --I;
        while ((I + 1) * (I + 2) / 2 <= sbi) ++I;
        const int64_t J = sbi - I * (I + 1) / 2;  // I >= J
        bi = I * SB_SUPER + lt / SB_SUPER;
        bj = J * SB_SUPER + lt % SB_SUPER;
    }
    return bi < t_tiles && bj <= bi;
}

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// A wave's 64 x 64 block as 4 x 4 blocks of 16 x 16: lane l holds, of block (a, b), accumulator
// e at output row row0 + 16 b + (l & 15) and column col0 + 16 a + 4 (l / 16) + e.
struct SbAcc {
    f32x4_t v[4][4];
};

struct SbFrags {
    f16x8_t fh[4], fl[4], sh[4], sl[4];
};

// The chunk's fragments from one LDS buffer (4 panels): first operand = column blocks, second
// (negated) = row blocks.
__device__ __forceinline__ void sb_read_frags(SbFrags& F, const sb_elem (*buf)[SB_PANEL], int fr,
                                              int sr, int lane) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        F.fh[q] = sb_frag(buf[2], fr + 16 * q + r, g);
        F.fl[q] = sb_frag(buf[3], fr + 16 * q + r, g);
        F.sh[q] = sb_neg(sb_frag(buf[0], sr + 16 * q + r, g));
        F.sl[q] = sb_neg(sb_frag(buf[1], sr + 16 * q + r, g));
    }
}

__device__ __forceinline__ void sb_products(SbAcc& A, const SbFrags& F) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            A.v[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.fl[a], F.sh[b], A.v[a][b], 0, 0, 0);
            A.v[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.fh[a], F.sl[b], A.v[a][b], 0, 0, 0);
            A.v[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.fh[a], F.sh[b], A.v[a][b], 0, 0, 0);
        }
}

// Ph / Pl point at row `off` of a split panel with row stride m_pad; the update covers m rows
// from there. ncols == 0: the lower triangle of the m x m block (super-block order);
// ncols > 0: the entries i >= j of its first ncols columns (the next block column of a
// look-ahead factorisation), tiles in row order. One workgroup per tile.
__global__ __launch_bounds__(256, 2) void syrk_f16x3_kernel(const sb_elem* __restrict__ Ph,
                                                             const sb_elem* __restrict__ Pl,
                                                             int64_t m, int64_t m_pad,
                                                             int64_t kb_pad, int64_t ncols,
                                                             float* __restrict__ C, int64_t ldc,
                                                             const int32_t* __restrict__ info,
                                                             const unsigned* __restrict__ hdr) {
    // buffers x {row hi, row lo, col hi, col lo}
    __shared__ __attribute__((aligned(16))) sb_elem lds[2][4][SB_PANEL];
    if (info && *info) return;
    const int64_t t_tiles = (m + SB_T - 1) / SB_T;
    int64_t bi, bj;
    if (!sb_tile(blockIdx.x, t_tiles, ncols, bi, bj)) return;
    // the panel is scaled by 2^e, so the products carry 2^(2e): C is scaled up on load and
    // back on store (exact powers of two)
    const int e2 = 2 * sb_scale_exp(hdr);
    const float up = ldexpf(1.0f, e2), down = ldexpf(1.0f, -e2);
    const int64_t ncol = ncols > 0 ? ncols : m;  // column bound
    const int64_t rows0 = bi * SB_T, cols0 = bj * SB_T;

    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int r = lane & 15, g = lane >> 4;
    // Wave w: first operand (accumulator rows) = output columns cols0 + 64 (w & 1),
    // second operand = output rows rows0 + 64 (w >> 1).
    const int fr = 64 * (w & 1), sr = 64 * (w >> 1);
    const int64_t col0 = cols0 + fr;
    const int64_t row0 = rows0 + sr;
    const bool live = row0 < m && col0 < ncol && col0 <= row0 + 63;
    // interior: the wave's 64 x 64 block lies inside the matrix and on or below the diagonal
    // in every entry, so its loads and stores need no masks.
    const bool interior = live && row0 + 63 < m && col0 + 63 < ncol && col0 + 63 <= row0;

    // first chunk in flight before C is read
    sb_load_lds(lds[0][0], Ph, m_pad, rows0, 0);
    sb_load_lds(lds[0][1], Pl, m_pad, rows0, 0);
    sb_load_lds(lds[0][2], Ph, m_pad, cols0, 0);
    sb_load_lds(lds[0][3], Pl, m_pad, cols0, 0);
    SbAcc acc;
    if (interior) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc.v[a][b][e] =
                        C[(row0 + 16 * b + r) + (col0 + 16 * a + 4 * g + e) * ldc] * up;
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t i = row0 + 16 * b + r;
                    const int64_t j = col0 + 16 * a + 4 * g + e;
                    acc.v[a][b][e] = (live && i < m && j <= i && j < ncol) ? C[i + j * ldc] * up : 0.0f;
                }
    }
    __builtin_amdgcn_s_waitcnt(0);  // this lane's copies have landed in LDS
    __syncthreads();

    const int64_t nchunks = kb_pad / SB_KC;
    int st = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
        if (c + 1 < nchunks) {  // the other buffer is free since the last barrier
            sb_load_lds(lds[st ^ 1][0], Ph, m_pad, rows0, c + 1);
            sb_load_lds(lds[st ^ 1][1], Pl, m_pad, rows0, c + 1);
            sb_load_lds(lds[st ^ 1][2], Ph, m_pad, cols0, c + 1);
            sb_load_lds(lds[st ^ 1][3], Pl, m_pad, cols0, c + 1);
        }
        SbFrags F;
        sb_read_frags(F, lds[st], fr, sr, lane);
        sb_products(acc, F);
        __builtin_amdgcn_s_waitcnt(0);  // the next chunk has landed in LDS
        __syncthreads();
        st ^= 1;
    }

    if (!live) return;
    if (interior) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    C[(row0 + 16 * b + r) + (col0 + 16 * a + 4 * g + e) * ldc] =
                        acc.v[a][b][e] * down;
        return;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t i = row0 + 16 * b + r;
            if (i >= m) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t j = col0 + 16 * a + 4 * g + e;
                if (j > i || j >= ncol) continue;
                C[i + j * ldc] = acc.v[a][b][e] * down;
            }
        }
}

// ------------------------------------------------------------------------------------------
// Wide tiles for the large trailing updates: 512 lanes own a 256 x 256 lower tile (4 x 2
// waves of 64 rows x 128 columns, 128 accumulators per lane), chunks double buffered in
// 128 KB of LDS, one workgroup per CU. Per flop the tile copies half the panel bytes of the
// 128 x 128 tile and issues a quarter fewer fragment reads per MFMA. Measured alone at m = 46,080, kb = 768: 4.6-4.8 ms against 5.1-5.3 ms for the 128 x 128
// tile; a 256 x 128 tile with a three-stage ring (chunks issued two ahead) took 5.7 ms. What
// the wide tile leaves exposed is C itself (read and written once per launch, 8.5 GB at that
// size), which one workgroup per CU cannot overlap with its own products.
// Tiles are grouped in 4 x 4 super-blocks, each XCD taking whole super-blocks as above.
// ------------------------------------------------------------------------------------------
static constexpr int SW_T = 256;
static constexpr int SW_PANEL = SW_T * SB_KC;  // 16 KB
static constexpr int SW_SUPER = 4;
static constexpr int SW_THREADS = 512;

__host__ __device__ inline int64_t sw_super_count(int64_t t) {
    const int64_t T = (t + SW_SUPER - 1) / SW_SUPER;
    return T * (T + 1) / 2;
}

// Workgroup v -> tile (bi, bj); false for padding and tiles above the diagonal. ncols == 0:
// the lower triangle in super-blocks as above. ncols > 0 (the first ncols columns of the m
// rows, a block column of the next panel): v -> (XCD x = v % 8, its l-th workgroup) -> tile row
// 8 (l / tc) + x, tile column l % tc, so one XCD runs the tc tiles of a tile row together (they
// share its row panel).
__device__ __forceinline__ bool sw_tile(int64_t v, int64_t t_tiles, int64_t ncols, int64_t& bi,
                                        int64_t& bj) {
    const int64_t x = v % SB_XCDS, l = v / SB_XCDS;
    if (ncols > 0) {
        const int64_t tc = (ncols + SW_T - 1) / SW_T;
        bi = (l / tc) * SB_XCDS + x;
        bj = l % tc;
        return bi < t_tiles && bj <= bi;
    }
    const int64_t sbi = (l / (SW_SUPER * SW_SUPER)) * SB_XCDS + x;
    const int64_t lt = l % (SW_SUPER * SW_SUPER);
    if (sbi >= sw_super_count(t_tiles)) return false;
    int64_t I = (int64_t)((sqrt(8.0 * (double)sbi + 1.0) - 1.0) * 0.5);
    while (I * (I + 1) / 2 > sbi) --I;
    while ((I + 1) * (I + 2) / 2 <= sbi) ++I;
    const int64_t J = sbi - I * (I + 1) / 2;
    bi = I * SW_SUPER + lt / SW_SUPER;
    bj = J * SW_SUPER + lt % SW_SUPER;
    return bi < t_tiles && bj <= bi;
}

// One 32-deep chunk of 256 rows (4 runs of 256 x 16 B) into an LDS panel, element (q, row) at
// (q * 256 + row) * 8; rows past rmax (the last row of the split) read row rmax, whose
// products land only in entries outside the update.
__device__ __forceinline__ void sw_load_lds(sb_elem* __restrict__ panel,
                                            const sb_elem* __restrict__ src, int64_t m_pad,
                                            int64_t row0, int64_t rmax, int64_t c) {
    const int t = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = t + SW_THREADS * h;  // 0 .. 1023
        const int q = p >> 8, row = p & 255;
        const int64_t rr = row0 + row < rmax ? row0 + row : rmax;
        const sb_elem* g = src + ((c * 4 + q) * m_pad + rr) * 8;
        __builtin_amdgcn_global_load_lds(
            (const void*)g,
            (__attribute__((address_space(3))) void*)(panel + (SW_THREADS * h + 64 * wave) * 8),
            16, 0, 0);
    }
}

__device__ __forceinline__ f16x8_t sw_frag(const sb_elem* __restrict__ panel, int row, int g) {
    return *reinterpret_cast<const f16x8_t*>(panel + (g * SW_T + row) * 8);
}

// A wave's 64 x 128 block as 8 x 4 blocks of 16 x 16: lane l holds, of block (a, b),
// accumulator e at output row row0 + 16 b + (l & 15) and column col0 + 16 a + 4 (l / 16) + e.
struct SwAcc {
    f32x4_t v[8][4];
};

// C(i, j) -= sum_k P(i, k) P(j, k), i >= j, over the lower triangle of the m x m block, or with
// ncols > 0 over its first ncols columns (COLS: a template argument so that the two uses carry
// different names in kernel traces); Ph / Pl at row `off` of the split, rmax = the last row of
// the split relative to that.
template <bool COLS>
__global__ __launch_bounds__(SW_THREADS, 1) void syrk_f16x3_wide_kernel(
    const sb_elem* __restrict__ Ph, const sb_elem* __restrict__ Pl, int64_t m, int64_t m_pad,
    int64_t rmax, int64_t kb_pad, int64_t ncols, float* __restrict__ C, int64_t ldc,
    const int32_t* __restrict__ info, const unsigned* __restrict__ hdr) {
    // buffers x {row hi, row lo, col hi, col lo}
    __shared__ __attribute__((aligned(16))) sb_elem lds[2][4][SW_PANEL];
    if (info && *info) return;
    const int64_t t_tiles = (m + SW_T - 1) / SW_T;
    const int64_t nc = COLS ? ncols : m;  // columns of the update
    int64_t bi, bj;
    if (!sw_tile(blockIdx.x, t_tiles, COLS ? ncols : 0, bi, bj)) return;
    const int e2 = 2 * sb_scale_exp(hdr);
    const float up = ldexpf(1.0f, e2), down = ldexpf(1.0f, -e2);
    const int64_t rows0 = bi * SW_T, cols0 = bj * SW_T;

    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int r = lane & 15, g = lane >> 4;
    // wave w: first operand (accumulator rows) = output columns cols0 + 128 (w & 1),
    // second operand = output rows rows0 + 64 (w >> 1)
    const int fr = 128 * (w & 1), sr = 64 * (w >> 1);
    const int64_t col0 = cols0 + fr;
    const int64_t row0 = rows0 + sr;
    const bool live = row0 < m && col0 < nc && col0 <= row0 + 63;
    const bool interior = live && row0 + 63 < m && col0 + 127 < nc && col0 + 127 <= row0;

    sw_load_lds(lds[0][0], Ph, m_pad, rows0, rmax, 0);
    sw_load_lds(lds[0][1], Pl, m_pad, rows0, rmax, 0);
    sw_load_lds(lds[0][2], Ph, m_pad, cols0, rmax, 0);
    sw_load_lds(lds[0][3], Pl, m_pad, cols0, rmax, 0);
    SwAcc acc;
    if (interior) {
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc.v[a][b][e] =
                        C[(row0 + 16 * b + r) + (col0 + 16 * a + 4 * g + e) * ldc] * up;
    } else {
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t i = row0 + 16 * b + r;
                    const int64_t j = col0 + 16 * a + 4 * g + e;
                    acc.v[a][b][e] =
                        (live && i < m && j < nc && j <= i) ? C[i + j * ldc] * up : 0.0f;
                }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    const int64_t nchunks = kb_pad / SB_KC;
    int st = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
        if (c + 1 < nchunks) {
            sw_load_lds(lds[st ^ 1][0], Ph, m_pad, rows0, rmax, c + 1);
            sw_load_lds(lds[st ^ 1][1], Pl, m_pad, rows0, rmax, c + 1);
            sw_load_lds(lds[st ^ 1][2], Ph, m_pad, cols0, rmax, c + 1);
            sw_load_lds(lds[st ^ 1][3], Pl, m_pad, cols0, rmax, c + 1);
        }
        f16x8_t sh[4], sl[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            sh[b] = sb_neg(sw_frag(lds[st][0], sr + 16 * b + r, g));
            sl[b] = sb_neg(sw_frag(lds[st][1], sr + 16 * b + r, g));
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const f16x8_t fh = sw_frag(lds[st][2], fr + 16 * a + r, g);
            const f16x8_t fl = sw_frag(lds[st][3], fr + 16 * a + r, g);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                acc.v[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fl, sh[b], acc.v[a][b], 0, 0, 0);
                acc.v[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fh, sl[b], acc.v[a][b], 0, 0, 0);
                acc.v[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fh, sh[b], acc.v[a][b], 0, 0, 0);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        st ^= 1;
    }

    if (!live) return;
    if (interior) {
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    C[(row0 + 16 * b + r) + (col0 + 16 * a + 4 * g + e) * ldc] =
                        acc.v[a][b][e] * down;
        return;
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t i = row0 + 16 * b + r;
            if (i >= m) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t j = col0 + 16 * a + 4 * g + e;
                if (j > i || j >= nc) continue;
                C[i + j * ldc] = acc.v[a][b][e] * down;
            }
        }
}

size_t f16x3_region_bytes(int64_t n, int64_t kb) {
    const int64_t m_pad = round_up(n > 0 ? n : 1, SB_T);
    const int64_t kb_pad = round_up(kb, SB_KC);
    return (size_t)(2 * m_pad * kb_pad) * sizeof(sb_elem);
}

size_t syrk_f16_workspace(int64_t n, int64_t kb) {
    return SB_HDR_BYTES + f16x3_region_bytes(n, kb);
}

int f16x3_prepare(const float* A, int64_t n, int64_t lda, void* ws, hipStream_t s) {
    unsigned* hdr = static_cast<unsigned*>(ws);
    hipLaunchKernelGGL(sb_hdr_clear_kernel, dim3(1), dim3(SB_HDR_BYTES / 4), 0, s, hdr);
    if (n > 0)
        hipLaunchKernelGGL(sb_diag_max_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                           A, n, lda, hdr);
    QMLH_CHECK_LAUNCH("split scale launch");
    return QMLH_OK;
}

static sb_elem* sb_panels(void* ws, size_t region) {
    return reinterpret_cast<sb_elem*>(static_cast<char*>(ws) + SB_HDR_BYTES + region);
}

int f16x3_split(const float* P, int64_t ldp, int64_t m, int64_t kb, void* ws, hipStream_t s,
                size_t region) {
    if (m <= 0 || kb <= 0) return QMLH_OK;
    const int64_t m_pad = round_up(m, SB_T);
    const int64_t kb_pad = round_up(kb, SB_KC);
    sb_elem* Ph = sb_panels(ws, region);
    sb_elem* Pl = Ph + m_pad * kb_pad;
    hipLaunchKernelGGL(split_panel_kernel, dim3((unsigned)ceil_div(m_pad, 256), (unsigned)(kb_pad / 8)),
                       dim3(256), 0, s, P, ldp, m, m_pad, kb, kb_pad, Ph, Pl,
                       static_cast<const unsigned*>(ws), 0);
    QMLH_CHECK_LAUNCH("split_panel launch");
    return QMLH_OK;
}

int f16x3_update(const void* ws, int64_t m_split, int64_t kb, int64_t off, int64_t m,
                  int64_t ncols, float* C, int64_t ldc, const int32_t* info, hipStream_t s,
                  size_t region, bool wide_cols) {
    if (m <= 0 || kb <= 0) return QMLH_OK;
    const int64_t m_pad = round_up(m_split, SB_T);
    const int64_t kb_pad = round_up(kb, SB_KC);
    if (off % SB_T != 0 || off + round_up(m, SB_T) > m_pad)
        return set_error(QMLH_ERR_ARG, "f16x3_update: rows %lld + %lld outside the split",
                         (long long)off, (long long)m);
    const sb_elem* Ph = sb_panels(const_cast<void*>(ws), region) + off * 8;
    const sb_elem* Pl = sb_panels(const_cast<void*>(ws), region) + m_pad * kb_pad + off * 8;
    if (ncols == 0 && m >= SW_MIN) {
        const int64_t tw = ceil_div(m, (int64_t)SW_T);
        const int64_t grid = ceil_div(sw_super_count(tw), (int64_t)SB_XCDS) * SB_XCDS * SW_SUPER *
                             SW_SUPER;
        if (grid > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "f16x3_update: grid too large");
        TimerScope ts(QMLH_TIMER_SYRK, (double)m * (double)(m + 1) * (double)kb, s);
        hipLaunchKernelGGL(syrk_f16x3_wide_kernel<false>, dim3((unsigned)grid), dim3(SW_THREADS),
                           0, s, Ph, Pl, m, m_pad, m_pad - off - 1, kb_pad, (int64_t)0, C, ldc,
                           info, static_cast<const unsigned*>(ws));
        QMLH_CHECK_LAUNCH("syrk_f16x3_wide launch");
        return QMLH_OK;
    }
    if (ncols > 0 && m >= SW_MIN && wide_cols) {
        const int64_t tc = ceil_div(ncols, (int64_t)SW_T);
        const int64_t grid = ceil_div(ceil_div(m, (int64_t)SW_T), (int64_t)SB_XCDS) * SB_XCDS * tc;
        if (grid > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "f16x3_update: grid too large");
        TimerScope ts(QMLH_TIMER_GEMM,
                      ((double)m * (double)ncols - (double)ncols * (double)(ncols - 1) / 2.0) *
                          2.0 * (double)kb,
                      s);
        hipLaunchKernelGGL(syrk_f16x3_wide_kernel<true>, dim3((unsigned)grid), dim3(SW_THREADS),
                           0, s, Ph, Pl, m, m_pad, m_pad - off - 1, kb_pad, ncols, C, ldc, info,
                           static_cast<const unsigned*>(ws));
        QMLH_CHECK_LAUNCH("syrk_f16x3_wide columns launch");
        return QMLH_OK;
    }
    const int64_t tm = ceil_div(m, (int64_t)SB_T);
    int64_t grid;
    double flops;
    if (ncols > 0) {
        grid = tm * ceil_div(ncols, (int64_t)SB_T);
        flops = ((double)m * (double)ncols - (double)ncols * (double)(ncols - 1) / 2.0) * 2.0 *
                (double)kb;
    } else {
        grid = ceil_div(sb_super_count(tm), (int64_t)SB_XCDS) * SB_XCDS * SB_SUPER * SB_SUPER;
        flops = (double)m * (double)(m + 1) * (double)kb;
    }
    if (grid > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "f16x3_update: grid too large");
    TimerScope ts(ncols > 0 ? QMLH_TIMER_GEMM : QMLH_TIMER_SYRK, flops, s);
    hipLaunchKernelGGL(syrk_f16x3_kernel, dim3((unsigned)grid), dim3(256), 0, s, Ph, Pl, m,
                       m_pad, kb_pad, ncols, C, ldc, info, static_cast<const unsigned*>(ws));
    QMLH_CHECK_LAUNCH("syrk_f16x3 launch");
    return QMLH_OK;
}

// C(i, j) -= sum_k X(i, k) P(j, k) for all i < r, j < m: X (r x kb) split with its own scale
// (hdr[1]), P a split of m rows (hdr[0]); the products carry 2^(e_x + e_p). 128 x 128 tiles, one
// per workgroup, the X tiles of one P column block adjacent in launch order (they share its
// chunks).
__global__ __launch_bounds__(256, 2) void gemm_f16x3_kernel(
    const sb_elem* __restrict__ Xh, const sb_elem* __restrict__ Xl, int64_t r, int64_t x_pad,
    const sb_elem* __restrict__ Ph, const sb_elem* __restrict__ Pl, int64_t m, int64_t m_pad,
    int64_t kb_pad, float* __restrict__ C, int64_t ldc, const int32_t* __restrict__ info,
    const unsigned* __restrict__ hdr) {
    __shared__ __attribute__((aligned(16))) sb_elem lds[2][4][SB_PANEL];
    if (info && *info) return;
    const int64_t tr = (r + SB_T - 1) / SB_T;
    const int64_t bi = blockIdx.x % tr, bj = blockIdx.x / tr;
    const int e2 = sb_scale_exp(hdr) + sb_scale_exp_abs(hdr);
    const float up = ldexpf(1.0f, e2), down = ldexpf(1.0f, -e2);
    const int64_t rows0 = bi * SB_T, cols0 = bj * SB_T;
    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int rr = lane & 15, g = lane >> 4;
    const int fr = 64 * (w & 1), sr = 64 * (w >> 1);
    const int64_t col0 = cols0 + fr, row0 = rows0 + sr;
    const bool live = row0 < r && col0 < m;
    const bool interior = row0 + 63 < r && col0 + 63 < m;
    sb_load_lds(lds[0][0], Xh, x_pad, rows0, 0);
    sb_load_lds(lds[0][1], Xl, x_pad, rows0, 0);
    sb_load_lds(lds[0][2], Ph, m_pad, cols0, 0);
    sb_load_lds(lds[0][3], Pl, m_pad, cols0, 0);
    SbAcc acc;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = row0 + 16 * b + rr, j = col0 + 16 * a + 4 * g + e;
                acc.v[a][b][e] = (interior || (i < r && j < m)) ? C[i + j * ldc] * up : 0.0f;
            }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int64_t nchunks = kb_pad / SB_KC;
    int st = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
        if (c + 1 < nchunks) {
            sb_load_lds(lds[st ^ 1][0], Xh, x_pad, rows0, c + 1);
            sb_load_lds(lds[st ^ 1][1], Xl, x_pad, rows0, c + 1);
            sb_load_lds(lds[st ^ 1][2], Ph, m_pad, cols0, c + 1);
            sb_load_lds(lds[st ^ 1][3], Pl, m_pad, cols0, c + 1);
        }
        SbFrags F;
        sb_read_frags(F, lds[st], fr, sr, lane);
        sb_products(acc, F);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        st ^= 1;
    }
    if (!live) return;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = row0 + 16 * b + rr, j = col0 + 16 * a + 4 * g + e;
                if (interior || (i < r && j < m)) C[i + j * ldc] = acc.v[a][b][e] * down;
            }
}

int f16x3_gemm_rows(const float* X, int64_t ldx, int64_t r, int64_t kb, void* ws,
                    size_t region_x, size_t region_p, int64_t m, float* C, int64_t ldc,
                    const int32_t* info, hipStream_t s) {
    if (r <= 0 || m <= 0 || kb <= 0) return QMLH_OK;
    unsigned* hdr = static_cast<unsigned*>(ws);
    const int64_t x_pad = round_up(r, SB_T), m_pad = round_up(m, SB_T);
    const int64_t kb_pad = round_up(kb, SB_KC);
    sb_elem* Xh = sb_panels(ws, region_x);
    sb_elem* Xl = Xh + x_pad * kb_pad;
    hipLaunchKernelGGL(sb_word_clear_kernel, dim3(1), dim3(1), 0, s, hdr, 1);
    hipLaunchKernelGGL(sb_maxabs_kernel, dim3((unsigned)ceil_div(r, 256),
                                              (unsigned)ceil_div(kb, SB_MAXABS_COLS)),
                       dim3(256), 0, s, X, ldx, r, kb, hdr);
    hipLaunchKernelGGL(split_panel_kernel, dim3((unsigned)ceil_div(x_pad, 256), (unsigned)(kb_pad / 8)),
                       dim3(256), 0, s, X, ldx, r, x_pad, kb, kb_pad, Xh, Xl,
                       static_cast<const unsigned*>(ws), 1);
    QMLH_CHECK_LAUNCH("extra rows split launch");
    const sb_elem* Ph = sb_panels(ws, region_p);
    const sb_elem* Pl = Ph + m_pad * kb_pad;
    const int64_t grid = ceil_div(r, (int64_t)SB_T) * ceil_div(m, (int64_t)SB_T);
    if (grid > 0x7fffffffLL) return set_error(QMLH_ERR_ARG, "f16x3_gemm_rows: grid too large");
    TimerScope ts(QMLH_TIMER_GEMM, 2.0 * (double)r * (double)m * (double)kb, s);
    hipLaunchKernelGGL(gemm_f16x3_kernel, dim3((unsigned)grid), dim3(256), 0, s, Xh, Xl, r, x_pad,
                       Ph, Pl, m, m_pad, kb_pad, C, ldc, info, static_cast<const unsigned*>(ws));
    QMLH_CHECK_LAUNCH("gemm_f16x3 launch");
    return QMLH_OK;
}

// C (m x m, lower) -= P P^T with P (m x kb) column-major FP32; ws holds
// syrk_f16_workspace(m, kb) bytes, prepared by f16x3_prepare for the matrix being factored.
int syrk_f16x3(const float* P, int64_t ldp, int64_t m, int64_t kb, float* C, int64_t ldc,
                const int32_t* info, void* ws, hipStream_t s) {
    if (m <= 0 || kb <= 0) return QMLH_OK;
    QMLH_TRY(f16x3_split(P, ldp, m, kb, ws, s, 0));
    return f16x3_update(ws, m, kb, 0, m, 0, C, ldc, info, s, 0, false);
}

}  // namespace qmlh
