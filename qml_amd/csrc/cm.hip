// Batched Coulomb-matrix generation: one 64-lane wavefront per molecule.
//
// Reproduces fgenerate_coulomb_matrix (qml/frepresentations.f90:52-140) and
// fgenerate_unsorted_coulomb_matrix (qml/frepresentations.f90:142-205) bit for bit:
//  * off-diagonal entries Z_i Z_j / sqrt((dx*dx + dy*dy) + dz*dz), i < j, with every product
//    and sum rounded separately (FP contraction is off in this file) and IEEE sqrt/division;
//  * diagonal 0.5 * Z**2.4 from a table the host fills with libm pow (integer Z < 256);
//  * row norms in the one-thread evaluation order of the reference's OpenMP reductions:
//    rn_k = d_k*d_k + (sum over partners j != k, ascending j, of M_kj^2);
//  * the ordering is the reference's repeated minloc (first index of the minimum, picked
//    entries replaced by huge()), emulated step by step as a wave-wide argmin, so ties and
//    non-finite norms order exactly as the Fortran loop orders them;
//  * packing cm[m(m+1)/2 + n] = M[s_m, s_n] for n <= m (0-based), zero beyond natoms.
#include "common.h"
#include "timer.h"

#include <float.h>

#pragma clang fp contract(off)

namespace qmlh {

__constant__ double c_cm_diag[256];

int upload_cm_diag_table(const double* table_h, int n) {
    return check_hip(hipMemcpyToSymbol(HIP_SYMBOL(c_cm_diag), table_h, sizeof(double) * n),
                     "upload Coulomb diagonal table");
}

static constexpr int CM_WAVE = 64;
static constexpr int CM_MAX_ATOMS = QMLH_CM_MAX_SIZE;
static constexpr int CM_PER_LANE = CM_MAX_ATOMS / CM_WAVE;

__device__ __forceinline__ double cm_diag(double z) {
    int zi = (int)z;
    if (z == (double)zi && zi >= 0 && zi < 256) return c_cm_diag[zi];
    return 0.5 * pow(z, 2.4);
}

// Z_i Z_j / |R_j - R_i| for atoms i < j held in LDS.
__device__ __forceinline__ double cm_pair(const double* sR, const double* sZ, int i, int j) {
    double dx = sR[3 * j + 0] - sR[3 * i + 0];
    double dy = sR[3 * j + 1] - sR[3 * i + 1];
    double dz = sR[3 * j + 2] - sR[3 * i + 2];
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return sZ[i] * sZ[j] / sqrt(s);
}

// Packed index p -> (m, n) with p = m(m+1)/2 + n, 0 <= n <= m.
__device__ __forceinline__ void cm_unpack(int p, int& m, int& n) {
    int mm = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (mm * (mm + 1) / 2 > p) --mm;
    while ((mm + 1) * (mm + 2) / 2 <= p) ++mm;
    m = mm;
    n = p - mm * (mm + 1) / 2;
}

// (value, index) argmin with the minloc tie rule: the smaller index wins on equal values.
__device__ __forceinline__ bool cm_better(double v, int i, double bv, int bi) {
    return (v < bv) || (v == bv && i < bi);
}

template <bool SORTED>
__global__ __launch_bounds__(CM_WAVE) void cm_kernel(const double* __restrict__ Z,
                                                      const double* __restrict__ R,
                                                      const int64_t* __restrict__ offsets,
                                                      int64_t nmol, int size, int64_t ldo,
                                                      double* __restrict__ out,
                                                      int32_t* __restrict__ status) {
    __shared__ double sR[3 * CM_MAX_ATOMS];
    __shared__ double sZ[CM_MAX_ATOMS];
    __shared__ double sD[CM_MAX_ATOMS];
    __shared__ int sOrder[CM_MAX_ATOMS];

    const int lane = threadIdx.x;
    const int d = size * (size + 1) / 2;

    for (int64_t mol = blockIdx.x; mol < nmol; mol += gridDim.x) {
        const int64_t a0 = offsets[mol];
        const int64_t natoms64 = offsets[mol + 1] - a0;
        double* row = out + mol * ldo;
        if (natoms64 > size || natoms64 < 0) {
            for (int64_t p = lane; p < ldo; p += CM_WAVE) row[p] = 0.0;
            if (lane == 0 && status) *status = 1;
            continue;
        }
        const int n = (int)natoms64;
        for (int a = lane; a < n; a += CM_WAVE) {
            sR[3 * a + 0] = R[3 * (a0 + a) + 0];
            sR[3 * a + 1] = R[3 * (a0 + a) + 1];
            sR[3 * a + 2] = R[3 * (a0 + a) + 2];
            double z = Z[a0 + a];
            sZ[a] = z;
            sD[a] = cm_diag(z);
        }
        __syncthreads();

        if (SORTED) {
            // Row norms: lane owns atoms lane, lane+64, ...
            double rn[CM_PER_LANE];
#pragma unroll
            for (int q = 0; q < CM_PER_LANE; ++q) {
                int k = lane + q * CM_WAVE;
                double v = DBL_MAX;
                if (k < n) {
                    double acc = 0.0;
                    for (int j = 0; j < n; ++j) {
                        if (j == k) continue;
                        double p = (j < k) ? cm_pair(sR, sZ, j, k) : cm_pair(sR, sZ, k, j);
                        double p2 = p * p;
                        acc = acc + p2;
                    }
                    double dk = sD[k];
                    double d2 = dk * dk;
                    v = d2 + acc;
                }
                rn[q] = v;
            }
            // Repeated minloc over all n entries (picked ones hold DBL_MAX), written from
            // the back: sorted_atoms(natoms - i + 1) = minloc(row_norms).
            for (int it = 0; it < n; ++it) {
                double bv = DBL_MAX;
                int bi = 0x7fffffff;
#pragma unroll
                for (int q = 0; q < CM_PER_LANE; ++q) {
                    int k = lane + q * CM_WAVE;
                    if (k < n && cm_better(rn[q], k, bv, bi)) {
                        bv = rn[q];
                        bi = k;
                    }
                }
#pragma unroll
                for (int off = CM_WAVE / 2; off > 0; off >>= 1) {
                    double ov = __shfl_xor(bv, off, CM_WAVE);
                    int oi = __shfl_xor(bi, off, CM_WAVE);
                    if (cm_better(ov, oi, bv, bi)) {
                        bv = ov;
                        bi = oi;
                    }
                }
                // Every lane now holds the same winner.
                if (lane == 0) sOrder[n - 1 - it] = bi;
#pragma unroll
                for (int q = 0; q < CM_PER_LANE; ++q)
                    if (lane + q * CM_WAVE == bi) rn[q] = DBL_MAX;
            }
            __syncthreads();
        }

        for (int p = lane; p < ldo; p += CM_WAVE) {
            double v = 0.0;
            if (p < d) {
                int m, c;
                cm_unpack(p, m, c);
                if (m < n) {
                    int i = SORTED ? sOrder[m] : m;
                    int j = SORTED ? sOrder[c] : c;
                    if (i == j) {
                        v = sD[i];
                    } else {
                        v = (i < j) ? cm_pair(sR, sZ, i, j) : cm_pair(sR, sZ, j, i);
                    }
                }
            }
            row[p] = v;
        }
        __syncthreads();
    }
}

}  // namespace qmlh

extern "C" int qmlh_cm_batch(const double* Z, const double* R, const int64_t* offsets,
                             int64_t nmol, int32_t size, int32_t sorted, double* out,
                             int64_t ldo, int32_t* status, qmlh_stream_t stream) {
    using namespace qmlh;
    if (nmol < 0) return set_error(QMLH_ERR_ARG, "cm_batch: nmol < 0");
    if (size < 1 || size > QMLH_CM_MAX_SIZE)
        return set_error(QMLH_ERR_ARG, "cm_batch: size %d outside [1, %d]", size,
                         QMLH_CM_MAX_SIZE);
    int64_t d = (int64_t)size * (size + 1) / 2;
    if (ldo < d) return set_error(QMLH_ERR_ARG, "cm_batch: ldo %lld < %lld", (long long)ldo,
                                  (long long)d);
    if (nmol == 0) return QMLH_OK;
    if (!Z || !R || !offsets || !out) return set_error(QMLH_ERR_ARG, "cm_batch: null pointer");
    int64_t grid = nmol < 262144 ? nmol : 262144;
    hipStream_t s = (hipStream_t)stream;
    TimerScope ts(QMLH_TIMER_CM, (double)nmol, s);
    if (sorted)
        hipLaunchKernelGGL(cm_kernel<true>, dim3((unsigned)grid), dim3(CM_WAVE), 0, s, Z, R,
                           offsets, nmol, (int)size, ldo, out, status);
    else
        hipLaunchKernelGGL(cm_kernel<false>, dim3((unsigned)grid), dim3(CM_WAVE), 0, s, Z, R,
                           offsets, nmol, (int)size, ldo, out, status);
    QMLH_CHECK_LAUNCH("cm_batch launch");
    return QMLH_OK;
}

// ------------------------------------------------------------------------------------------
// Atomic (local) Coulomb matrices: one 64-lane wavefront per output row (central atom).
//
// Reproduces fgenerate_local_coulomb_matrix (qml/frepresentations.f90:207-413, row-norm
// ordering) and fgenerate_atomic_coulomb_matrix (qml/frepresentations.f90:415-598, distance
// ordering) for one central atom k of one molecule:
//  * distances |R_b - R_a| for a < b summed x, y, z in order (lines 279-285), so the matrix
//    of distances is exactly symmetric;
//  * cutoff_count(k) = 1 + #{x != k : d(x, k) < central cut-off} rows, an error beyond `size`;
//  * row-norm ordering: M_ij(k) for i < j is ((f_ik Z_i) Z_j / d_ij) times the interaction
//    decay and then the central decay of j, each a separate rounding (lines 311-351); the
//    diagonal is ((f_ik f_ik) 0.5) Z_i^2.4; row norms are sums of squares in ascending j (the
//    one-thread order of the reduction at lines 311-351); the central atom is picked first,
//    then repeated maxloc with the picked entry set to zero (lines 355-373);
//  * distance ordering: repeated minloc over d(., k), picked entries set to huge (lines
//    535-546); entry (m, c) = (f_{s_m k} P(s_m, s_c)) times the central decay of s_c, with P
//    the interaction-cut pair matrix (lines 510-529 and 556-586);
//  * decay factor 0.5 (cos(pi ((r - c) + w) / w) + 1) with the device cos (libm's and the
//    device's cos may differ in the last place; every other operation is IEEE-exact and
//    matches the Fortran bit for bit).
// Atoms of the molecule are staged in LDS; a lane owns atoms lane, lane + 64, ...
// ------------------------------------------------------------------------------------------
namespace qmlh {

static constexpr int ACM_MAX_ATOMS = QMLH_ACM_MAX_ATOMS;
static constexpr int ACM_PER_LANE = ACM_MAX_ATOMS / CM_WAVE;

struct AcmParams {
    double cc, cd, ic, id;  // normalised central / interaction cut-off and decay
};

__device__ __forceinline__ double acm_decay(double r, double c, double w) {
    const double pi = 3.141592653589793;  // 4 atan(1) in FP64
    return 0.5 * (cos(pi * (r - c + w) / w) + 1.0);
}

// |R_b - R_a| with the smaller index first.
__device__ __forceinline__ double acm_dist(const double* sR, int a, int b) {
    if (a == b) return 0.0;
    int lo = a < b ? a : b, hi = a < b ? b : a;
    double dx = sR[3 * hi + 0] - sR[3 * lo + 0];
    double dy = sR[3 * hi + 1] - sR[3 * lo + 1];
    double dz = sR[3 * hi + 2] - sR[3 * lo + 2];
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return sqrt(s);
}

// Row-norm ordering's matrix entry M(a, b) for central atom with distances sDk.
__device__ __forceinline__ double acm_rn_entry(const double* sR, const double* sZ,
                                               const double* sDk, int a, int b,
                                               const AcmParams& p) {
    int i = a < b ? a : b, j = a < b ? b : a;
    double di = sDk[i];
    if (di > p.cc) return 0.0;
    double pref = 1.0;
    if (di > p.cc - p.cd) pref = acm_decay(di, p.cc, p.cd);
    if (i == j) return pref * pref * 0.5 * (2.0 * cm_diag(sZ[i]));
    double dj = sDk[j];
    if (dj > p.cc) return 0.0;
    double dij = acm_dist(sR, i, j);
    if (dij > p.ic) return 0.0;
    double v = pref * sZ[i] * sZ[j] / dij;
    if (dij > p.ic - p.id) v = v * 0.5 * (cos(3.141592653589793 * (dij - p.ic + p.id) / p.id) + 1);
    if (dj > p.cc - p.cd) v = v * 0.5 * (cos(3.141592653589793 * (dj - p.cc + p.cd) / p.cd) + 1);
    return v;
}

// Distance ordering's interaction-cut pair matrix P(a, b).
__device__ __forceinline__ double acm_pair(const double* sR, const double* sZ, int a, int b,
                                           const AcmParams& p) {
    if (a == b) return cm_diag(sZ[a]);
    int i = a < b ? a : b, j = a < b ? b : a;
    double dij = acm_dist(sR, i, j);
    if (dij > p.ic) return 0.0;
    double v = sZ[i] * sZ[j] / dij;
    if (dij > p.ic - p.id) v = v * 0.5 * (cos(3.141592653589793 * (dij - p.ic + p.id) / p.id) + 1);
    return v;
}

// status bits: 1 a central atom needs more than `size` rows; 2 a molecule has more than
// QMLH_ACM_MAX_ATOMS atoms; 4 a central atom or molecule index is out of range. The row of
// such a case is written as zeros.
template <bool ROWNORM>
__global__ __launch_bounds__(CM_WAVE) void atomic_cm_kernel(
    const double* __restrict__ Z, const double* __restrict__ R,
    const int64_t* __restrict__ offsets, int64_t nmol, const int64_t* __restrict__ central,
    const int64_t* __restrict__ cmol, int64_t nrows, int size, AcmParams p, int64_t ldo,
    double* __restrict__ out, int32_t* __restrict__ status) {
    __shared__ double sR[3 * ACM_MAX_ATOMS];
    __shared__ double sZ[ACM_MAX_ATOMS];
    __shared__ double sDk[ACM_MAX_ATOMS];
    __shared__ int sOrder[ACM_MAX_ATOMS];

    const int lane = threadIdx.x;
    const int d = size * (size + 1) / 2;

    for (int64_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        double* dst = out + row * ldo;
        const int64_t mol = cmol[row];
        int bad = 0;
        int64_t a0 = 0, n64 = 0, k64 = central[row];
        if (mol < 0 || mol >= nmol) {
            bad = 4;
        } else {
            a0 = offsets[mol];
            n64 = offsets[mol + 1] - a0;
            if (n64 > ACM_MAX_ATOMS || n64 < 0) bad = 2;
            else if (k64 < 0 || k64 >= n64) bad = 4;
        }
        if (bad) {
            for (int64_t q = lane; q < ldo; q += CM_WAVE) dst[q] = 0.0;
            if (lane == 0 && status) atomicOr(status, bad);
            continue;
        }
        const int n = (int)n64, k = (int)k64;
        for (int a = lane; a < n; a += CM_WAVE) {
            sR[3 * a + 0] = R[3 * (a0 + a) + 0];
            sR[3 * a + 1] = R[3 * (a0 + a) + 1];
            sR[3 * a + 2] = R[3 * (a0 + a) + 2];
            sZ[a] = Z[a0 + a];
        }
        __syncthreads();
        int cnt_lane = 0;
        for (int a = lane; a < n; a += CM_WAVE) {
            double dk = acm_dist(sR, a, k);
            sDk[a] = dk;
            if (a != k && dk < p.cc) ++cnt_lane;
        }
#pragma unroll
        for (int off = CM_WAVE / 2; off > 0; off >>= 1) cnt_lane += __shfl_xor(cnt_lane, off, CM_WAVE);
        const int cnt = 1 + cnt_lane;
        __syncthreads();
        if (cnt > size) {
            for (int64_t q = lane; q < ldo; q += CM_WAVE) dst[q] = 0.0;
            if (lane == 0 && status) atomicOr(status, 1);
            __syncthreads();
            continue;
        }

        // Ordering key per owned atom: row norm (picked -> 0, maximum first) or distance
        // (picked -> huge, minimum first).
        double key[ACM_PER_LANE];
#pragma unroll
        for (int q = 0; q < ACM_PER_LANE; ++q) {
            const int x = lane + q * CM_WAVE;
            double v = 0.0;
            if (x < n) {
                if (ROWNORM) {
                    if (x == k) {
                        v = DBL_MAX;
                    } else {
                        double acc = 0.0;
                        for (int j = 0; j < n; ++j) {
                            double e = acm_rn_entry(sR, sZ, sDk, x, j, p);
                            double e2 = e * e;
                            acc = acc + e2;
                        }
                        v = acc;
                    }
                } else {
                    v = sDk[x];
                }
            }
            key[q] = v;
        }
        for (int it = 0; it < cnt; ++it) {
            double bv = ROWNORM ? -1.0 : DBL_MAX;
            int bi = 0x7fffffff;
#pragma unroll
            for (int q = 0; q < ACM_PER_LANE; ++q) {
                const int x = lane + q * CM_WAVE;
                if (x < n) {
                    bool better = ROWNORM ? (key[q] > bv || (key[q] == bv && x < bi))
                                          : (key[q] < bv || (key[q] == bv && x < bi));
                    if (better) {
                        bv = key[q];
                        bi = x;
                    }
                }
            }
#pragma unroll
            for (int off = CM_WAVE / 2; off > 0; off >>= 1) {
                double ov = __shfl_xor(bv, off, CM_WAVE);
                int oi = __shfl_xor(bi, off, CM_WAVE);
                bool better = ROWNORM ? (ov > bv || (ov == bv && oi < bi))
                                      : (ov < bv || (ov == bv && oi < bi));
                if (better) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (bi >= n) bi = 0;  // only with NaN keys (NaN coordinates): keep indices valid
            if (lane == 0) sOrder[it] = bi;
#pragma unroll
            for (int q = 0; q < ACM_PER_LANE; ++q)
                if (lane + q * CM_WAVE == bi) key[q] = ROWNORM ? 0.0 : DBL_MAX;
        }
        __syncthreads();

        for (int64_t q = lane; q < ldo; q += CM_WAVE) {
            double v = 0.0;
            if (q < d) {
                int m, c;
                cm_unpack((int)q, m, c);
                if (m < cnt) {
                    const int i = sOrder[m], j = sOrder[c];
                    if (ROWNORM) {
                        v = acm_rn_entry(sR, sZ, sDk, i, j, p);
                    } else {
                        const double di = sDk[i];
                        if (!(di > p.cc)) {
                            double pref = 1.0;
                            if (di > p.cc - p.cd) pref = acm_decay(di, p.cc, p.cd);
                            v = pref * acm_pair(sR, sZ, i, j, p);
                            const double dj = sDk[j];
                            if (dj > p.cc - p.cd)
                                v = v * 0.5 * (cos(3.141592653589793 * (dj - p.cc + p.cd) / p.cd) + 1);
                        }
                    }
                }
            }
            dst[q] = v;
        }
        __syncthreads();
    }
}

// Molecules of more than ACM_MAX_ATOMS atoms: the same rows as atomic_cm_kernel, with the
// molecule, the distances to the central atom and the ordering keys held in this workgroup's
// slice of global scratch instead of LDS and registers (one wavefront per row; the keys of
// picked atoms are overwritten in place, as the selection above does in registers). Rows of
// smaller molecules are left to atomic_cm_kernel.
__host__ __device__ inline int64_t acm_large_stride(int64_t max_atoms, int size) {
    return 6 * max_atoms + size;  // R (3 n), Z, distances, keys (n each), order (size)
}

template <bool ROWNORM>
__global__ __launch_bounds__(CM_WAVE) void atomic_cm_large_kernel(
    const double* __restrict__ Z, const double* __restrict__ R,
    const int64_t* __restrict__ offsets, int64_t nmol, const int64_t* __restrict__ central,
    const int64_t* __restrict__ cmol, int64_t nrows, int size, AcmParams p, int64_t ldo,
    double* __restrict__ out, int32_t* __restrict__ status, int64_t max_atoms,
    double* __restrict__ ws) {
    const int lane = threadIdx.x;
    const int d = size * (size + 1) / 2;
    double* gR = ws + (int64_t)blockIdx.x * acm_large_stride(max_atoms, size);
    double* gZ = gR + 3 * max_atoms;
    double* gDk = gZ + max_atoms;
    double* gKey = gDk + max_atoms;
    double* gOrd = gKey + max_atoms;  // atom indices, stored as doubles

    for (int64_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        const int64_t mol = cmol[row];
        if (mol < 0 || mol >= nmol) continue;  // reported by atomic_cm_kernel
        const int64_t a0 = offsets[mol];
        const int64_t n = offsets[mol + 1] - a0;
        if (n <= ACM_MAX_ATOMS) continue;  // atomic_cm_kernel's row
        double* dst = out + row * ldo;
        const int64_t k = central[row];
        if (n > max_atoms || k < 0 || k >= n) {
            for (int64_t q = lane; q < ldo; q += CM_WAVE) dst[q] = 0.0;
            if (lane == 0 && status) atomicOr(status, n > max_atoms ? 2 : 4);
            continue;
        }
        for (int64_t a = lane; a < n; a += CM_WAVE) {
            gR[3 * a + 0] = R[3 * (a0 + a) + 0];
            gR[3 * a + 1] = R[3 * (a0 + a) + 1];
            gR[3 * a + 2] = R[3 * (a0 + a) + 2];
            gZ[a] = Z[a0 + a];
        }
        __syncthreads();
        int cnt_lane = 0;
        for (int64_t a = lane; a < n; a += CM_WAVE) {
            const double dk = acm_dist(gR, (int)a, (int)k);
            gDk[a] = dk;
            if (a != k && dk < p.cc) ++cnt_lane;
        }
#pragma unroll
        for (int off = CM_WAVE / 2; off > 0; off >>= 1) cnt_lane += __shfl_xor(cnt_lane, off, CM_WAVE);
        const int cnt = 1 + cnt_lane;
        __syncthreads();
        if (cnt > size) {
            for (int64_t q = lane; q < ldo; q += CM_WAVE) dst[q] = 0.0;
            if (lane == 0 && status) atomicOr(status, 1);
            __syncthreads();
            continue;
        }
        // ordering keys: row norm (picked -> 0, maximum first) or distance (picked -> huge)
        for (int64_t x = lane; x < n; x += CM_WAVE) {
            double v;
            if (ROWNORM) {
                if (x == k) {
                    v = DBL_MAX;
                } else {
                    double acc = 0.0;
                    for (int64_t j = 0; j < n; ++j) {
                        const double e = acm_rn_entry(gR, gZ, gDk, (int)x, (int)j, p);
                        const double e2 = e * e;
                        acc = acc + e2;
                    }
                    v = acc;
                }
            } else {
                v = gDk[x];
            }
            gKey[x] = v;
        }
        __syncthreads();
        for (int it = 0; it < cnt; ++it) {
            double bv = ROWNORM ? -1.0 : DBL_MAX;
            int64_t bi = INT64_MAX;
            for (int64_t x = lane; x < n; x += CM_WAVE) {
                const double kv = gKey[x];
                const bool better = ROWNORM ? (kv > bv || (kv == bv && x < bi))
                                            : (kv < bv || (kv == bv && x < bi));
                if (better) {
                    bv = kv;
                    bi = x;
                }
            }
#pragma unroll
            for (int off = CM_WAVE / 2; off > 0; off >>= 1) {
                const double ov = __shfl_xor(bv, off, CM_WAVE);
                const int64_t oi = __shfl_xor(bi, off, CM_WAVE);
                const bool better = ROWNORM ? (ov > bv || (ov == bv && oi < bi))
                                            : (ov < bv || (ov == bv && oi < bi));
                if (better) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (bi >= n) bi = 0;  // only with NaN keys (NaN coordinates): keep indices valid
            if (lane == 0) {
                gOrd[it] = (double)bi;
                gKey[bi] = ROWNORM ? 0.0 : DBL_MAX;
            }
            __syncthreads();  // the picked key is visible before the next scan
        }
        for (int64_t q = lane; q < ldo; q += CM_WAVE) {
            double v = 0.0;
            if (q < d) {
                int m, c;
                cm_unpack((int)q, m, c);
                if (m < cnt) {
                    const int i = (int)gOrd[m], j = (int)gOrd[c];
                    if (ROWNORM) {
                        v = acm_rn_entry(gR, gZ, gDk, i, j, p);
                    } else {
                        const double di = gDk[i];
                        if (!(di > p.cc)) {
                            double pref = 1.0;
                            if (di > p.cc - p.cd) pref = acm_decay(di, p.cc, p.cd);
                            v = pref * acm_pair(gR, gZ, i, j, p);
                            const double dj = gDk[j];
                            if (dj > p.cc - p.cd)
                                v = v * 0.5 * (cos(3.141592653589793 * (dj - p.cc + p.cd) / p.cd) + 1);
                        }
                    }
                }
            }
            dst[q] = v;
        }
        __syncthreads();  // this row's scratch reads are done before the next row's stores
    }
}

static constexpr int64_t ACM_LARGE_GRID = 1024;

static AcmParams acm_params(double cc, double cd, double ic, double id) {
    // Cut-off normalisation of qml/frepresentations.f90:254-273 (identical at 460-479).
    AcmParams p{cc, cd, ic, id};
    if (p.cc < 0) p.cc = DBL_MAX;
    if (p.ic < 0 || p.ic > 2 * p.cc) p.ic = 2 * p.cc;
    if (p.cd < 0) p.cd = 0.0;
    else if (p.cd > p.cc) p.cd = p.cc;
    if (p.id < 0) p.id = 0.0;
    else if (p.id > p.ic) p.id = p.ic;
    return p;
}

}  // namespace qmlh

extern "C" int qmlh_atomic_cm_batch(const double* Z, const double* R, const int64_t* offsets,
                                    int64_t nmol, const int64_t* central, const int64_t* cmol,
                                    int64_t nrows, int32_t size, int32_t sorting,
                                    double central_cutoff, double central_decay,
                                    double interaction_cutoff, double interaction_decay,
                                    double* out, int64_t ldo, int32_t* status,
                                    qmlh_stream_t stream) {
    using namespace qmlh;
    if (nrows < 0 || nmol < 0) return set_error(QMLH_ERR_ARG, "atomic_cm_batch: negative count");
    if (size < 1 || size > QMLH_CM_MAX_SIZE)
        return set_error(QMLH_ERR_ARG, "atomic_cm_batch: size %d outside [1, %d]", size,
                         QMLH_CM_MAX_SIZE);
    if (sorting != QMLH_ACM_ROW_NORM && sorting != QMLH_ACM_DISTANCE)
        return set_error(QMLH_ERR_ARG, "atomic_cm_batch: unknown sorting %d", sorting);
    int64_t d = (int64_t)size * (size + 1) / 2;
    if (ldo < d) return set_error(QMLH_ERR_ARG, "atomic_cm_batch: ldo %lld < %lld",
                                  (long long)ldo, (long long)d);
    if (nrows == 0) return QMLH_OK;
    if (!Z || !R || !offsets || !central || !cmol || !out)
        return set_error(QMLH_ERR_ARG, "atomic_cm_batch: null pointer");
    const AcmParams p = acm_params(central_cutoff, central_decay, interaction_cutoff,
                                   interaction_decay);
    int64_t grid = nrows < 262144 ? nrows : 262144;
    hipStream_t s = (hipStream_t)stream;
    TimerScope ts(QMLH_TIMER_CM, (double)nrows, s);
    if (sorting == QMLH_ACM_ROW_NORM)
        hipLaunchKernelGGL(atomic_cm_kernel<true>, dim3((unsigned)grid), dim3(CM_WAVE), 0, s, Z,
                           R, offsets, nmol, central, cmol, nrows, (int)size, p, ldo, out, status);
    else
        hipLaunchKernelGGL(atomic_cm_kernel<false>, dim3((unsigned)grid), dim3(CM_WAVE), 0, s,
                           Z, R, offsets, nmol, central, cmol, nrows, (int)size, p, ldo, out,
                           status);
    QMLH_CHECK_LAUNCH("atomic_cm_batch launch");
    return QMLH_OK;
}

extern "C" size_t qmlh_atomic_cm_large_workspace(int64_t max_atoms, int32_t size) {
    if (max_atoms <= 0 || size < 1) return 0;
    return (size_t)qmlh::ACM_LARGE_GRID * (size_t)qmlh::acm_large_stride(max_atoms, size) *
           sizeof(double);
}

extern "C" int qmlh_atomic_cm_large(const double* Z, const double* R, const int64_t* offsets,
                                    int64_t nmol, const int64_t* central, const int64_t* cmol,
                                    int64_t nrows, int32_t size, int32_t sorting,
                                    double central_cutoff, double central_decay,
                                    double interaction_cutoff, double interaction_decay,
                                    double* out, int64_t ldo, int32_t* status, int64_t max_atoms,
                                    void* ws, size_t ws_bytes, qmlh_stream_t stream) {
    using namespace qmlh;
    if (nrows < 0 || nmol < 0 || max_atoms < 0)
        return set_error(QMLH_ERR_ARG, "atomic_cm_large: negative count");
    if (size < 1 || size > QMLH_CM_MAX_SIZE)
        return set_error(QMLH_ERR_ARG, "atomic_cm_large: size %d outside [1, %d]", size,
                         QMLH_CM_MAX_SIZE);
    if (sorting != QMLH_ACM_ROW_NORM && sorting != QMLH_ACM_DISTANCE)
        return set_error(QMLH_ERR_ARG, "atomic_cm_large: unknown sorting %d", sorting);
    const int64_t d = (int64_t)size * (size + 1) / 2;
    if (ldo < d) return set_error(QMLH_ERR_ARG, "atomic_cm_large: ldo < size (size + 1) / 2");
    if (nrows == 0 || max_atoms <= QMLH_ACM_MAX_ATOMS) return QMLH_OK;
    if (max_atoms > (int64_t)1 << 30) return set_error(QMLH_ERR_ARG, "atomic_cm_large: max_atoms too large");
    if (!Z || !R || !offsets || !central || !cmol || !out || !ws)
        return set_error(QMLH_ERR_ARG, "atomic_cm_large: null pointer");
    if (ws_bytes < qmlh_atomic_cm_large_workspace(max_atoms, size))
        return set_error(QMLH_ERR_WORKSPACE, "atomic_cm_large: workspace too small");
    const AcmParams p = acm_params(central_cutoff, central_decay, interaction_cutoff,
                                   interaction_decay);
    const int64_t grid = nrows < ACM_LARGE_GRID ? nrows : ACM_LARGE_GRID;
    hipStream_t s = (hipStream_t)stream;
    double* w = static_cast<double*>(ws);
    if (sorting == QMLH_ACM_ROW_NORM)
        hipLaunchKernelGGL(atomic_cm_large_kernel<true>, dim3((unsigned)grid), dim3(CM_WAVE), 0, s,
                           Z, R, offsets, nmol, central, cmol, nrows, (int)size, p, ldo, out,
                           status, max_atoms, w);
    else
        hipLaunchKernelGGL(atomic_cm_large_kernel<false>, dim3((unsigned)grid), dim3(CM_WAVE), 0,
                           s, Z, R, offsets, nmol, central, cmol, nrows, (int)size, p, ldo, out,
                           status, max_atoms, w);
    QMLH_CHECK_LAUNCH("atomic_cm_large launch");
    return QMLH_OK;
}
