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
