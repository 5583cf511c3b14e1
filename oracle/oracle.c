/*
 * CPU oracle for the KRR hot path: a plain-C restatement of the reference Fortran loops.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product path (qml_amd/) may load or call this
 * library; it is the checker the HIP kernels are compared against, and the CPU baseline that
 * bench.py times beside them.
 *
 * Built with -ffp-contract=off so every a*b+c is two roundings, exactly as the Fortran is
 * evaluated on an x86-64 target without FMA. Each function cites the reference loop it follows.
 *
 * Conventions: representations are row-major (n, d) float64 arrays (numpy C order); kernel and
 * distance matrices are written column-major (na, nb), i.e. element (i, j) at K[i + j*na],
 * which is the Fortran-ordered array the reference returns (qml/kernels.py:58).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* Pair entry Z_i Z_j / |R_j - R_i| for i < j, summed x, y, z in order
 * (qml/frepresentations.f90:104-105). */
static double pair_entry(const double *Z, const double *R, int64_t i, int64_t j) {
    double dx = R[3 * j + 0] - R[3 * i + 0];
    double dy = R[3 * j + 1] - R[3 * i + 1];
    double dz = R[3 * j + 2] - R[3 * i + 2];
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return Z[i] * Z[j] / sqrt(s);
}

/* Sorted (row-norm) Coulomb matrix of one molecule: qml/frepresentations.f90:52-140.
 * Row norms use the serial evaluation order of the two OpenMP reduction loops at one thread:
 * rn_k = d_k*d_k + (sum over j != k, ascending j, of M_kj^2), the sum starting from zero
 * (the reduction's private copy) and added to the diagonal term at the end (lines 91-114).
 * The order is the repeated minloc of lines 116-121 written from the back.
 * Returns 0, or -1 when natoms > size (the reference writes out of bounds there). */
int oracle_cm_sorted(const double *Z, const double *R, int64_t natoms, int64_t size, double *cm) {
    int64_t d = size * (size + 1) / 2;
    if (natoms > size) return -1;
    double *pd = (double *)malloc(sizeof(double) * (size_t)(natoms * natoms + 1));
    double *rn = (double *)malloc(sizeof(double) * (size_t)(natoms + 1));
    double *acc = (double *)malloc(sizeof(double) * (size_t)(natoms + 1));
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(natoms + 1));
    for (int64_t i = 0; i < natoms; ++i) {
        double p = 0.5 * pow(Z[i], 2.4);
        rn[i] = p * p; /* 0 + p*p, exact */
        pd[i * natoms + i] = p;
        acc[i] = 0.0;
    }
    for (int64_t i = 0; i < natoms; ++i) {
        for (int64_t j = i + 1; j < natoms; ++j) {
            double p = pair_entry(Z, R, i, j);
            pd[i * natoms + j] = p;
            pd[j * natoms + i] = p;
            double p2 = p * p;
            acc[j] = acc[j] + p2;
            acc[i] = acc[i] + p2;
        }
    }
    for (int64_t i = 0; i < natoms; ++i) rn[i] = rn[i] + acc[i];
    const double huge = DBL_MAX;
    for (int64_t it = 0; it < natoms; ++it) {
        int64_t best = 0;
        for (int64_t j = 1; j < natoms; ++j)
            if (rn[j] < rn[best]) best = j;
        order[natoms - 1 - it] = best;
        rn[best] = huge;
    }
    memset(cm, 0, sizeof(double) * (size_t)d);
    for (int64_t m = 0; m < natoms; ++m) {
        int64_t i = order[m];
        int64_t base = m * (m + 1) / 2;
        for (int64_t n = 0; n <= m; ++n) cm[base + n] = pd[i * natoms + order[n]];
    }
    free(pd); free(rn); free(acc); free(order);
    return 0;
}

/* Unsorted Coulomb matrix: qml/frepresentations.f90:142-205. */
int oracle_cm_unsorted(const double *Z, const double *R, int64_t natoms, int64_t size, double *cm) {
    int64_t d = size * (size + 1) / 2;
    if (natoms > size) return -1;
    memset(cm, 0, sizeof(double) * (size_t)d);
    for (int64_t m = 0; m < natoms; ++m) {
        int64_t base = m * (m + 1) / 2;
        for (int64_t n = 0; n < m; ++n) cm[base + n] = pair_entry(Z, R, n, m);
        cm[base + m] = 0.5 * pow(Z[m], 2.4);
    }
    return 0;
}

/* Batch of molecules in ragged layout; molecule m owns atoms offsets[m]..offsets[m+1]-1.
 * Output row m is cm[m*d : (m+1)*d]. OpenMP over molecules (each molecule serial). */
int oracle_cm_batch(const double *Z, const double *R, const int64_t *offsets, int64_t nmol,
                    int64_t size, int sorted, double *out) {
    int64_t d = size * (size + 1) / 2;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
    for (int64_t m = 0; m < nmol; ++m) {
        int64_t a0 = offsets[m], n = offsets[m + 1] - offsets[m];
        int rc = sorted ? oracle_cm_sorted(Z + a0, R + 3 * a0, n, size, out + m * d)
                        : oracle_cm_unsorted(Z + a0, R + 3 * a0, n, size, out + m * d);
        if (rc) bad |= 1;
    }
    return bad ? -1 : 0;
}

/* Laplacian kernel: qml/fkernels.f90:374-400. inv = -1/sigma; K(j,i) = exp(inv * sum|a-b|),
 * the sum serial over features; OpenMP over the columns i = 1..nb as the reference. */
void oracle_laplacian(const double *A, int64_t na, const double *B, int64_t nb, int64_t d,
                      double sigma, double *K) {
    double inv = -1.0 / sigma;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i) {
        const double *b = B + i * d;
        for (int64_t j = 0; j < na; ++j) {
            const double *a = A + j * d;
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) s = s + fabs(a[k] - b[k]);
            K[j + i * na] = exp(inv * s);
        }
    }
}

/* Gaussian kernel: qml/fkernels.f90:340-372. inv = -0.5/(sigma*sigma); direct differences. */
void oracle_gaussian(const double *A, int64_t na, const double *B, int64_t nb, int64_t d,
                     double sigma, double *K) {
    double inv = -0.5 / (sigma * sigma);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i) {
        const double *b = B + i * d;
        for (int64_t j = 0; j < na; ++j) {
            const double *a = A + j * d;
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) {
                double t = a[k] - b[k];
                s = s + t * t;
            }
            K[j + i * na] = exp(inv * s);
        }
    }
}

/* Manhattan distance: qml/fdistance.f90:23-46. */
void oracle_manhattan(const double *A, int64_t na, const double *B, int64_t nb, int64_t d,
                      double *D) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i) {
        const double *b = B + i * d;
        for (int64_t j = 0; j < na; ++j) {
            const double *a = A + j * d;
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) s = s + fabs(a[k] - b[k]);
            D[j + i * na] = s;
        }
    }
}

/* L2 distance: qml/fdistance.f90:49-80. */
void oracle_l2(const double *A, int64_t na, const double *B, int64_t nb, int64_t d, double *D) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i) {
        const double *b = B + i * d;
        for (int64_t j = 0; j < na; ++j) {
            const double *a = A + j * d;
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) {
                double t = a[k] - b[k];
                s = s + t * t;
            }
            D[j + i * na] = sqrt(s);
        }
    }
}
