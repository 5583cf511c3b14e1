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

/* ------------------------------------------------------------------------------------------
 * Atomic (local) Coulomb matrices: qml/frepresentations.f90:207-413 (row-norm ordering,
 * fgenerate_local_coulomb_matrix) and 415-598 (distance ordering,
 * fgenerate_atomic_coulomb_matrix), one central atom at a time, one thread.
 * ------------------------------------------------------------------------------------------ */

/* Cut-off parameter normalisation shared by both routines (lines 254-273 / 460-479). */
void oracle_acm_params(double *cc, double *cd, double *ic, double *id) {
    const double huge = DBL_MAX;
    if (*cc < 0) *cc = huge;
    if ((*ic < 0) || (*ic > 2 * *cc)) *ic = 2 * *cc;
    if (*cd < 0) *cd = 0.0;
    else if (*cd > *cc) *cd = *cc;
    if (*id < 0) *id = 0.0;
    else if (*id > *ic) *id = *ic;
}

/* |R_b - R_a| for a < b, coordinates summed x, y, z in order (lines 279-285). */
static double acm_dist(const double *R, int64_t a, int64_t b) {
    if (a == b) return 0.0;
    if (a > b) { int64_t t = a; a = b; b = t; }
    double dx = R[3 * b + 0] - R[3 * a + 0];
    double dy = R[3 * b + 1] - R[3 * a + 1];
    double dz = R[3 * b + 2] - R[3 * a + 2];
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return sqrt(s);
}

/* 0.5 (cos(pi (r - c + w) / w) + 1), the decay factor of lines 321-322 / 565-567. */
static double acm_decay(double r, double c, double w) {
    const double pi = 4.0 * atan(1.0);
    return 0.5 * (cos(pi * (r - c + w) / w) + 1.0);
}

/* Row-norm ordered matrix of central atom k (0-based) into cm[d]; returns 0, or the needed
 * size when the atoms within the central cut-off exceed `size` (the reference stops there). */
static int64_t acm_rownorm_one(const double *Z, const double *R, int64_t n, int64_t k,
                               int64_t size, double cc, double cd, double ic, double id,
                               double *cm, double *P, double *rn, double *dk, int64_t *order) {
    int64_t cnt = 1;
    for (int64_t x = 0; x < n; ++x) {
        dk[x] = acm_dist(R, x, k);
        if (x != k && dk[x] < cc) ++cnt;
    }
    if (cnt > size) return cnt;
    for (int64_t x = 0; x < n * n; ++x) P[x] = 0.0;
    for (int64_t x = 0; x < n; ++x) rn[x] = 0.0;
    /* lines 311-351: rows i inside the central cut-off; pair (i, j > i) inside both cut-offs. */
    for (int64_t i = 0; i < n; ++i) {
        if (dk[i] > cc) continue;
        double pref = 1.0;
        if (dk[i] > cc - cd) pref = acm_decay(dk[i], cc, cd);
        double pn = pref * pref * 0.5 * pow(Z[i], 2.4);
        P[i * n + i] = pn;
        rn[i] = rn[i] + pn * pn;
        for (int64_t j = i + 1; j < n; ++j) {
            if (dk[j] > cc) continue;
            double dij = acm_dist(R, i, j);
            if (dij > ic) continue;
            double p = pref * Z[i] * Z[j] / dij;
            if (dij > ic - id) p = p * 0.5 * (cos(4.0 * atan(1.0) * (dij - ic + id) / id) + 1);
            if (dk[j] > cc - cd) p = p * 0.5 * (cos(4.0 * atan(1.0) * (dk[j] - cc + cd) / cd) + 1);
            P[i * n + j] = p;
            P[j * n + i] = p;
            double p2 = p * p;
            rn[i] = rn[i] + p2;
            rn[j] = rn[j] + p2;
        }
    }
    /* lines 355-373: central atom first, then repeated maxloc (first index of the maximum),
     * picked entries set to zero. */
    rn[k] = DBL_MAX;
    for (int64_t it = 0; it < cnt; ++it) {
        int64_t best = 0;
        for (int64_t x = 1; x < n; ++x)
            if (rn[x] > rn[best]) best = x;
        order[it] = best;
        rn[best] = 0.0;
    }
    /* lines 383-395: packed lower triangle of the reordered matrix. */
    memset(cm, 0, sizeof(double) * (size_t)(size * (size + 1) / 2));
    for (int64_t m = 0; m < cnt; ++m) {
        int64_t base = m * (m + 1) / 2;
        for (int64_t c = 0; c <= m; ++c) cm[base + c] = P[order[m] * n + order[c]];
    }
    return 0;
}

/* Distance ordered matrix of central atom k (lines 481-592). */
static int64_t acm_distance_one(const double *Z, const double *R, int64_t n, int64_t k,
                                int64_t size, double cc, double cd, double ic, double id,
                                double *cm, double *P, double *tmp, double *dk,
                                int64_t *order) {
    int64_t cnt = 1;
    for (int64_t x = 0; x < n; ++x) {
        dk[x] = acm_dist(R, x, k);
        if (x != k && dk[x] < cc) ++cnt;
    }
    if (cnt > size) return cnt;
    /* lines 510-529: pair matrix with the interaction cut-off only. */
    for (int64_t i = 0; i < n; ++i) {
        P[i * n + i] = 0.5 * pow(Z[i], 2.4);
        for (int64_t j = i + 1; j < n; ++j) {
            double dij = acm_dist(R, i, j);
            double p = 0.0;
            if (!(dij > ic)) {
                p = Z[i] * Z[j] / dij;
                if (dij > ic - id) p = p * 0.5 * (cos(4.0 * atan(1.0) * (dij - ic + id) / id) + 1);
            }
            P[i * n + j] = p;
            P[j * n + i] = p;
        }
    }
    /* lines 535-546: repeated minloc over the distances to k, picked entries set to huge. */
    for (int64_t x = 0; x < n; ++x) tmp[x] = dk[x];
    for (int64_t it = 0; it < cnt; ++it) {
        int64_t best = 0;
        for (int64_t x = 1; x < n; ++x)
            if (tmp[x] < tmp[best]) best = x;
        order[it] = best;
        tmp[best] = DBL_MAX;
    }
    /* lines 556-586: rows outside the central cut-off stay zero; entry (m, c) carries the
     * central decay of the row atom as a prefactor and that of the column atom last. */
    memset(cm, 0, sizeof(double) * (size_t)(size * (size + 1) / 2));
    for (int64_t m = 0; m < cnt; ++m) {
        int64_t i = order[m];
        if (dk[i] > cc) continue;
        double pref = 1.0;
        if (dk[i] > cc - cd) pref = acm_decay(dk[i], cc, cd);
        int64_t base = m * (m + 1) / 2;
        for (int64_t c = 0; c <= m; ++c) {
            int64_t j = order[c];
            double p = pref * P[i * n + j];
            if (dk[j] > cc - cd) p = p * 0.5 * (cos(4.0 * atan(1.0) * (dk[j] - cc + cd) / cd) + 1);
            cm[base + c] = p;
        }
    }
    return 0;
}

/* One molecule, `ncent` central atoms (0-based indices) -> out (ncent, size(size+1)/2).
 * sorting: 0 row-norm, 1 distance. Cut-offs are normalised as the reference does. Returns 0,
 * or the size the first offending central atom needs. */
int64_t oracle_atomic_cm(const double *Z, const double *R, int64_t n, const int64_t *cent,
                         int64_t ncent, int64_t size, int sorting, double cc, double cd,
                         double ic, double id, double *out) {
    oracle_acm_params(&cc, &cd, &ic, &id);
    int64_t d = size * (size + 1) / 2;
    double *P = (double *)malloc(sizeof(double) * (size_t)(n * n + 1));
    double *rn = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *dk = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t rc = 0;
    for (int64_t l = 0; l < ncent && rc == 0; ++l) {
        if (sorting == 0)
            rc = acm_rownorm_one(Z, R, n, cent[l], size, cc, cd, ic, id, out + l * d, P, rn, dk,
                                 order);
        else
            rc = acm_distance_one(Z, R, n, cent[l], size, cc, cd, ic, id, out + l * d, P, rn,
                                  dk, order);
    }
    free(P); free(rn); free(dk); free(order);
    return rc;
}

/* ------------------------------------------------------------------------------------------
 * Local (atom-summed) kernels: qml/fkernels.f90:23-106 (Gaussian) and 108-196 (Laplacian).
 *   K(s, a, b) = sum over atoms j of b (outer), i of a (inner) of exp(D(i, j) * inv_s)
 * with D the squared L2 (Gaussian, inv = -0.5/sigma^2) or L1 (Laplacian, inv = -1/sigma)
 * distance summed serially over the features, and the atom-pair sum in the column-major order
 * of the Fortran sum() over atomic_distance(:ni, :nj). Output Fortran-ordered
 * (nsig, nmA, nmB): element (s, a, b) at K[s + nsig (a + nmA b)].
 * kind: 0 Laplacian, 1 Gaussian. OpenMP over molecules of B.
 * ------------------------------------------------------------------------------------------ */
void oracle_local_kernels(int kind, const double *A, const int64_t *offA, int64_t nmA,
                          const double *B, const int64_t *offB, int64_t nmB, int64_t d,
                          const double *sigmas, int64_t nsig, double *K) {
    double *inv = (double *)malloc(sizeof(double) * (size_t)(nsig + 1));
    for (int64_t s = 0; s < nsig; ++s)
        inv[s] = kind == 1 ? -0.5 / (sigmas[s] * sigmas[s]) : -1.0 / sigmas[s];
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = 0; b < nmB; ++b) {
        for (int64_t a = 0; a < nmA; ++a) {
            double *k = K + nsig * (a + nmA * b);
            for (int64_t s = 0; s < nsig; ++s) k[s] = 0.0;
            for (int64_t j = offB[b]; j < offB[b + 1]; ++j) {
                const double *y = B + j * d;
                for (int64_t i = offA[a]; i < offA[a + 1]; ++i) {
                    const double *x = A + i * d;
                    double dist = 0.0;
                    if (kind == 1) {
                        for (int64_t f = 0; f < d; ++f) {
                            double t = x[f] - y[f];
                            dist = dist + t * t;
                        }
                    } else {
                        for (int64_t f = 0; f < d; ++f) dist = dist + fabs(x[f] - y[f]);
                    }
                    for (int64_t s = 0; s < nsig; ++s) k[s] = k[s] + exp(dist * inv[s]);
                }
            }
        }
    }
    free(inv);
}

/* ------------------------------------------------------------------------------------------
 * Other global kernels and the p-distance: qml/fkernels.f90:403-531, qml/fdistance.f90:83-155.
 * Same conventions as above (rows of A, B; K column-major (na, nb)).
 * ------------------------------------------------------------------------------------------ */

/* flinear_kernel (403-424): K(j, i) = dot_product(a_j, b_i), a serial sum. */
void oracle_linear(const double *A, int64_t na, const double *B, int64_t nb, int64_t d,
                   double *K) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i)
        for (int64_t j = 0; j < na; ++j) {
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) s = s + A[j * d + k] * B[i * d + k];
            K[j + i * na] = s;
        }
}

/* Fortran x**m for an integer m >= 1 by binary powering. */
static double powi_f(double x, int m) {
    double r = (m & 1) ? x : 1.0;
    while (m >>= 1) {
        x = x * x;
        if (m & 1) r = r * x;
    }
    return r;
}

/* fsargan_kernel (491-531): d = sum |a - b|; prefactor(m) = gamma_m (-inv d)**m;
 * K = exp(inv d) (1 + sum(prefactor)), inv = -1/sigma. */
void oracle_sargan(const double *A, int64_t na, const double *B, int64_t nb, int64_t d,
                   double sigma, const double *gammas, int64_t ng, double *K) {
    double inv = -1.0 / sigma;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i)
        for (int64_t j = 0; j < na; ++j) {
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) s = s + fabs(A[j * d + k] - B[i * d + k]);
            double pre = 0.0;
            for (int64_t m = 0; m < ng; ++m) pre = pre + gammas[m] * powi_f(-inv * s, (int)(m + 1));
            K[j + i * na] = exp(inv * s) * (1 + pre);
        }
}

/* fmatern_kernel_l2 (427-488): d2 = sum (a - b)^2, d = sqrt(d2); order 0: exp(inv d),
 * inv = -1/sigma; 1: exp(inv d)(1 - inv d), inv = -sqrt(3)/sigma; otherwise
 * exp(inv d)(1 - inv d + inv2 d2), inv = -sqrt(5)/sigma, inv2 = 5/(3 sigma^2). */
void oracle_matern_l2(const double *A, int64_t na, const double *B, int64_t nb, int64_t d,
                      double sigma, int order, double *K) {
    double inv = order == 0 ? -1.0 / sigma : order == 1 ? -sqrt(3.0) / sigma : -sqrt(5.0) / sigma;
    double inv2 = 5.0 / (3.0 * sigma * sigma);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i)
        for (int64_t j = 0; j < na; ++j) {
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) {
                double t = A[j * d + k] - B[i * d + k];
                s = s + t * t;
            }
            double dd = sqrt(s);
            double v;
            if (order == 0) v = exp(inv * dd);
            else if (order == 1) v = exp(inv * dd) * (1.0 - inv * dd);
            else v = exp(inv * dd) * (1.0 - inv * dd + inv2 * s);
            K[j + i * na] = v;
        }
}

/* fp_distance_integer (121-155) when p_is_int, else fp_distance_double (83-118):
 * D = (sum |a - b|**p)**(1/p). */
void oracle_p_distance(const double *A, int64_t na, const double *B, int64_t nb, int64_t d,
                       double p, int p_is_int, double *D) {
    double inv_p = 1.0 / p;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i)
        for (int64_t j = 0; j < na; ++j) {
            double s = 0.0;
            for (int64_t k = 0; k < d; ++k) {
                double t = fabs(A[j * d + k] - B[i * d + k]);
                s = s + (p_is_int ? powi_f(t, (int)p) : pow(t, p));
            }
            D[j + i * na] = pow(s, inv_p);
        }
}
