/*
 * libqmlhip: C ABI of the MI355X (gfx950) KRR hot path.
 *
 * Conventions
 *  - Every pointer is a device pointer (HBM) unless its name ends in _h (host memory).
 *    The caller owns every buffer; no entry point allocates device memory.
 *  - Representations are row-major (n, ld) FP64: row i at X + i*ld, features 0..d-1.
 *  - Kernel, distance and Cholesky matrices are column-major: (i, j) at K[i + j*ldk],
 *    i.e. the Fortran-ordered (na, nb) array the reference returns (qml/kernels.py:58).
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = the null stream) and returns
 *    a status: 0 ok, < 0 error (message in qmlh_last_error(), thread local). Numerical outcomes
 *    that LAPACK reports as `info` are written to a device int32 the caller reads back.
 *  - Calls on different streams may run concurrently; the library keeps no mutable global
 *    state besides the constant diagonal table of qmlh_init.
 */
#ifndef QMLHIP_H
#define QMLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *qmlh_stream_t;

enum {
    QMLH_OK = 0,
    QMLH_ERR_ARG = -1,
    QMLH_ERR_HIP = -2,
    QMLH_ERR_UNSUPPORTED = -3,
    QMLH_ERR_WORKSPACE = -4
};

/* Library version (major*10000 + minor*100 + patch). */
int qmlh_version(void);

/* Message of the last failing call on this thread ("" when none). */
const char *qmlh_last_error(void);

/* Upload the Coulomb diagonal table 0.5*Z**2.4 for integer Z in [0, 255] computed with the
 * host libm pow, the same call the reference Fortran makes (qml/frepresentations.f90:95).
 * Called once per device before qmlh_cm_batch; idempotent. */
int qmlh_init(void);

/* ---------------------------------------------------------------------------------------- *
 * Coulomb matrices.
 * Replaces fgenerate_coulomb_matrix (qml/frepresentations.f90:52-140, sorted = 1) and
 * fgenerate_unsorted_coulomb_matrix (qml/frepresentations.f90:142-205, sorted = 0), which
 * qml/representations.py:109-115 calls once per molecule; here one call covers a batch.
 * Z: (total_atoms,) nuclear charges as FP64 (the f2py coercion of the reference).
 * R: (total_atoms, 3) coordinates, row-major. offsets: (nmol + 1,) int64, molecule m owns
 * atoms offsets[m] .. offsets[m+1]-1. out: (nmol, ldo) row-major, ldo >= size*(size+1)/2;
 * columns beyond size*(size+1)/2 are written as zero. size <= QMLH_CM_MAX_SIZE.
 * status (device int32, may be NULL): set to 1 when some molecule has natoms > size, a case
 * the reference leaves undefined (it writes out of bounds); that molecule's row is zeroed.
 * ---------------------------------------------------------------------------------------- */
#define QMLH_CM_MAX_SIZE 256
int qmlh_cm_batch(const double *Z, const double *R, const int64_t *offsets, int64_t nmol,
                  int32_t size, int32_t sorted, double *out, int64_t ldo, int32_t *status,
                  qmlh_stream_t stream);

/* ---------------------------------------------------------------------------------------- *
 * Atomic (local) Coulomb matrices: one row per central atom.
 * Replaces fgenerate_local_coulomb_matrix (qml/frepresentations.f90:207-413, sorting
 * QMLH_ACM_ROW_NORM) and fgenerate_atomic_coulomb_matrix (qml/frepresentations.f90:415-598,
 * QMLH_ACM_DISTANCE), which qml/representations.py:223-231 calls once per molecule with that
 * molecule's central atom indices; here one call covers central atoms of many molecules.
 * Z, R, offsets: as qmlh_cm_batch. central: (nrows,) int64 index of the central atom within its
 * molecule (0-based); cmol: (nrows,) int64 molecule of each row. out: (nrows, ldo) row-major,
 * row r = the packed lower triangle of central atom central[r] of molecule cmol[r]; columns
 * beyond size*(size+1)/2 are zero. The four cut-off arguments mean what the reference's do and
 * are normalised the same way (negative central cut-off: none; negative decays: zero; ...).
 * Molecules may hold up to QMLH_ACM_MAX_ATOMS atoms. status (device int32, may be NULL): bit 1
 * when a central atom has more atoms within the central cut-off than `size` rows (the
 * reference stops the program there), bit 2 when a molecule exceeds QMLH_ACM_MAX_ATOMS, bit 4
 * when a central or molecule index is out of range; such rows are written as zeros.
 * ---------------------------------------------------------------------------------------- */
#define QMLH_ACM_MAX_ATOMS 256
enum { QMLH_ACM_ROW_NORM = 0, QMLH_ACM_DISTANCE = 1 };
int qmlh_atomic_cm_batch(const double *Z, const double *R, const int64_t *offsets, int64_t nmol,
                         const int64_t *central, const int64_t *cmol, int64_t nrows,
                         int32_t size, int32_t sorting, double central_cutoff,
                         double central_decay, double interaction_cutoff,
                         double interaction_decay, double *out, int64_t ldo, int32_t *status,
                         qmlh_stream_t stream);
/* Rows of molecules with more than QMLH_ACM_MAX_ATOMS atoms (up to max_atoms), which
 * qmlh_atomic_cm_batch reports with status bit 2 and leaves zero: the same rows, with each
 * molecule staged in a slice of the workspace instead of LDS. Same arguments and status bits
 * (bit 2 now for molecules above max_atoms); rows of smaller molecules are not touched, so a
 * batch of mixed sizes is one qmlh_atomic_cm_batch call plus this one. ws:
 * qmlh_atomic_cm_large_workspace(max_atoms, size) bytes. No counterpart of its own in the
 * reference, whose fgenerate_*_atomic_coulomb_matrix (qml/frepresentations.f90:207-598) has
 * no atom cap. */
size_t qmlh_atomic_cm_large_workspace(int64_t max_atoms, int32_t size);
int qmlh_atomic_cm_large(const double *Z, const double *R, const int64_t *offsets, int64_t nmol,
                         const int64_t *central, const int64_t *cmol, int64_t nrows, int32_t size,
                         int32_t sorting, double central_cutoff, double central_decay,
                         double interaction_cutoff, double interaction_decay, double *out,
                         int64_t ldo, int32_t *status, int64_t max_atoms, void *ws,
                         size_t ws_bytes, qmlh_stream_t stream);

/* ---------------------------------------------------------------------------------------- *
 * Pairwise kernels and distances, VALU tile kernels (direct differences).
 * kind QMLH_PAIR_LAPLACIAN: K(i,j) = exp((-1/sigma) * sum_k |A_ik - B_jk|)
 *      replaces flaplacian_kernel (qml/fkernels.f90:374-400) behind qml/kernels.py:36-63.
 * kind QMLH_PAIR_GAUSSIAN:  K(i,j) = exp((-0.5/(sigma*sigma)) * sum_k (A_ik - B_jk)^2)
 *      replaces fgaussian_kernel (qml/fkernels.f90:340-372) behind qml/kernels.py:65-92.
 * kind QMLH_PAIR_MANHATTAN: D(i,j) = sum_k |A_ik - B_jk|
 *      replaces fmanhattan_distance (qml/fdistance.f90:23-46) behind qml/distance.py:28-59.
 * kind QMLH_PAIR_L2: D(i,j) = sqrt(sum_k (A_ik - B_jk)^2)
 *      replaces fl2_distance (qml/fdistance.f90:49-80) behind qml/distance.py:61-92.
 * The feature sum runs k = 0..d-1 in order for every entry, as the reference loops do.
 * flags QMLH_SYMMETRIC: A == B (same rows); only tiles on or above the diagonal are computed
 * and mirrored, which is value-identical because |a-b| = |b-a| and (a-b)^2 = (b-a)^2.
 * ---------------------------------------------------------------------------------------- */
enum {
    QMLH_PAIR_LAPLACIAN = 0,
    QMLH_PAIR_GAUSSIAN = 1,
    QMLH_PAIR_MANHATTAN = 2,
    QMLH_PAIR_L2 = 3
};
enum { QMLH_SYMMETRIC = 1 };
/* QMLH_LOWER (with QMLH_SYMMETRIC; qmlh_gaussian_mfma, qmlh_pairwise, qmlh_pairwise_ext):
 * write only the lower triangle i >= j, computing the tiles on or below the diagonal without
 * the transposed stores; the upper triangle of K is left untouched. For callers that only
 * read the lower triangle (a Cholesky factorisation of K + lambda I, the CG's matvec); no
 * counterpart in the reference, whose kernels return the whole matrix. */
enum { QMLH_LOWER = 2 };
int qmlh_pairwise(int32_t kind, const double *A, int64_t na, int64_t lda, const double *B,
                  int64_t nb, int64_t ldb, int64_t d, double sigma, double *K, int64_t ldk,
                  int32_t flags, qmlh_stream_t stream);

/* Pairwise kernels with parameters (VALU tile kernels, direct differences), the epilogue
 * variants of the same feature sums. params_h: host array.
 * kind QMLH_PAIR_SARGAN: params = [sigma, gamma_1 .. gamma_ng] (ng <= 16):
 *      K(i,j) = exp(-d/sigma) (1 + sum_m gamma_m (d/sigma)^m), d = sum_k |A_ik - B_jk|;
 *      replaces fsargan_kernel (qml/fkernels.f90:491-531) behind qml/kernels.py:117-150 (and
 *      matern_kernel's L1 form, qml/kernels.py:152-216).
 * kind QMLH_PAIR_MATERN_L2: params = [sigma, order (0, 1, 2)], d = sqrt(sum_k (A_ik - B_jk)^2):
 *      exp(-d/s), exp(-sqrt3 d/s)(1 + sqrt3 d/s), exp(-sqrt5 d/s)(1 + sqrt5 d/s + 5 d^2/(3 s^2));
 *      replaces fmatern_kernel_l2 (qml/fkernels.f90:427-488).
 * kind QMLH_PAIR_P_INT: params = [p], integer p >= 1: D(i,j) = (sum_k |A_ik - B_jk|^p)^(1/p)
 *      (p = 1, 3, 4 by products, other p by pow); replaces fp_distance_integer
 *      (qml/fdistance.f90:121-155).
 * kind QMLH_PAIR_P_REAL: params = [p], p > 0, real powers; replaces fp_distance_double
 *      (qml/fdistance.f90:83-118). Both behind qml/distance.py:94-147 p_distance.
 * flags QMLH_SYMMETRIC and QMLH_LOWER as for qmlh_pairwise. */
enum {
    QMLH_PAIR_SARGAN = 4,
    QMLH_PAIR_MATERN_L2 = 5,
    QMLH_PAIR_P_INT = 6,
    QMLH_PAIR_P_REAL = 7
};
int qmlh_pairwise_ext(int32_t kind, const double *A, int64_t na, int64_t lda, const double *B,
                      int64_t nb, int64_t ldb, int64_t d, const double *params_h,
                      int32_t nparams, double *K, int64_t ldk, int32_t flags,
                      qmlh_stream_t stream);

/* Squared row norms norms[i] = sum_k X_ik^2 in FP64 (k in order). */
int qmlh_row_sqnorms(const double *X, int64_t n, int64_t ldx, int64_t d, double *norms,
                     qmlh_stream_t stream);

/* Gaussian kernel in the expanded form on FP64 MFMA (v_mfma_f64_16x16x4f64):
 * K(i,j) = exp((-0.5/(sigma*sigma)) * max(nA_i + nB_j - 2 A_i.B_j, 0)), with nA, nB from
 * qmlh_row_sqnorms. Same result as fgaussian_kernel (qml/fkernels.f90:340-372) to within
 * 4 u (|A_i|^2 + |B_j|^2) / (2 sigma^2) in the exponent; the drop-in picks this form only
 * where that bound is below 1e-13. Requires lda, ldb multiples of 4, d <= lda, ldb and the
 * pad columns [d, round_up(d, 16)) zero (or readable and zero in both operands).
 * With QMLH_SYMMETRIC the diagonal is exactly 1; with QMLH_SYMMETRIC | QMLH_LOWER only the
 * lower triangle is written. */
int qmlh_gaussian_mfma(const double *A, int64_t na, int64_t lda, const double *normsA,
                       const double *B, int64_t nb, int64_t ldb, const double *normsB,
                       int64_t d, double sigma, double *K, int64_t ldk, int32_t flags,
                       qmlh_stream_t stream);

/* Linear kernel K(i,j) = sum_k A_ik B_jk on FP64 MFMA (the same cross-term tiles as the
 * Gaussian); replaces flinear_kernel (qml/fkernels.f90:403-424) behind qml/kernels.py:94-115.
 * Same layout rules as qmlh_gaussian_mfma; QMLH_SYMMETRIC computes upper tiles and mirrors. */
int qmlh_linear_mfma(const double *A, int64_t na, int64_t lda, const double *B, int64_t nb,
                     int64_t ldb, int64_t d, double *K, int64_t ldk, int32_t flags,
                     qmlh_stream_t stream);

/* ---------------------------------------------------------------------------------------- *
 * Cholesky factorisation and solve.
 * qmlh_potrf replaces LAPACK dpotrf("U") at qml/fcho_solve.f90:35. The reference factors the
 * upper triangle (numpy indices i <= j) of a matrix Fortran sees column-major; the same bytes
 * read row-major (a C-ordered numpy array) are that matrix's transpose, so this routine
 * factors the LOWER triangle of a column-major matrix in place: A = L L^T, L in the lower
 * triangle, the strictly upper part untouched. info (device int32): 0, or j+1 when the
 * leading minor of order j+1 is not positive definite (LAPACK's convention); the remaining
 * columns are then left unfactored, as dpotrf leaves them.
 * qmlh_potrs replaces dpotrs("U", n, 1) at qml/fcho_solve.f90:46: x := (L L^T)^-1 x; ws holds
 * qmlh_potrs_workspace(n) bytes of scratch.
 * ---------------------------------------------------------------------------------------- */
int qmlh_potrf(double *A, int64_t n, int64_t lda, int32_t *info, qmlh_stream_t stream);
size_t qmlh_potrs_workspace(int64_t n);
int qmlh_potrs(const double *L, int64_t n, int64_t lda, double *x, void *ws, size_t ws_bytes,
               qmlh_stream_t stream);

/* A(i,i) += value for i < n (the K[diag] += lambda step of tests/test_energy_krr_cmat.py:104). */
int qmlh_add_diagonal(double *A, int64_t n, int64_t lda, double value, qmlh_stream_t stream);

/* Inverse of a symmetric positive definite matrix.
 * Replaces fcho_invert (qml/fcho_solve.f90:54-73: LAPACK dpotrf("L") + dpotri("L")) behind
 * qml/math.py:33-58 cho_invert. On entry the lower triangle of the column-major A holds the
 * matrix; on exit A holds the whole symmetric inverse (both triangles, as cho_invert returns it
 * after its lower-to-upper copy). info: dpotrf's (0, or the order of the first non-positive
 * leading minor; the inverse is then not meaningful, as in the reference). ws: device
 * workspace of qmlh_cho_invert_workspace(n) bytes, 16-byte aligned. */
size_t qmlh_cho_invert_workspace(int64_t n);
int qmlh_cho_invert(double *A, int64_t n, int64_t lda, int32_t *info, void *ws, size_t ws_bytes,
                    qmlh_stream_t stream);

/* Symmetric indefinite factorisation with Bunch-Kaufman pivoting, A = L D L^T (lower), and
 * solves with it. Replace LAPACK dsytrf("L") / dsytrs("L") in fbkf_solve and fbkf_invert
 * (qml/fcho_solve.f90:76-150) behind qml/math.py:100-164 bkf_invert / bkf_solve (whose inverse
 * here is the solve with the identity). qmlh_bkf_factor overwrites the lower triangle of the
 * column-major A with L and D as dsytrf does, writes the 1-based pivots to ipiv (device int32,
 * n; negative for both columns of a 2 x 2 block) and dsytrf's info (first zero pivot, or 0);
 * one workgroup walks the columns, so it suits orders up to a few thousand. ws: device
 * workspace of qmlh_bkf_workspace(n) bytes. qmlh_bkf_solve: B (n x nrhs column-major, ldb)
 * := A^-1 B with those factors, one workgroup per column. */
size_t qmlh_bkf_workspace(int64_t n);
int qmlh_bkf_factor(double *A, int64_t n, int64_t lda, int32_t *ipiv, int32_t *info, void *ws,
                    size_t ws_bytes, qmlh_stream_t stream);
int qmlh_bkf_solve(const double *A, int64_t n, int64_t lda, const int32_t *ipiv, double *B,
                   int64_t ldb, int64_t nrhs, qmlh_stream_t stream);

/* ---------------------------------------------------------------------------------------- *
 * Mixed precision (no single reference counterpart: together these replace the FP64
 * dpotrf/dpotrs pair of qml/fcho_solve.f90:35,46 in the KRR fit at FP64 accuracy).
 *
 * qmlh_spotrf: the Cholesky of qmlh_potrf on FP32 storage; trailing updates on FP32 MFMA,
 * diagonal blocks and substitutions in FP64 arithmetic. Same info convention.
 *
 * qmlh_gaussian_deflated: from one FP64 MFMA pass over X (n, ldx), with norms from
 * qmlh_row_sqnorms, s = -0.5/sigma^2 and c = x_i.x_j, writes
 *   u_i = exp(s * norms_i)                                               (n FP64)
 *   K(i, j) = exp(s * max(n_i + n_j - 2c, 0)) + lambda [i==j]     (lower, FP64; K may be NULL)
 *   B32(i, j) = u_i u_j (expm1(c/sigma^2) - c/sigma^2) + (lambda + shift) [i==j] (lower, FP32)
 *   V(i, :) = u_i [1, x_i / sigma]            (n x (d+1), column-major FP32; V may be NULL)
 * so that K + lambda I = V V^T + B32 - shift I exactly: the constant and linear terms of
 * exp(c/sigma^2) as an explicit rank-(d+1) factor, and a positive semidefinite remainder with
 * small entries that a low-precision factorisation resolves relative to lambda.
 *
 * qmlh_woodbury_setup: for the qmlh_spotrf factor L of B32 and V (n x r, r <= 1024),
 * overwrites V with Y = L^-1 V and writes C = chol(I + Y^T Y) (r x r, lower, FP64); ws holds
 * qmlh_woodbury_workspace(n) bytes (the FP64 inverses of L's 64 x 64 diagonal blocks).
 *
 * qmlh_pcg_solve: FP64 preconditioned CG on A x = y (A symmetric, column-major; only its lower
 * triangle is read), with
 * preconditioner M = L L^T + V V^T applied as M^-1 r = L^-T (q - Y C^-T C^-1 Y^T q),
 * q = L^-1 r (r = 0: M = L L^T); Y is FP32 (y_fp64 = 0) or FP64 (y_fp64 = 1). Iterates until |y - A x| <= tol |y| or max_iter. The stop
 * test runs on the CG's recursive residual and is confirmed on the true residual y - A x,
 * recomputed from A: above tol, the CG restarts from the true residual while each restart at
 * least halves it and the backward error (below) is above the unit roundoff u = 2^-53 (below
 * it, x solves a system within the rounding of A's own FP64 entries). relres_h receives the recomputed |y - A x| / |y|; backward_h (nullable) the
 * normwise backward error |y - A x| / (|A|_inf |x| + |y|) when relres > tol, else -1 (not
 * computed). Synchronises the stream once per iteration.
 * ---------------------------------------------------------------------------------------- */
int qmlh_spotrf(float *A, int64_t n, int64_t lda, int32_t *info, qmlh_stream_t stream);
/* qmlh_spotrf with the O(n^3) trailing updates in split FP16: with the panel scaled by a power
 * of two 2^e set from the largest diagonal entry (so it lies below 2^14), P P^T ~ 2^-2e (Ph Ph^T
 * + Ph Pl^T + Pl Ph^T), Ph = fp16(2^e P), Pl = fp16(2^e P - Ph), on FP16 MFMA with FP32
 * accumulation, about 2^-22 relative per update (the dropped Pl Pl^T term); for factors that
 * only precondition (qmlh_pcg_solve).
 * ws: qmlh_spotrf_f16x3_workspace(n) bytes, 16-byte aligned. */
size_t qmlh_spotrf_f16x3_workspace(int64_t n);
int qmlh_spotrf_f16x3(float *A, int64_t n, int64_t lda, int32_t *info, void *ws,
                       size_t ws_bytes, qmlh_stream_t stream);
int qmlh_gaussian_deflated(const double *X, int64_t n, int64_t ldx, const double *norms,
                           int64_t d, double sigma, double lambda, double shift, double *K,
                           int64_t ldk, float *B32, int64_t ldb, float *V, int64_t ldv,
                           double *u, qmlh_stream_t stream);
size_t qmlh_woodbury_workspace(int64_t n);
int qmlh_woodbury_setup(const float *L, int64_t n, int64_t ldl, float *V, int64_t ldv,
                        int64_t r, double *C, int64_t ldc, int32_t *info, void *ws,
                        size_t ws_bytes, qmlh_stream_t stream);
/* qmlh_spotrf (f16x3 = 0) or qmlh_spotrf_f16x3 (f16x3 != 0; ws as there, else ws may be
 * NULL) of A that also overwrites V (n x r, column-major, ldv >= n) with L^-1 V in the same
 * pass: V^T is placed in the r rows below the matrix (lda >= n + r; A[n:n+r, :] is
 * overwritten), solved as extra panel rows while the trailing updates run (their own updates in
 * FP32 MFMA; with f16x3 from order 4096 and r <= 1536, split-FP16 products with a scale of
 * their own, sharing the factor's workspace), and copied back. The factor of A is the same, bit for bit, as without V. On a
 * failed factorisation (info != 0) V is unspecified. Replaces qmlh_spotrf* + the solve of
 * qmlh_woodbury_setup. pivot_floor > 0 selects static pivoting, for factors that only
 * precondition: a finite pivot at or below pivot_floor is replaced by pivot_floor (the factor
 * is then of A + a diagonal perturbation in those columns) instead of failing; NaN pivots still
 * set info. pivot_floor = 0: LAPACK's rule. */
int qmlh_spotrf_lowrank(float *A, int64_t n, int64_t lda, float *V, int64_t ldv, int64_t r,
                        int32_t f16x3, double pivot_floor, int32_t *info, void *ws,
                        size_t ws_bytes, qmlh_stream_t stream);
/* qmlh_spotrf_lowrank with shift on breakdown (breakdown_shift > 0; for factors that only
 * precondition; no counterpart in the reference, which has no low-precision factor): a pivot
 * at or below breakdown_pivot (>= 0; for A = B + lambda I with B positive semidefinite every
 * exact pivot is at least lambda, so a smaller one is rounding) counts as a breakdown; where a
 * 64 x 64 diagonal block has one it is factored again with its diagonal raised by
 * breakdown_shift (then 10 breakdown_shift), and every later diagonal block is raised by the
 * shift that succeeded, so one pass yields the exact factor of
 * A + diag(0, .., 0, s1, .., s2, ..) (zero before the first failing block, nondecreasing)
 * instead of stopping at the first failed pivot. *shift_dev (device, one double) ends as the
 * last shift (0: no breakdown). A block that fails at 10 breakdown_shift sets info as a failed
 * pivot does. breakdown_shift = 0: exactly qmlh_spotrf_lowrank. */
int qmlh_spotrf_lowrank_shift(float *A, int64_t n, int64_t lda, float *V, int64_t ldv,
                              int64_t r, int32_t f16x3, double pivot_floor,
                              double breakdown_pivot, double breakdown_shift, double *shift_dev,
                              int32_t *info, void *ws, size_t ws_bytes, qmlh_stream_t stream);
/* C = chol(I + Y^T Y) (r x r, lower, FP64; r <= 1024) for Y (n x r, column-major FP32): the
 * rest of qmlh_woodbury_setup once Y = L^-1 V is known. The Gram matrix is summed in FP64 over
 * row slices in a fixed order (deterministic); ws holds
 * qmlh_woodbury_capacitance_workspace(n, r) bytes. */
size_t qmlh_woodbury_capacitance_workspace(int64_t n, int64_t r);
int qmlh_woodbury_capacitance(const float *Y, int64_t n, int64_t ldy, int64_t r, double *C,
                              int64_t ldc, int32_t *info, void *ws, size_t ws_bytes,
                              qmlh_stream_t stream);
/* The same capacitance factor for an FP64 Y (qmlh_trsm_f32l). */
int qmlh_woodbury_capacitance_f64(const double *Y, int64_t n, int64_t ldy, int64_t r, double *C,
                                  int64_t ldc, int32_t *info, void *ws, size_t ws_bytes,
                                  qmlh_stream_t stream);
/* V(i, 0) = u_i, V(i, 1 + k) = u_i X(i, k) / sigma (n x (d+1), column-major FP64; u as written
 * by qmlh_gaussian_deflated): the deflation's rank-(d+1) term at full precision. */
int qmlh_gaussian_lowrank(const double *X, int64_t n, int64_t ldx, int64_t d, double sigma,
                          const double *u, double *V, int64_t ldv, qmlh_stream_t stream);
/* V := L^-1 V for an FP32 lower factor L (column-major) and an FP64 right-hand side V (n x nrhs,
 * column-major), in FP64 arithmetic (FP64 inverses of the 64 x 64 diagonal blocks, FP64 MFMA
 * block updates on L widened to FP64 one 256-wide block column at a time): L^-1 V of the
 * stored factor to FP64 accuracy, for the Woodbury term of qmlh_pcg_solve with y_fp64 = 1.
 * ws: qmlh_trsm_f32l_workspace(n) bytes. */
size_t qmlh_trsm_f32l_workspace(int64_t n);
int qmlh_trsm_f32l(const float *L, int64_t n, int64_t ldl, double *V, int64_t ldv, int64_t nrhs,
                   void *ws, size_t ws_bytes, qmlh_stream_t stream);
/* The hybrid Woodbury term: Y (n x r, column-major FP64) gets columns k .. r-1 widened from
 * the factorisation's FP32 solve Y32 = L^-1 V (qmlh_spotrf_lowrank), and columns 0 .. k-1
 * solved again in FP64 arithmetic with the FP32 factor L from the FP64 right-hand sides Y holds
 * there on entry (qmlh_gaussian_lowrank's leading columns; u first). One single-launch solve
 * per FP64 column; from order 1024 in 128-row blocks whose diagonal step is a product with the
 * FP64 inverse of the block (computed once per call, as qmlh_pcg_solve's preconditioner does).
 * ws: qmlh_lowrank_hybrid_workspace(n) bytes. */
size_t qmlh_lowrank_hybrid_workspace(int64_t n);
int qmlh_lowrank_hybrid(const float *L, int64_t n, int64_t ldl, const float *Y32, int64_t ldy32,
                        double *Y, int64_t ldy, int64_t r, int64_t k, void *ws, size_t ws_bytes,
                        qmlh_stream_t stream);
/* B(i, j) = (float)(A(i, j) + shift [i == j]) on the lower triangle (column-major). */
int qmlh_lower_to_f32(const double *A, int64_t n, int64_t lda, float *B, int64_t ldb,
                      double shift, qmlh_stream_t stream);
size_t qmlh_pcg_workspace(int64_t n, int64_t r);
int qmlh_pcg_solve(const double *A, int64_t n, int64_t lda, const float *L, int64_t ldl,
                   const void *Y, int64_t ldy, int64_t r, int32_t y_fp64, const double *C,
                   int64_t ldc,
                   const double *y, double *x, double tol, int32_t max_iter, void *ws,
                   size_t ws_bytes, int32_t *iters_h, double *relres_h, double *backward_h,
                   qmlh_stream_t stream);

/* ---------------------------------------------------------------------------------------- *
 * Fused KRR prediction y_q = sum_j k(Xq_q, Xt_j) alpha_j without materialising the
 * (m, n) kernel; the composition Ks.T.dot(alpha) of tests/test_energy_krr_cmat.py:108-109.
 * kind: QMLH_PAIR_LAPLACIAN (VALU) or QMLH_PAIR_GAUSSIAN (MFMA expanded form; needs the
 * norms and the padding rules of qmlh_gaussian_mfma). Every query's sum over the training
 * set uses a fixed chunking that does not depend on m or on the other queries, so a query's
 * result is bitwise the same alone or in any batch. ws: workspace of at least
 * qmlh_predict_workspace(m, n) bytes.
 * ---------------------------------------------------------------------------------------- */
size_t qmlh_predict_workspace(int64_t m, int64_t n);
int qmlh_krr_predict(int32_t kind, const double *Xq, int64_t m, int64_t ldq,
                     const double *normsQ, const double *Xt, int64_t n, int64_t ldt,
                     const double *normsT, const double *alpha, int64_t d, double sigma,
                     double *y, void *ws, size_t ws_bytes, qmlh_stream_t stream);
/* qmlh_krr_predict without a workspace argument, allocating nothing: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The per-chunk partials live in a
 * library-owned device array of qmlh_krr_predict_noalloc_capacity() bytes (16 MiB) that each
 * GPU holds until the process ends, resolved for a device by qmlh_init (or by the first call
 * on it). Calls needing more -- ceil(n / 128) * m doubles -- return QMLH_ERR_ARG. Results are
 * bitwise those of qmlh_krr_predict. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams at once) race. */
size_t qmlh_krr_predict_noalloc_capacity(void);
int qmlh_krr_predict_noalloc(int32_t kind, const double *Xq, int64_t m, int64_t ldq,
                             const double *normsQ, const double *Xt, int64_t n, int64_t ldt,
                             const double *normsT, const double *alpha, int64_t d, double sigma,
                             double *y, qmlh_stream_t stream);

/* ---------------------------------------------------------------------------------------- *
 * Batched KRR fits: many independent problems of uneven size -- a learning curve over
 * training-set sizes, a (sigma, lambda) grid -- fitted together. Each problem is the plain
 * FP64 fit the reference's users write per problem (qml/kernels.py:33-87 laplacian_kernel /
 * gaussian_kernel on rows of a representation array, K[diag] += lambda, qml/math.py:61-100
 * cho_solve); the reference has no batched entry point, this replaces their loop.
 *
 * Problem p trains on rows x_row0 .. x_row0 + n - 1 of the shared row-major pool X (x_rows
 * rows, ldx >= d), reads its labels y[y_offset .. y_offset + n) and writes alpha[alpha_offset ..
 * alpha_offset + n) (alpha ranges must not overlap); info[p] follows LAPACK (0, or the first
 * failing column + 1). norms (squared row norms of X, qmlh_row_sqnorms) are needed when any
 * problem is Gaussian. nb (64, 128, 256 or 512) is the Cholesky block size. branches (1 .. 4): the
 * problems are split into that many groups of about equal n^3 work whose factorisations step
 * on streams of their own (the plan creates branches - 1 streams at the default priority and
 * forks from / joins back to the caller's stream), so small launches of one group overlap the
 * large ones of another; results do not depend on it.
 *
 * qmlh_krr_batched_create plans the launches and uploads the job tables into ws (a blocking
 * copy; ws 256-byte aligned, qmlh_krr_batched_workspace bytes; it also holds the factors).
 * qmlh_krr_batched_run enqueues the whole fit (kernel launches only, no host synchronisation,
 * no allocation); qmlh_krr_batched_capture records that sequence into a hipGraph on a
 * non-default stream, and qmlh_krr_batched_launch replays it as one graph launch. Every run
 * recomputes everything from X and y. qmlh_krr_batched_factor gives the element offset into
 * ws and the leading dimension of problem p's lower Cholesky factor.
 * ---------------------------------------------------------------------------------------- */
typedef struct qmlh_problem {
    int64_t n;
    int64_t x_row0;
    int64_t y_offset;
    int64_t alpha_offset;
    double sigma;
    double lambda;
    int32_t kind; /* QMLH_PAIR_GAUSSIAN or QMLH_PAIR_LAPLACIAN */
    int32_t reserved;
} qmlh_problem;

typedef struct qmlh_batched_plan qmlh_batched_plan;

size_t qmlh_krr_batched_workspace(const qmlh_problem *probs, int32_t nprob, int32_t nb,
                                  int32_t branches);
int qmlh_krr_batched_create(const qmlh_problem *probs, int32_t nprob, int32_t nb,
                            int32_t branches, const double *X, int64_t x_rows, int64_t ldx, int64_t d,
                            const double *norms, const double *y, int64_t y_len, double *alpha,
                            int64_t alpha_len, int32_t *info, void *ws, size_t ws_bytes,
                            qmlh_batched_plan **plan);
int qmlh_krr_batched_run(qmlh_batched_plan *plan, qmlh_stream_t stream);
int qmlh_krr_batched_capture(qmlh_batched_plan *plan, qmlh_stream_t stream);
int qmlh_krr_batched_launch(qmlh_batched_plan *plan, qmlh_stream_t stream);
int64_t qmlh_krr_batched_launches(const qmlh_batched_plan *plan);
int qmlh_krr_batched_factor(const qmlh_batched_plan *plan, int32_t p, int64_t *offset,
                            int64_t *ld);
void qmlh_krr_batched_destroy(qmlh_batched_plan *plan);

/* ---------------------------------------------------------------------------------------- *
 * Local (atom-summed) kernels between molecules, several sigmas per call:
 *   K_s(a, b) = sum over atoms i of molecule a, j of molecule b of exp(inv_s * D(A_i, B_j))
 * Replaces fget_local_kernels_gaussian / fget_local_kernels_laplacian (qml/fkernels.f90:23-196)
 * behind qml/kernels.py:218-298, and fget_vector_kernels_gaussian / _laplacian
 * (qml/fkernels.f90:198-337) behind qml/wrappers.py:33-97, which give the same values from
 * zero-padded per-molecule arrays.
 * kind QMLH_PAIR_GAUSSIAN: D = squared L2 distance, inv_s = -0.5/sigma_s^2. With normsA and
 *      normsB (|x|^2 per atom row, qmlh_row_sqnorms) the distance is the expanded form on FP64
 *      MFMA; with both NULL, direct differences on the VALU.
 * kind QMLH_PAIR_LAPLACIAN: D = L1 distance, inv_s = -1/sigma_s (VALU, direct differences).
 * A: (atomsA, lda) row-major atom representations; offsA: (nmA + 1) int64 atom offsets of A's
 * molecules; panelsA: (npA + 1) int64 panel starts from qmlh_local_panels (copied to the
 * device). Same for B. sigmas_h: host array of nsig > 0 values. K_s(a, b) is written at
 * K[a + b*ldk + s*sstride] (column-major per sigma, the reference's kernels(s, :, :)); every
 * entry is written. flags QMLH_SYMMETRIC: A and B are the same molecules (same pointers and
 * panels); only panel pairs P <= Q are computed and mirrored, so every K_s is exactly
 * symmetric.
 * ---------------------------------------------------------------------------------------- */
#define QMLH_LOCAL_PANEL_ATOMS 128
#define QMLH_LOCAL_PANEL_MOLS 16
int qmlh_local_kernels(int32_t kind, const double *A, int64_t atomsA, int64_t lda,
                       const double *normsA, const int64_t *offsA, const int64_t *panelsA,
                       int64_t npA, const double *B, int64_t atomsB, int64_t ldb,
                       const double *normsB, const int64_t *offsB, const int64_t *panelsB,
                       int64_t npB, int64_t d, const double *sigmas_h, int32_t nsig, double *K,
                       int64_t ldk, int64_t sstride, int32_t flags, qmlh_stream_t stream);

/* Panel partition for qmlh_local_kernels (host only, no device work; no counterpart in the
 * reference, whose loops run molecule pair by molecule pair): consecutive molecules are grouped
 * into panels of at most QMLH_LOCAL_PANEL_ATOMS atoms and QMLH_LOCAL_PANEL_MOLS molecules; a
 * molecule with more atoms than QMLH_LOCAL_PANEL_ATOMS is a panel alone. offsets_h: (nmol + 1)
 * host int64 atom offsets; panels_h: (nmol + 1) host int64 output, panels_h[p] = first molecule
 * of panel p, panels_h[np] = nmol. Returns np >= 0, or an error status < 0. */
int64_t qmlh_local_panels(const int64_t *offsets_h, int64_t nmol, int64_t *panels_h);

/* ---------------------------------------------------------------------------------------- *
 * Measurement hooks (no counterpart in the reference; used by bench.py). When enabled, each
 * heavy launch is bracketed by HIP events on its stream and its algorithmic work is added to
 * its class: flops for the MFMA classes and the Cholesky, pair-features for the VALU pairwise
 * class, molecules for the Coulomb matrices. qmlh_timer_read waits for the recorded events.
 * ---------------------------------------------------------------------------------------- */
enum {
    QMLH_TIMER_SYRK = 0,      /* potrf trailing updates (lower, FP64 MFMA)        flops */
    QMLH_TIMER_GEMM = 1,      /* potrf panel/diagonal-block GEMM updates         flops */
    QMLH_TIMER_GAUSS = 2,     /* Gaussian kernel matrix, FP64 MFMA               flops */
    QMLH_TIMER_PAIR = 3,      /* VALU pairwise kernels                  pair-features */
    QMLH_TIMER_PREDICT = 4,   /* fused predict                            query-train */
    QMLH_TIMER_POTRF = 5,     /* whole factorisation                      n^3/3 flops */
    QMLH_TIMER_POTRS = 6,     /* both triangular solves                     2 n^2    */
    QMLH_TIMER_CM = 7,        /* Coulomb matrices                           molecules */
    QMLH_TIMER_LOCAL = 8,     /* local kernels                  atom-pair features    */
    QMLH_TIMER_PCG = 9,       /* mixed-fit refinement, qmlh_pcg_solve          iterations */
    QMLH_TIMER_CLASSES = 10
};
int qmlh_timer_enable(int32_t on);
int qmlh_timer_read(int32_t timer_class, double *total_ms_h, int64_t *launches_h,
                    double *work_h);

#ifdef __cplusplus
}
#endif

#endif /* QMLHIP_H */
