// Internal dense linear-algebra building blocks shared by the Cholesky, the refinement and
// the batched solver (implemented in potrf.hip).
#pragma once

#include "common.h"

namespace qmlh {

// Lower Cholesky in place (column-major), storage T in {double, float}; LAPACK info rules.
// For float, a non-null split_ws (qmlh_spotrf_f16x3_workspace(n) bytes) selects split-FP16
// trailing updates instead of FP32 MFMA ones. r > 0: the r rows below the matrix
// (A[n:n+r, 0:n], lda >= n + r) are solved along, ending as W L^-T.
// Pivot handling of the diagonal blocks. floor > 0: static pivoting (a pivot at or below it is
// replaced by it, not a failure). shift_acc non-null and bshift > 0: shift on breakdown (a
// pivot at or below bpivot fails; a block with a failed pivot is re-factored with its diagonal
// raised by bshift, then 10 bshift, and every later block gets the same shift; *shift_acc, zero
// on entry, ends as the last shift; a block that fails at 10 bshift sets info).
struct PivotRule {
    double floor = 0.0;
    double bshift = 0.0;
    double* shift_acc = nullptr;
    double bpivot = 0.0;
};

template <typename T>
int potrf_impl(T* A, int64_t n, int64_t lda, int32_t* info, hipStream_t s,
               void* split_ws = nullptr, int64_t r = 0, PivotRule pr = PivotRule());

// Workspace of the split: a header (the scale), then regions of f16x3_region_bytes(n, kb)
// each holding the hi/lo panels of one split of up to n rows and kb features.
size_t f16x3_region_bytes(int64_t n, int64_t kb);
size_t syrk_f16_workspace(int64_t n, int64_t kb);
// Record the scale of the split panels (from the largest diagonal entry of the matrix about
// to be factored) in ws; before the first split.
int f16x3_prepare(const float* A, int64_t n, int64_t lda, void* ws, hipStream_t s);
// Split P (m x kb, FP32 column-major) into the scaled FP16 hi/lo panels of the region that
// starts `region` bytes after the header of ws.
int f16x3_split(const float* P, int64_t ldp, int64_t m, int64_t kb, void* ws, hipStream_t s,
                size_t region);
// C -= P P^T in split FP16 from a split of m_split rows, for the m rows starting at row off
// (a multiple of 128): ncols == 0 the lower triangle of the m x m block, ncols > 0 the entries
// i >= j of its first ncols columns. wide_cols: ncols > 0 on 256 x 256 tiles (one workgroup per
// CU) from order 4096 on, else on 128 x 128 tiles (two per CU).
int f16x3_update(const void* ws, int64_t m_split, int64_t kb, int64_t off, int64_t m,
                  int64_t ncols, float* C, int64_t ldc, const int32_t* info, hipStream_t s,
                  size_t region, bool wide_cols = false);
// C (r x m) -= X P^T: X (r x kb, FP32 column-major) split with its own scale into the region at
// region_x (its entries need not be bounded by the factor's diagonal), P the split of m rows in
// the region at region_p (kb features, as f16x3_split left it); three FP16 products per term.
int f16x3_gemm_rows(const float* X, int64_t ldx, int64_t r, int64_t kb, void* ws,
                    size_t region_x, size_t region_p, int64_t m, float* C, int64_t ldc,
                    const int32_t* info, hipStream_t s);
// C (m x m, lower) -= P P^T, P (m x kb) FP32 column-major, as three FP16 products.
int syrk_f16x3(const float* P, int64_t ldp, int64_t m, int64_t kb, float* C, int64_t ldc,
                const int32_t* info, void* ws, hipStream_t s);

// Synchronisation words the single-launch triangular solve needs for order n.
size_t trsv_sync_ints(int64_t n);

// Single-launch triangular solve (forward z = L^-1 rhs, or backward x = L^-T rhs) in 64-row
// blocks with FP64 substitution.
template <typename T>
int trsv_syncfree(const T* L, int64_t n, int64_t ldl, const double* rhs, double* out,
                  int32_t* sync, bool forward, hipStream_t s);

// Orders from which the triangular solves run as one sync-free launch.
static constexpr int64_t TRSV_SYNCFREE_MIN = 1024;
// 128-row-block solves for a factor whose 128 x 128 diagonal inverses are known: workspace
// doubles, the inverses (from the 64 x 64 ones, written to dinv64 first), and the solve.
size_t trsv_dinv128_doubles(int64_t n);
template <typename T>
int trsv_inverse128(const T* L, int64_t n, int64_t ldl, double* dinv64, double* dinv128,
                    hipStream_t s);
template <typename T>
int trsv_syncfree128(const T* L, int64_t n, int64_t ldl, const double* rhs, double* out,
                     int32_t* sync, bool forward, hipStream_t s, const double* dinv128);
// Doubles trsv_diag_inverse writes for order n, and the inverses themselves.
size_t trsv_dinv_doubles(int64_t n);
template <typename T>
int trsv_diag_inverse(const T* L, int64_t n, int64_t ldl, double* dinv, hipStream_t s);

// x := (L L^T)^-1 x; tmp is n doubles of scratch, sync trsv_sync_ints(n) ints.
template <typename T>
int potrs_impl(const T* L, int64_t n, int64_t lda, double* x, double* tmp, int32_t* sync,
               hipStream_t s);

// z := L^-1 r (r may be clobbered); one sync-free launch for large n when sync is given.
template <typename T>
int trsv_lower_fwd(const T* L, int64_t n, int64_t lda, double* r, double* z, int32_t* sync,
                   hipStream_t s);

// x := L^-T z (z may be clobbered).
template <typename T>
int trsv_lower_bwd(const T* L, int64_t n, int64_t lda, double* z, double* x, int32_t* sync,
                   hipStream_t s);

// V := L^-1 V for an (n, nrhs) column-major right-hand side, L lower (both float or double);
// dinv: the FP64 inverses of L's 64 x 64 diagonal blocks (trsv_diag_inverse).
template <typename T>
int trsm_lower_fwd_multi(const T* L, int64_t n, int64_t ldl, T* V, int64_t ldv, int64_t nrhs,
                         const double* dinv, hipStream_t s, bool lower_rhs = false);

}  // namespace qmlh
