// Shared helpers for libqmlhip: status codes, the thread-local error message, launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "qmlhip.h"

namespace qmlh {

// Record a message for qmlh_last_error() and return `code`.
int set_error(int code, const char* fmt, ...);

// Fold a HIP status into the library status convention.
inline int check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) return set_error(QMLH_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return QMLH_OK;
}

#define QMLH_CHECK_LAUNCH(what)                                                      \
    do {                                                                             \
        int _rc = ::qmlh::check_hip(hipGetLastError(), what);                        \
        if (_rc) return _rc;                                                         \
    } while (0)

#define QMLH_TRY(expr)                                                               \
    do {                                                                             \
        int _rc = (expr);                                                            \
        if (_rc) return _rc;                                                         \
    } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// FP64 MFMA operand/accumulator vector types.
typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

}  // namespace qmlh
