// Library-wide entry points: version, error text, one-time device set-up.
#include "common.h"

#include <math.h>
#include <mutex>
#include <string.h>

namespace qmlh {

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// Defined in cm.hip: uploads the diagonal table to the current device.
int upload_cm_diag_table(const double* table_h, int n);
// Defined in predict.hip: resolves the current device's copy of the prediction scratch.
int predict_scratch_setup(double** out);

}  // namespace qmlh

extern "C" int qmlh_version(void) { return 100; }

extern "C" const char* qmlh_last_error(void) { return qmlh::g_err; }

extern "C" int qmlh_init(void) {
    // 0.5 * Z**2.4 with the host libm pow, as qml/frepresentations.f90:95 evaluates it.
    static double table[256];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int z = 0; z < 256; ++z) table[z] = 0.5 * pow((double)z, 2.4);
    });
    QMLH_TRY(qmlh::upload_cm_diag_table(table, 256));
    return qmlh::predict_scratch_setup(nullptr);
}
