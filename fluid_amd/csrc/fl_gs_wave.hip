// Temporal-blocked Gauss-Seidel kernel (placeholder until the systolic kernel lands).
#include <hip/hip_runtime.h>

#include "fl_internal.h"

namespace fl {

bool gs_wave_supported(Ctx*, int) { return false; }
int gs_wave_solve(Ctx*, int, float) { return FL_ERR_UNSUPPORTED; }
int pressure_free(Ctx*) { return FL_OK; }

}  // namespace fl
