// Per-phase event timing hooks (see fl_profile.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace fl {

// call kinds reported by fl_profile_read
constexpr int PROF_DF_ASSEMBLE = 1;   // phases: qp, stiffness, traction, gather
constexpr int PROF_DF_RESIDUAL = 2;   // phases: qp, traction, gather

void prof_begin(hipStream_t st, int kind);
void prof_mark(hipStream_t st);

}  // namespace fl
