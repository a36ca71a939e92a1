// Process-wide execution context of the gs_native classes. The reference issues all work on the
// legacy default stream; here every class uses one shared gsn_context (one HIP stream), so the
// renderer's outputs are ordered before the solver's reads exactly as in the reference.
#pragma once

#include "gsn_hip.h"

namespace gs_native {

// Returns the shared context, creating it on first use (nullptr if no device is available).
gsn_context* DefaultContext();

} // namespace gs_native
