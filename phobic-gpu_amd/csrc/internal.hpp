// internal.hpp -- library-internal helpers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/phobic_hip.h"

namespace phobic {

// The stream a context's host-buffer builds run on.
hipStream_t ctx_stream(phobic_ctx* ctx);

}  // namespace phobic
