// internal.hpp -- library-internal helpers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/phobic_hip.h"

namespace phobic {

// The stream a context's host-buffer builds run on.
hipStream_t ctx_stream(phobic_ctx* ctx);

// Stream-ordered device memory from the library's pool of the current
// device.  Freed blocks stay in the pool (no device-wide synchronisation on
// free), so concurrent builds on other streams are not stalled.
hipError_t dev_alloc(void** p, size_t bytes, hipStream_t s);
hipError_t dev_free(void* p, hipStream_t s);
// Returns the pool's unused memory to the device (after synchronisation).
void dev_trim();

}  // namespace phobic
