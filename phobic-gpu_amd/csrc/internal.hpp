// internal.hpp -- library-internal helpers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/phobic_hip.h"

namespace phobic {

// The stream a context's host-buffer builds run on.
hipStream_t ctx_stream(phobic_ctx* ctx);
int ctx_device(const phobic_ctx* ctx);

// Looks up (once per device; the device must be current) the device address
// of the fixed scratch array of the allocation-free entry point.
int noalloc_init(int device);

// phobic_build_u64_device with every scratch region carved from
// [base, base + bytes) (device memory) instead of the context pool: no
// allocation.  fulcs_host records the fulcrum table the region at base
// holds (uploaded only when the configuration's table differs).
// PHOBIC_E_SCRATCH_LIMIT when the build needs more than `bytes`.
int build_u64_in_scratch(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n, const phobic_config* cfg,
                         const phobic_outputs* d_out, uint32_t* d_status, hipStream_t s, void* base, size_t bytes,
                         uint32_t* fulcs_host);

// Stream-ordered device memory from the library's pool of the current
// device.  Freed blocks stay in the pool (no device-wide synchronisation on
// free), so concurrent builds on other streams are not stalled.
hipError_t dev_alloc(void** p, size_t bytes, hipStream_t s);
hipError_t dev_free(void* p, hipStream_t s);
// Returns the pool's unused memory to the device (after synchronisation).
void dev_trim();

}  // namespace phobic
