// noalloc.hip -- the allocation-free build entry point.
//
// phobic_build_u64_device_noalloc runs the same kernels as
// phobic_build_u64_device, with all scratch carved from the file-scope
// device array below instead of the context's memory pool.  Each device
// holds its own copy of the array (at its own address) from the first use
// of the library on it until the process ends.  Host code never passes the
// array's name to a kernel or copy: it uses the address hipGetSymbolAddress
// returns on each device, looked up once when the first context on that
// device is created.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/phobic_hip.h"
#include "internal.hpp"
#include "kernels.hpp"

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[PHOBIC_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};
// the fulcrum table the array of each device holds (all zero = none yet)
uint32_t g_fulcs[kMaxDevices][phobic::kFulcsTable] = {};

}  // namespace

int phobic::noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return PHOBIC_E_NO_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[device]) return PHOBIC_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return PHOBIC_E_HIP;
  g_scratch_ptr[device] = p;
  return PHOBIC_OK;
}

extern "C" int phobic_build_u64_device_noalloc(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n,
                                               const phobic_config* cfg, const phobic_outputs* d_out,
                                               uint32_t* d_status, void* stream) {
  if (!ctx) return PHOBIC_E_INVALID_CONFIG;
  const int dev = phobic::ctx_device(ctx);
  if (dev < 0 || dev >= kMaxDevices) return PHOBIC_E_NO_DEVICE;
  void* p;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    p = g_scratch_ptr[dev];
  }
  if (!p) return PHOBIC_E_HIP;  // phobic_ctx_create looks it up
  return phobic::build_u64_in_scratch(ctx, d_keys, n, cfg, d_out, d_status, (hipStream_t)stream, p,
                                      PHOBIC_NOALLOC_SCRATCH_BYTES, g_fulcs[dev]);
}
