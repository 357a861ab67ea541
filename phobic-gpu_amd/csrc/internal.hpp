// internal.hpp -- library-internal helpers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "../../include/phobic_hip.h"
#include "kernels.hpp"

namespace phobic {

// The stream a context's host-buffer builds run on.
hipStream_t ctx_stream(phobic_ctx* ctx);
int ctx_device(const phobic_ctx* ctx);

// A group: several independent key sets built as one flattened launch
// (every set's outputs equal building it alone with its own
// configuration).  The sets share the partition size, the load factor and
// the bucketer (a custom fulcrum table only when every set has the same
// configuration); lambda and the XXH3 seed may differ from set to set, with
// at most kMaxFulcTabs distinct lambdas.  Set s holds keys
// [key_off[s], key_off[s+1]) of the group's key array; its outputs in the
// group's output block are
//   pilots         block.pilots + pilot_base[s]        (np_s * B_s)
//   offsets        block.offsets + part_base[s] + s    (np_s + 1)
//   remap_offsets  block.remap_offsets + part_base[s] + s (np_s + 1, alpha < 1)
//   remap          block.remap + remap_base[s]          (alpha < 1)
// with block.pilots holding pilot_base[sets] entries, block.offsets and
// block.remap_offsets np + sets, block.remap remap_base[sets].
struct GroupLayout {
  uint32_t sets = 0, np = 0, B = 0, a_ppm = 1000000;  // B: the largest of the sets'
  uint64_t n = 0;
  std::vector<uint64_t> key_off;
  std::vector<uint32_t> part_base, remap_base;
  // per set: its pilots start at pilot_base[s] (sets + 1 entries); when the
  // sets' configurations differ (mixed), its seed, buckets per partition and
  // fulcrum table (ntabs tables of kFulcsTable entries in ftabs)
  std::vector<uint64_t> pilot_base;
  bool mixed = false;
  std::vector<uint64_t> seed;
  std::vector<uint32_t> setB, ftab, ftabs;
  uint32_t ntabs = 0;
  // u64 sets read where they lie: per set the device address of its keys
  // (empty: the sets are concatenated in the group's key array)
  std::vector<const uint64_t*> set_keys;
};
// cfgs[s]: the configuration of set s
int group_layout(const phobic_config* const* cfgs, const uint64_t* set_sizes, uint32_t sets, GroupLayout* gl);
// Enqueues the group's build on s (kind: u64 keys, or bytes with d_offs the
// group's n + 1 string offsets into d_keys).
int group_build(phobic_ctx* ctx, const phobic_config* cfg, int kind, const void* d_keys, const uint64_t* d_offs,
                const GroupLayout& gl, const phobic_outputs& block, uint32_t* d_status, hipStream_t s);
// Device bytes the context's scratch would grow by for a build of n keys /
// a group (0 when it already holds enough; from_empty: as if it held
// none), counted in the pool's units.
uint64_t build_scratch_growth(const phobic_ctx* ctx, const phobic_config* cfg, uint64_t n, bool from_empty = false);
uint64_t group_scratch_growth(const phobic_ctx* ctx, const phobic_config* cfg, const GroupLayout& gl,
                              bool from_empty = false);
// Device bytes the context holds (scratch and staging), in the pool's units,
// and their release (no build may be running on the context).
uint64_t ctx_scratch_bytes(const phobic_ctx* ctx);
void ctx_trim(phobic_ctx* ctx);
// An allocation of b bytes in the pool's units, and the bytes the current
// device's pool has handed out.
uint64_t pool_bytes(uint64_t b);
uint64_t pool_used_bytes();

// Writes bytes [off, off + len) of a staged upload to dst (host memory).
// Called from several threads at once, for disjoint ranges.
using StageFill = std::function<void(char* dst, uint64_t off, uint64_t len)>;
// The `bytes` bytes that fill produces, into dst (device memory), through the
// context's two pinned host chunks: while one chunk's DMA runs on s, the
// next chunk is filled.  Returns once fill has produced every byte (the
// caller's host data is then no longer read); the copies are ordered on s.
int stage_to_device(phobic_ctx* ctx, void* dst, uint64_t bytes, const StageFill& fill, hipStream_t s);
// Allocates the context's pinned chunks now instead of at first use.
int ctx_prepare_staging(phobic_ctx* ctx);

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
