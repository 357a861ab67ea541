// kernels.hpp -- argument blocks and launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace phobic {

constexpr uint32_t kFulcsTable = 2048;  // shaders/constants.glsl:1 FULCS_INTER

enum : int { kKindU64 = 0, kKindBytes = 1, kKindKey = 2 };

// Error bits raised by kernels (bit index = phobic_status code).
enum : uint32_t {
  kErrPartitionOverflow = 5,
  kErrBucketOverflow = 6,
  kErrDuplicateKeys = 7,
  kErrPilotOverflow = 8,
};

struct KeySource {
  const uint64_t* u64;    // kKindU64
  const uint8_t* bytes;   // kKindBytes: concatenated strings
  const uint64_t* offs;   // kKindBytes: n+1 offsets into bytes
  const uint4* keys;      // kKindKey: pre-hashed 16-byte keys
};

struct ScatterArgs {
  KeySource src;
  uint64_t n;
  uint64_t seed;
  uint32_t np, cap;
  const uint32_t* fulcs;  // 2048 entries
  uint32_t* pcount;       // np, zeroed before the launch
  uint64_t* rec_lower;    // np*cap: lower1 | lower2 << 32
  uint16_t* rec_bkt;      // np*cap: bucket inside the partition
  uint32_t* status;       // error bit mask
};

struct SearchArgs {
  const uint64_t* rec_lower;
  const uint16_t* rec_bkt;
  const uint32_t* pcount;
  uint32_t np, B, cap;
  uint32_t* pilots;  // np*B, interleaved p + b*np
  uint32_t* status;
};

struct EvalArgs {
  const uint64_t* keys;
  uint64_t count;
  uint64_t seed;
  uint32_t np;
  const uint32_t* fulcs;
  const uint32_t* pilots;
  const uint32_t* offsets;
  uint32_t* out;
};

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_hash_scatter(int kind, const ScatterArgs& a, hipStream_t s);
hipError_t launch_partition_offsets(const uint32_t* counts, uint32_t np, uint32_t* offsets, hipStream_t s);
size_t search_lds_bytes(uint32_t cap, uint32_t B);
hipError_t launch_partition_search(const SearchArgs& a, hipStream_t s);
hipError_t launch_eval_u64(const EvalArgs& a, hipStream_t s);
hipError_t launch_check_bijection(const uint32_t* pos, uint64_t count, uint64_t n, uint32_t* bitmap,
                                  unsigned long long* bad, hipStream_t s);
hipError_t launch_splitmix64(uint64_t seed, uint64_t n, uint64_t* out, hipStream_t s);

}  // namespace phobic
