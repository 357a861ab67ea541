// kernels.hpp -- argument blocks and launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace phobic {

constexpr uint32_t kFulcsTable = 2048;  // shaders/constants.glsl:1 FULCS_INTER

// key kinds: u64 keys, byte strings, pre-hashed Keys (the reference's
// nohash), and (evaluation only) 16-byte keys hashed with XXH3:
// std::pair<u64, u64> and an xxhash-hashed Key (hasher.hpp:51-66)
enum : int { kKindU64 = 0, kKindBytes = 1, kKindKey = 2, kKindPair = 3 };

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

constexpr uint32_t kMaxFanout = 1024;  // child bins per partitioning level

// Records a partitioning level writes.
enum : int {
  kFmtRaw = 0,    // the u64 key itself: plain u64 builds re-hash it at the next level
  kFmtRec = 1,    // lower64 + meta32 (partition inside the bin << 16 | bucket)
  kFmtFinal = 2,  // lower64 + bucket16: the last level, which the search reads
};

// Level 1 of the partitioning: hash keys and route them to level-1 bins.
struct HashRouteArgs {
  KeySource src;
  uint64_t n;
  uint64_t seed;
  uint32_t np, cap;
  uint32_t G;             // partitions per level-1 bin
  uint32_t fanout;        // number of level-1 bins, <= kMaxFanout
  const uint32_t* fulcs;  // 2048 entries
  uint32_t* child_count;  // per level-1 bin, zeroed before the launch
  uint64_t* out_lo;       // raw keys, or lower1 | lower2 << 32
  void* out_meta;         // u32 meta (kFmtRec) or u16 bucket (kFmtFinal)
  uint32_t* status;       // error bit mask
  // batched builds (sets > 0): key i of set s lies in [set_key_off[s],
  // set_key_off[s+1]); its partition is set_part_base[s] + hi32(x * np_s)
  uint32_t sets;
  const uint64_t* set_key_off;
  const uint32_t* set_part_base;
  // sets of several configurations (null: every set uses seed and fulcs):
  // set s hashes with set_seed[s] and maps buckets with fulcrum table
  // set_ftab[s] of the ntabs tables in ftabs (kFulcsTable entries each)
  const uint64_t* set_seed;
  const uint32_t* set_ftab;
  const uint32_t* ftabs;
  uint32_t ntabs;
  // u64 sets read in place (null: the keys are src.u64, concatenated):
  // key i of set s is set_keys[s][i - set_key_off[s]]
  const uint64_t* const* set_keys;
};

// Most fulcrum tables one launch holds (sets that differ in lambda).
constexpr uint32_t kMaxFulcTabs = 4;

// Levels 2..L: split every parent bin into `fanout` child bins.
struct SplitArgs {
  const uint64_t* in_lo;
  const void* in_meta;    // u32 meta of kFmtRec input
  const uint32_t* parent_count;
  uint64_t parent_cap;    // G_parent * cap
  uint32_t slices;        // tiles per parent bin (grid = parents * slices)
  uint32_t fanout;        // children per parent
  uint32_t G_parent;      // partitions per parent bin
  uint32_t G_child;       // partitions per child bin
  uint32_t cap;
  // kFmtRaw input: the key is hashed again
  uint32_t np;
  uint64_t seed;
  const uint32_t* fulcs;
  uint32_t* child_count;  // per child bin, zeroed before the launch
  uint64_t* out_lo;
  void* out_meta;
  uint32_t* status;
};

// Table size of a partition with ps keys at load factor a_ppm / 1e6:
// ceil(ps * 1e6 / a_ppm); ps itself when a_ppm = 1e6 (the reference).
__host__ __device__ inline uint32_t table_size(uint32_t ps, uint32_t a_ppm) {
  return a_ppm >= 1000000u ? ps : (uint32_t)(((uint64_t)ps * 1000000u + a_ppm - 1) / a_ppm);
}

struct SearchArgs {
  const uint64_t* rec_lower;
  const uint16_t* rec_bucket;  // bucket of each record
  const uint32_t* pcount;
  uint64_t* sorted;  // np*cap scratch: keys of each partition in bucket order
  uint32_t np, B, cap;
  uint32_t a_ppm;    // load factor in ppm (1e6 = alpha 1)
  uint32_t tcap;     // table_size(cap, a_ppm)
  uint32_t* pilots;  // np*B, interleaved p + b*np
  uint32_t* remap;   // alpha < 1: remap entries, else null
  const uint32_t* remap_offsets;
  uint32_t* status;
  // batched builds (sets > 0): global partition p belongs to the set s with
  // set_part_base[s] <= p < set_part_base[s+1]; its pilots go to
  // pilots + B*set_part_base[s], interleaved by that set's partition count.
  // Sets of several configurations: set s has set_B[s] <= B buckets per
  // partition and its pilots start at pilots + set_pilot_base[s] (null: B
  // and B*set_part_base[s]); B sizes the LDS layout.
  uint32_t sets;
  const uint32_t* set_part_base;
  const uint32_t* set_B;
  const uint64_t* set_pilot_base;
  uint32_t mode;  // mask of kSearchSinglePilot / kSearchNarrow
};

// Pilot-search variants (identical results, different speed).
enum : uint32_t {
  kSearchSinglePilot = 1,  // one pilot per hash round for every bucket
  kSearchNarrow = 2,       // scan passes of 2048 offsets for every table size
};

struct EvalArgs {
  KeySource src;  // u64 / bytes + offs / keys (pairs: keys, hashed)
  uint64_t count;
  uint64_t seed;
  uint32_t np;
  uint32_t a_ppm;
  const uint32_t* fulcs;
  const uint32_t* pilots;
  const uint32_t* offsets;
  const uint32_t* remap;
  const uint32_t* remap_offsets;
  uint32_t* out;
};

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
// Routing tiles: `threads` (512 or 1024) times `per` records per thread
// (512: 8, 12, 16, 20 or 24; 1024: 4, 6, 8, 10 or 12); out_fmt / in_fmt are
// kFmt* record formats.  A launch takes the largest per, at most the one
// asked for, whose tile fits the workgroup's LDS.
hipError_t launch_hash_route(int kind, int out_fmt, const HashRouteArgs& a, uint32_t threads, uint32_t per,
                             hipStream_t s);
hipError_t launch_split_level(int in_fmt, int out_fmt, const SplitArgs& a, uint32_t parents, uint32_t threads,
                              uint32_t per, hipStream_t s);
// records per thread a split level of `threads` uses when asked for `per`
// (the slices per parent bin are sized for it)
uint32_t split_tile_per(uint32_t threads, uint32_t per, int in_fmt, int out_fmt, uint32_t fanout);
// exclusive scans of counts into offsets[np+1] and, when a_ppm > 0, of
// table_size(count) - count into remap_offsets[np+1]; `totals` holds
// 2 * partition_scan_tiles(np) words of scratch
uint32_t partition_scan_tiles(uint32_t np);
hipError_t launch_partition_offsets(const uint32_t* counts, uint32_t np, uint32_t a_ppm, uint32_t* offsets,
                                    uint32_t* remap_offsets, uint32_t* totals, hipStream_t s);
// waves = partitions (wave64 each) per workgroup: 1, 2, 4 or 8
size_t search_lds_bytes(uint32_t tcap, uint32_t B, uint32_t waves);
hipError_t launch_partition_search(const SearchArgs& a, uint32_t waves, hipStream_t s);
hipError_t launch_eval(int kind, const EvalArgs& a, hipStream_t s);
// per-set exclusive offsets from the global ones: set s's np_s + 1 entries
// start at out + set_part_base[s] + s
hipError_t launch_batch_offsets(const uint32_t* global_offsets, const uint32_t* set_part_base, uint32_t sets,
                                uint32_t total_entries, uint32_t* out, hipStream_t s);
// batched alpha < 1 builds: where partition p's remap entries start in the
// remap output, for p < np: set_remap_base[s] + goff[p] - goff[pb[s]] with
// s the set owning p and goff the global remap offsets
hipError_t launch_batch_remap_bases(const uint32_t* goff, const uint32_t* set_part_base,
                                    const uint32_t* set_remap_base, uint32_t sets, uint32_t np, uint32_t* out,
                                    hipStream_t s);
hipError_t launch_check_bijection(const uint32_t* pos, uint64_t count, uint64_t n, uint32_t* bitmap,
                                  unsigned long long* bad, hipStream_t s);
hipError_t launch_copy16(void* dst, const void* src, uint64_t words, hipStream_t s);
hipError_t launch_splitmix64(uint64_t seed, uint64_t n, uint64_t* out, hipStream_t s);
hipError_t launch_gen_strings(uint64_t seed, uint64_t n, const uint64_t* offs, uint8_t* bytes, hipStream_t s);

}  // namespace phobic
