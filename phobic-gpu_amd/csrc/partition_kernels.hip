// partition_kernels.hip -- hashing and multi-level radix partitioning of keys
// into fixed-capacity partition regions (gfx950).
//
// The reference counts bucket sizes and scatters keys with one global atomic
// per key (bucket_sizes.comp:21-35, redistribute_keys.comp:21-53).  Random
// per-key global atomics sustain only a few tens of G/s on MI355X, so here
// keys are routed to their partition in L levels of fan-out f <= 128: every
// level loads a tile of 512 * per records (per = 8, 12 or 16 per thread),
// sorts it by child bin in LDS, takes one global atomic per (tile, child
// bin) and writes coalesced runs.  Longer runs per bin pay: 512-thread tiles
// of 6144 records split 20% faster than 256-thread tiles of 3072.
//
// Level l's bins cover G_l = f^(L-l) consecutive partitions and hold at most
// G_l * (P + P/2) records, so a full bin implies a partition with more than
// P + P/2 keys (PHOBIC_E_PARTITION_OVERFLOW, the reference's
// partitionMaxSize(), mphf_config.h:61-64).  The last level (G = 1) leaves
// every partition p at records [p*cap, p*cap + count[p]).
//
// Record = lower64 (lower1 | lower2 << 32) + meta32 (partition index inside
// the current bin << 16 | bucket index inside the partition).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_hash.hpp"
#include "kernels.hpp"
#include "xcd.hpp"

namespace phobic {
namespace {

constexpr uint32_t kThreads = 512;

// kPer records per thread, kPer * 256 per tile
template <uint32_t kPer>
struct TileLds {
  static constexpr uint32_t kTile = kThreads * kPer;
  uint64_t lo[kTile];
  uint32_t meta[kTile];
  uint8_t child[kTile];
  // per child bin: records, then (once wave 0 has read them) lim: the
  // tile's record d of the bin goes to output index base + d for d < lim,
  // records from lim on are past the bin's capacity
  uint32_t hist[kMaxFanout];
  uint32_t lstart[kMaxFanout];
  uint64_t base[kMaxFanout];
};

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Routes the block's tile (record k of thread t valid when valid[k],
// `total` valid records) to child bins child_first + child[k] of capacity
// child_cap.  Wave 0 scans the tile's bin histogram and takes the bins'
// global ranges with one atomic each; the scatter into LDS overlaps those
// atomics.
template <uint32_t kPer>
__device__ __forceinline__ void route_tile(TileLds<kPer>& s, const uint64_t (&lo)[kPer], const uint32_t (&meta)[kPer],
                                           const uint32_t (&child)[kPer], const bool (&valid)[kPer],
                                           uint32_t total, uint64_t child_first, uint64_t child_cap,
                                           uint32_t* __restrict__ child_count, uint64_t* __restrict__ out_lo,
                                           uint32_t* __restrict__ out_meta, uint32_t* __restrict__ status) {
  static_assert(kMaxFanout == 128, "wave 0 scans two bins per lane");
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < kMaxFanout; i += kThreads) s.hist[i] = 0;
  __syncthreads();
  uint32_t rank[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++)
    if (valid[k]) rank[k] = atomicAdd(&s.hist[child[k]], 1u);
  __syncthreads();
  uint32_t g0 = 0, g1 = 0, v0 = 0, v1 = 0, l0 = 0;
  if (tid < 64) {  // bins 2 tid and 2 tid + 1 (empty beyond the fan-out)
    v0 = s.hist[2 * tid];
    v1 = s.hist[2 * tid + 1];
    const uint32_t pair = v0 + v1;
    const uint32_t incl = wave_incl_scan(pair);
    l0 = incl - pair;
    s.lstart[2 * tid] = l0;
    s.lstart[2 * tid + 1] = l0 + v0;
    if (v0) g0 = atomicAdd(&child_count[child_first + 2 * tid], v0);
    if (v1) g1 = atomicAdd(&child_count[child_first + 2 * tid + 1], v1);
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    if (valid[k]) {
      const uint32_t d = s.lstart[child[k]] + rank[k];
      s.lo[d] = lo[k];
      s.meta[d] = meta[k];
      s.child[d] = (uint8_t)child[k];
    }
  }
  if (tid < 64) {
    // record d of bin c goes to bin slot g_c + (d - lstart_c), inside the
    // bin while that is below child_cap
    auto bin = [&](uint32_t c, uint32_t g, uint32_t v, uint32_t l) {
      s.base[c] = (child_first + c) * child_cap + g - l;
      s.hist[c] = l + (uint32_t)min((uint64_t)v, child_cap > g ? child_cap - g : 0);
    };
    bin(2 * tid, g0, v0, l0);
    bin(2 * tid + 1, g1, v1, l0 + v0);
  }
  __syncthreads();
  bool over = false;
  for (uint32_t d = tid; d < total; d += kThreads) {
    const uint32_t c = s.child[d];
    if (d < s.hist[c]) {
      const uint64_t at = s.base[c] + d;
      out_lo[at] = s.lo[d];
      out_meta[at] = s.meta[d];
    } else {
      over = true;
    }
  }
  if (over) atomicOr(status, 1u << kErrPartitionOverflow);
  __syncthreads();
}

template <int KIND>
__device__ __forceinline__ U128 hash_key(const KeySource& src, uint64_t i, uint64_t seed) {
  if constexpr (KIND == kKindU64) {
    return xxh3_128_u64(src.u64[i], seed);
  } else if constexpr (KIND == kKindBytes) {
    const uint64_t a = src.offs[i], b = src.offs[i + 1];
    return xxh3_128_bytes(src.bytes + a, b - a, seed);
  } else {
    const uint4 k = src.keys[i];
    return U128{(uint64_t)k.x | ((uint64_t)k.y << 32), (uint64_t)k.z | ((uint64_t)k.w << 32)};
  }
}

}  // namespace

// Level 1: hash raw keys, map to (partition, bucket) and route to the
// level-1 bins (partitions [c*G, (c+1)*G)).
template <int KIND, uint32_t kPer>
__global__ __launch_bounds__(kThreads) void k_hash_route(HashRouteArgs a) {
  constexpr uint32_t kTile = kThreads * kPer;
  __shared__ TileLds<kPer> s;
  __shared__ uint32_t fl[kFulcs];
  for (uint32_t i = threadIdx.x; i < kFulcs; i += kThreads) fl[i] = a.fulcs[i];
  __syncthreads();
  const uint64_t ntiles = (a.n + kTile - 1) / kTile;
  const float rG = 1.0f / (float)a.G;
  // u64 keys: the next tile's keys are loaded while this tile is routed
  uint64_t kv[kPer];
  auto load_u64 = [&](uint64_t t, uint64_t (&dst)[kPer]) {
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint64_t i = t * kTile + k * kThreads + threadIdx.x;
      dst[k] = i < a.n ? a.src.u64[i] : 0;
    }
  };
  if (KIND == kKindU64 && blockIdx.x < ntiles) load_u64(blockIdx.x, kv);
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    uint32_t set = 0;
    if (a.sets) {  // set of this thread's first key: last s with set_key_off[s] <= i
      const uint64_t i0 = t * kTile + threadIdx.x;
      uint32_t lo_s = 0, hi_s = a.sets;
      while (hi_s - lo_s > 1) {
        const uint32_t mid = (lo_s + hi_s) / 2;
        if (a.set_key_off[mid] <= i0) lo_s = mid;
        else hi_s = mid;
      }
      set = lo_s;
    }
    // first key offset of the next set, and the set's partition range
    uint64_t set_end = ~0ull;
    uint32_t pb = 0, pn = a.np;
    if (a.sets) {
      if (set + 1 < a.sets) set_end = a.set_key_off[set + 1];
      pb = a.set_part_base[set];
      pn = a.set_part_base[set + 1] - pb;
    }
    // issued after every other load this tile waits for: vector memory
    // counters retire in order, so a wait on a later load would wait for
    // these too
    uint64_t nk[kPer];
    if (KIND == kKindU64 && t + gridDim.x < ntiles) load_u64(t + gridDim.x, nk);
    uint64_t lo[kPer];
    uint32_t meta[kPer], child[kPer];
    bool valid[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint64_t i = t * kTile + k * kThreads + threadIdx.x;
      valid[k] = i < a.n;
      lo[k] = 0;
      meta[k] = 0;
      child[k] = 0;
      if (valid[k]) {
        const U128 h = KIND == kKindU64 ? xxh3_128_u64(kv[k], a.seed) : hash_key<KIND>(a.src, i, a.seed);
        uint32_t p;
        if (a.sets && i >= set_end) {  // a later set (loads only at a boundary)
          do {
            set++;
            set_end = set + 1 < a.sets ? a.set_key_off[set + 1] : ~0ull;
          } while (i >= set_end);
          pb = a.set_part_base[set];
          pn = a.set_part_base[set + 1] - pb;
        }
        p = pb + partition_of((uint32_t)h.lo, pn);
        const uint32_t b = bucket_of((uint32_t)(h.lo >> 32), fl);
        // level-1 bin c = p / G and r = p mod G: the f32 quotient estimate
        // is within one of c (c < 2^8, relative error below 2^-21), G <
        // 2^16, so q * G is a 24-bit product and one step corrects it
        uint32_t c = (uint32_t)(__uint2float_rn(p) * rG);
        int32_t r = (int32_t)(p - __umul24(c, a.G));
        if (r < 0) {
          c--;
          r += (int32_t)a.G;
        } else if (r >= (int32_t)a.G) {
          c++;
          r -= (int32_t)a.G;
        }
        lo[k] = h.hi;
        meta[k] = ((uint32_t)r << 16) | b;
        child[k] = c;
      }
    }
    const uint32_t total = (uint32_t)min((uint64_t)kTile, a.n - t * kTile);  // valid records: a prefix
    route_tile(s, lo, meta, child, valid, total, 0, (uint64_t)a.G * a.cap, a.child_count, a.out_lo, a.out_meta,
               a.status);
    if (KIND == kKindU64) {
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) kv[k] = nk[k];
    }
  }
}

// Level l > 1: split every parent bin (G_parent partitions) into `fanout`
// child bins of G_child partitions.  Block (b, s) handles tile s of parent b.
template <uint32_t kPer>
__global__ __launch_bounds__(kThreads) void k_split_level(SplitArgs a) {
  constexpr uint32_t kTile = kThreads * kPer;
  __shared__ TileLds<kPer> s;
  // neighbouring tiles of a parent on one XCD: their runs in each child bin
  // are adjacent, so the lines they share are merged in one L2
  const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t parent = bid / a.slices;
  const uint64_t tile0 = (uint64_t)(bid % a.slices) * kTile;
  const uint32_t cnt = min(a.parent_count[parent], (uint32_t)min((uint64_t)0xFFFFFFFFu, a.parent_cap));
  if (tile0 >= cnt) return;
  const uint64_t in0 = (uint64_t)parent * a.parent_cap;
  uint64_t lo[kPer];
  uint32_t m[kPer], meta[kPer], child[kPer];
  bool valid[kPer];
  // all loads first (past the end: record tile0, which exists), then the
  // child bins: pl / G_child = umulhi(pl, floor(2^32 / G_child) + 1) for
  // pl < 2^16 and 2 <= G_child < 2^16 (the error stays below 1 / G_child)
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    const uint64_t i = tile0 + k * kThreads + threadIdx.x;
    valid[k] = i < cnt;
    const uint64_t at = in0 + (valid[k] ? i : tile0);
    lo[k] = a.in_lo[at];
    m[k] = a.in_meta[at];
  }
  const uint32_t G = a.G_child;
  const uint32_t inv = G > 1 ? (uint32_t)(0xFFFFFFFFu / G) + 1 : 0;
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    const uint32_t pl = m[k] >> 16;
    const uint32_t c = G > 1 ? __umulhi(pl, inv) : pl;
    meta[k] = ((pl - c * G) << 16) | (m[k] & 0xFFFFu);
    child[k] = valid[k] ? c : 0;
  }
  const uint32_t total = (uint32_t)min((uint64_t)kTile, cnt - tile0);  // valid records: a prefix
  route_tile(s, lo, meta, child, valid, total, (uint64_t)parent * a.fanout, (uint64_t)a.G_child * a.cap, a.child_count,
             a.out_lo, a.out_meta, a.status);
}

uint32_t route_tile_records(uint32_t per) { return kThreads * (per == 8 ? 8 : per == 12 ? 12 : 16); }

template <uint32_t kPer>
static hipError_t hash_route_per(int kind, const HashRouteArgs& a, hipStream_t s) {
  constexpr uint32_t kTile = kThreads * kPer;
  const uint64_t ntiles = (a.n + kTile - 1) / kTile;
  const uint32_t grid = (uint32_t)(ntiles < 4096 ? ntiles : 4096);
  switch (kind) {
    case kKindU64: hipLaunchKernelGGL((k_hash_route<kKindU64, kPer>), dim3(grid), dim3(kThreads), 0, s, a); break;
    case kKindBytes: hipLaunchKernelGGL((k_hash_route<kKindBytes, kPer>), dim3(grid), dim3(kThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_hash_route<kKindKey, kPer>), dim3(grid), dim3(kThreads), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_hash_route(int kind, const HashRouteArgs& a, uint32_t per, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  // A string tile waits on two dependent loads per key (offsets, then
  // bytes): tiles of 8 records per thread, two workgroups per CU, hide that
  // better than longer runs per bin pay (5.2 against 5.4 ms for 2e8
  // strings with 16, 6.1 ms with 20).
  if (kind == kKindBytes) per = 8;
  return per == 8 ? hash_route_per<8>(kind, a, s)
       : per == 12 ? hash_route_per<12>(kind, a, s)
       : per == 20 ? hash_route_per<20>(kind, a, s)
                   : hash_route_per<16>(kind, a, s);
}

hipError_t launch_split_level(const SplitArgs& a, uint32_t parents, uint32_t per, hipStream_t s) {
  if (parents == 0) return hipSuccess;
  if (per == 8)
    hipLaunchKernelGGL(k_split_level<8>, dim3(parents * a.slices), dim3(kThreads), 0, s, a);
  else if (per == 12)
    hipLaunchKernelGGL(k_split_level<12>, dim3(parents * a.slices), dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL(k_split_level<16>, dim3(parents * a.slices), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace phobic
