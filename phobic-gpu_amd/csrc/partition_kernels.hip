// partition_kernels.hip -- hashing and multi-level radix partitioning of keys
// into fixed-capacity partition regions (gfx950).
//
// The reference counts bucket sizes and scatters keys with one global atomic
// per key (bucket_sizes.comp:21-35, redistribute_keys.comp:21-53).  Random
// per-key global atomics sustain only a few tens of G/s on MI355X, so here
// keys are routed to their partition in L levels of fan-out f <= 1024: every
// level loads a tile of 512 * per records, sorts it by child bin in LDS,
// takes one global atomic per (tile, child bin) and writes coalesced runs.
//
// Level l's bins cover G_l = f^(L-l) consecutive partitions and hold at most
// G_l * (P + P/2) records, so a full bin implies a partition with more than
// P + P/2 keys (PHOBIC_E_PARTITION_OVERFLOW, the reference's
// partitionMaxSize(), mphf_config.h:61-64).  The last level (G = 1) leaves
// every partition p at records [p*cap, p*cap + count[p]).
//
// What a level writes (kFmt*):
//   plain u64 builds carry the 8-byte key itself through every level but the
//   last (each level hashes it again: XXH3 of 8 bytes is a few multiplies,
//   cheaper than moving a hash), and the last level writes
//   lower64 + the 16-bit bucket: 8 + 8 bytes per key per inner level and
//   8 + 10 at the last one;
//   byte strings, pre-hashed keys and batched builds hash once at level 1 and
//   carry lower64 + meta32 (partition inside the bin << 16 | bucket) to the
//   last level, which writes lower64 + bucket16.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_hash.hpp"
#include "kernels.hpp"
#include "xcd.hpp"

namespace phobic {
namespace {

// Routing workgroups have kT = 512 or 1024 threads.
constexpr uint32_t kMaxWaves = 16;
constexpr uint32_t kLdsLimit = 160 * 1024;  // LDS of one gfx950 workgroup
static_assert(kMaxFanout <= 2 * 512, "the bin scan holds two bins per thread");

__host__ __device__ constexpr uint32_t meta_bytes(int fmt) {
  return fmt == kFmtRec ? 4u : fmt == kFmtFinal ? 2u : 0u;
}

template <int FMT>
struct Meta {
  using T = uint32_t;
};
template <>
struct Meta<kFmtFinal> {
  using T = uint16_t;
};

// Dynamic LDS of a routing tile of `tile` records into F bins (byte offsets):
//   lo[tile] u64 | meta[tile] (4, 2 or 0 bytes) | child[tile] u16 |
//   base[F] u64 | hist[F] u32 | lstart[F] u32 | wsum[kMaxWaves] u32 |
//   fulcrums (2048 u32) when the level maps keys to buckets
struct TileLayout {
  uint32_t lo, meta, child, base, hist, lstart, wsum, fulcs, bytes;
};
__host__ __device__ inline TileLayout tile_layout(uint32_t tile, int fmt, uint32_t F, bool fulcs) {
  TileLayout l;
  uint32_t at = 0;
  auto take = [&at](uint32_t bytes, uint32_t align) {
    at = (at + align - 1) & ~(align - 1);
    const uint32_t r = at;
    at += bytes;
    return r;
  };
  l.lo = take(tile * 8, 16);
  l.meta = take(tile * meta_bytes(fmt), 4);
  l.child = take(tile * 2, 4);
  l.base = take(F * 8, 8);
  l.hist = take(F * 4, 4);
  l.lstart = take(F * 4, 4);
  l.wsum = take(kMaxWaves * 4, 4);
  l.fulcs = take(fulcs ? kFulcs * 4 : 0, 4);
  l.bytes = at;
  return l;
}

template <int FMT>
struct Tile {
  uint64_t* lo;
  typename Meta<FMT>::T* meta;
  uint16_t* child;
  uint64_t* base;
  uint32_t* hist;  // per bin: records of the tile, then the bin's limit
  uint32_t* lstart;
  uint32_t* wsum;
  uint32_t* fulcs;
};

template <int FMT>
__device__ __forceinline__ Tile<FMT> tile_at(uint8_t* lds, const TileLayout& l) {
  Tile<FMT> t;
  t.lo = reinterpret_cast<uint64_t*>(lds + l.lo);
  t.meta = reinterpret_cast<typename Meta<FMT>::T*>(lds + l.meta);
  t.child = reinterpret_cast<uint16_t*>(lds + l.child);
  t.base = reinterpret_cast<uint64_t*>(lds + l.base);
  t.hist = reinterpret_cast<uint32_t*>(lds + l.hist);
  t.lstart = reinterpret_cast<uint32_t*>(lds + l.lstart);
  t.wsum = reinterpret_cast<uint32_t*>(lds + l.wsum);
  t.fulcs = reinterpret_cast<uint32_t*>(lds + l.fulcs);
  return t;
}

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
// `total` valid records, which form a prefix of the tile) to child bins
// child_first + child[k] (child[k] < F) of capacity child_cap.  The tile's
// bin histogram is scanned by the whole block, two bins per thread; each
// thread takes its bins' global ranges with one atomic each, in flight
// while the block finishes the scan and scatters the tile into LDS in bin
// order.  Records then leave in runs, one per bin.
template <uint32_t kT, uint32_t kPer, int FMT>
__device__ __forceinline__ void route_tile(const Tile<FMT>& s, const uint64_t (&lo)[kPer], const uint32_t (&meta)[kPer],
                                           const uint32_t (&child)[kPer], const bool (&valid)[kPer], uint32_t total,
                                           uint32_t F, uint64_t child_first, uint64_t child_cap,
                                           uint32_t* __restrict__ child_count, uint64_t* __restrict__ out_lo,
                                           void* __restrict__ out_meta, uint32_t* __restrict__ status) {
  using MT = typename Meta<FMT>::T;
  constexpr bool kMeta = FMT != kFmtRaw;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (uint32_t i = tid; i < F; i += kT) s.hist[i] = 0;
  __syncthreads();
  uint32_t rank[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++)
    if (valid[k]) rank[k] = atomicAdd(&s.hist[child[k]], 1u);
  __syncthreads();
  const uint32_t c0 = 2 * tid, c1 = c0 + 1;
  const uint32_t v0 = c0 < F ? s.hist[c0] : 0, v1 = c1 < F ? s.hist[c1] : 0;
  const uint32_t pair = v0 + v1;
  const uint32_t incl = wave_incl_scan(pair);
  if (lane == 63) s.wsum[wid] = incl;
  uint32_t g0 = 0, g1 = 0;
  if (v0) g0 = atomicAdd(&child_count[child_first + c0], v0);
  if (v1) g1 = atomicAdd(&child_count[child_first + c1], v1);
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wid; w++) before += s.wsum[w];
  const uint32_t l0 = before + incl - pair;
  if (c0 < F) s.lstart[c0] = l0;
  if (c1 < F) s.lstart[c1] = l0 + v0;
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    if (valid[k]) {
      const uint32_t d = s.lstart[child[k]] + rank[k];
      s.lo[d] = lo[k];
      if constexpr (kMeta) s.meta[d] = (MT)meta[k];
      s.child[d] = (uint16_t)child[k];
    }
  }
  // record d of bin c goes to bin slot g_c + (d - lstart_c), inside the bin
  // while that is below child_cap
  auto bin = [&](uint32_t c, uint32_t g, uint32_t v, uint32_t l) {
    s.base[c] = (child_first + c) * child_cap + g - l;
    s.hist[c] = l + (uint32_t)min((uint64_t)v, child_cap > g ? child_cap - g : 0);
  };
  if (c0 < F) bin(c0, g0, v0, l0);
  if (c1 < F) bin(c1, g1, v1, l0 + v0);
  __syncthreads();
  bool over = false;
  for (uint32_t d = tid; d < total; d += kT) {
    const uint32_t c = s.child[d];
    if (d < s.hist[c]) {
      const uint64_t at = s.base[c] + d;
      out_lo[at] = s.lo[d];
      if constexpr (kMeta) static_cast<MT*>(out_meta)[at] = s.meta[d];
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

// Level 1: hash raw keys, map each to its partition (and, unless the level
// writes raw keys, its bucket) and route it to the level-1 bins (partitions
// [c*G, (c+1)*G)).
template <int KIND, uint32_t kT, uint32_t kPer, int OUT>
__global__ __launch_bounds__(kT) void k_hash_route(HashRouteArgs a) {
  constexpr uint32_t kTile = kT * kPer;
  constexpr bool kBuckets = OUT != kFmtRaw;
  static_assert(OUT != kFmtRaw || KIND == kKindU64, "only u64 keys travel raw");
  extern __shared__ __align__(16) uint8_t lds[];
  const TileLayout lay = tile_layout(kTile, OUT, a.fanout, kBuckets);
  const Tile<OUT> s = tile_at<OUT>(lds, lay);
  if constexpr (kBuckets) {
    for (uint32_t i = threadIdx.x; i < kFulcs; i += kT) s.fulcs[i] = a.fulcs[i];
    __syncthreads();
  }
  const uint64_t ntiles = (a.n + kTile - 1) / kTile;
  const float rG = 1.0f / (float)a.G;
  // u64 keys: the next tile's keys are loaded while this tile is routed
  uint64_t kv[kPer];
  auto load_u64 = [&](uint64_t t, uint64_t (&dst)[kPer]) {
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint64_t i = t * kTile + k * kT + threadIdx.x;
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
      const uint64_t i = t * kTile + k * kT + threadIdx.x;
      valid[k] = i < a.n;
      lo[k] = 0;
      meta[k] = 0;
      child[k] = 0;
      if (valid[k]) {
        const U128 h = KIND == kKindU64 ? xxh3_128_u64(kv[k], a.seed) : hash_key<KIND>(a.src, i, a.seed);
        if (a.sets && i >= set_end) {  // a later set (loads only at a boundary)
          do {
            set++;
            set_end = set + 1 < a.sets ? a.set_key_off[set + 1] : ~0ull;
          } while (i >= set_end);
          pb = a.set_part_base[set];
          pn = a.set_part_base[set + 1] - pb;
        }
        const uint32_t p = pb + partition_of((uint32_t)h.lo, pn);
        // level-1 bin c = p / G and r = p mod G: the f32 quotient estimate
        // is within one of c (c < 2^11, relative error below 2^-22), G <
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
        if constexpr (OUT == kFmtRaw) {
          lo[k] = kv[k];
        } else {
          lo[k] = h.hi;
          const uint32_t b = bucket_of((uint32_t)(h.lo >> 32), s.fulcs);
          meta[k] = OUT == kFmtRec ? ((uint32_t)r << 16) | b : b;
        }
        child[k] = min(c, a.fanout - 1);
      }
    }
    const uint32_t total = (uint32_t)min((uint64_t)kTile, a.n - t * kTile);  // valid records: a prefix
    route_tile<kT, kPer, OUT>(s, lo, meta, child, valid, total, a.fanout, 0, (uint64_t)a.G * a.cap, a.child_count,
                          a.out_lo, a.out_meta, a.status);
    if (KIND == kKindU64) {
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) kv[k] = nk[k];
    }
  }
}

// Level l > 1: split every parent bin (G_parent partitions) into `fanout`
// child bins of G_child partitions.  Block (b, s) handles tile s of parent b.
template <uint32_t kT, uint32_t kPer, int IN, int OUT>
__global__ __launch_bounds__(kT) void k_split_level(SplitArgs a) {
  constexpr uint32_t kTile = kT * kPer;
  constexpr bool kBuckets = IN == kFmtRaw && OUT == kFmtFinal;
  static_assert(IN != kFmtRaw || OUT != kFmtRec, "raw keys go to raw or final records");
  static_assert(IN != kFmtRec || OUT != kFmtRaw, "hashed records cannot become keys again");
  extern __shared__ __align__(16) uint8_t lds[];
  // neighbouring tiles of a parent on one XCD: their runs in each child bin
  // are adjacent, so the lines they share are merged in one L2
  const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t parent = bid / a.slices;
  const uint64_t tile0 = (uint64_t)(bid % a.slices) * kTile;
  const uint32_t cnt = min(a.parent_count[parent], (uint32_t)min((uint64_t)0xFFFFFFFFu, a.parent_cap));
  if (tile0 >= cnt) return;  // the whole block leaves together
  const TileLayout lay = tile_layout(kTile, OUT, a.fanout, kBuckets);
  const Tile<OUT> s = tile_at<OUT>(lds, lay);
  const uint64_t in0 = (uint64_t)parent * a.parent_cap;
  uint64_t lo[kPer];
  uint32_t m[kPer], meta[kPer], child[kPer];
  bool valid[kPer];
  // all loads first (past the end: record tile0, which exists)
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    const uint64_t i = tile0 + k * kT + threadIdx.x;
    valid[k] = i < cnt;
    const uint64_t at = in0 + (valid[k] ? i : tile0);
    lo[k] = a.in_lo[at];
    m[k] = IN == kFmtRec ? static_cast<const uint32_t*>(a.in_meta)[at] : 0u;
  }
  if constexpr (kBuckets) {
    for (uint32_t i = threadIdx.x; i < kFulcs; i += kT) s.fulcs[i] = a.fulcs[i];
    __syncthreads();
  }
  // child bin pl / G_child = umulhi(pl, floor(2^32 / G_child) + 1) for
  // pl < 2^16 and 2 <= G_child < 2^16 (the error stays below 1 / G_child)
  const uint32_t G = a.G_child;
  const uint32_t inv = G > 1 ? (uint32_t)(0xFFFFFFFFu / G) + 1 : 0;
  const uint32_t first = parent * a.G_parent;  // the parent's first partition
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    uint32_t pl;
    if constexpr (IN == kFmtRaw) {
      const U128 h = xxh3_128_u64(lo[k], a.seed);
      pl = partition_of((uint32_t)h.lo, a.np) - first;
      if constexpr (OUT == kFmtFinal) {
        meta[k] = bucket_of((uint32_t)(h.lo >> 32), s.fulcs);
        lo[k] = h.hi;
      }
    } else {
      pl = m[k] >> 16;
    }
    const uint32_t c = G > 1 ? __umulhi(pl, inv) : pl;
    if constexpr (IN == kFmtRec) meta[k] = OUT == kFmtRec ? ((pl - c * G) << 16) | (m[k] & 0xFFFFu) : m[k] & 0xFFFFu;
    child[k] = valid[k] ? min(c, a.fanout - 1) : 0;
  }
  const uint32_t total = (uint32_t)min((uint64_t)kTile, cnt - tile0);  // valid records: a prefix
  route_tile<kT, kPer, OUT>(s, lo, meta, child, valid, total, a.fanout, (uint64_t)parent * a.fanout,
                        (uint64_t)a.G_child * a.cap, a.child_count, a.out_lo, a.out_meta, a.status);
}

namespace {

size_t route_lds(uint32_t tile, int out_fmt, uint32_t F, bool fulcs) {
  return tile_layout(tile, out_fmt, F, fulcs).bytes;
}

// Tile shapes (threads, records per thread) a routing launch can take.
bool valid_shape(uint32_t threads, uint32_t per) {
  return threads == 512 ? (per == 8 || per == 12 || per == 16 || per == 20 || per == 24)
                        : threads == 1024 && (per == 4 || per == 6 || per == 8 || per == 10 || per == 12);
}

// The largest records per thread, at most `per`, whose tile of `threads`
// threads fits the workgroup's LDS.
uint32_t fitting_per(uint32_t threads, uint32_t per, int out_fmt, uint32_t F, bool fulcs) {
  for (uint32_t p : {24u, 20u, 16u, 12u, 10u, 8u, 6u, 4u})
    if (p <= per && valid_shape(threads, p) && route_lds(threads * p, out_fmt, F, fulcs) <= kLdsLimit) return p;
  return threads == 512 ? 8 : 4;
}

template <int KIND, uint32_t kT, uint32_t kPer, int OUT>
hipError_t hash_route_go(const HashRouteArgs& a, hipStream_t s) {
  constexpr uint32_t kTile = kT * kPer;
  const uint64_t ntiles = (a.n + kTile - 1) / kTile;
  const uint32_t grid = (uint32_t)(ntiles < 4096 ? ntiles : 4096);
  const size_t lds = route_lds(kTile, OUT, a.fanout, OUT != kFmtRaw);
  hipLaunchKernelGGL((k_hash_route<KIND, kT, kPer, OUT>), dim3(grid), dim3(kT), lds, s, a);
  return hipGetLastError();
}

template <int KIND, int OUT>
hipError_t hash_route_shape(uint32_t threads, uint32_t per, const HashRouteArgs& a, hipStream_t s) {
  if (threads == 1024) {
    switch (per) {
      case 12: return hash_route_go<KIND, 1024, 12, OUT>(a, s);
      case 10: return hash_route_go<KIND, 1024, 10, OUT>(a, s);
      case 8: return hash_route_go<KIND, 1024, 8, OUT>(a, s);
      case 6: return hash_route_go<KIND, 1024, 6, OUT>(a, s);
      default: return hash_route_go<KIND, 1024, 4, OUT>(a, s);
    }
  }
  switch (per) {
    case 24: return hash_route_go<KIND, 512, 24, OUT>(a, s);
    case 20: return hash_route_go<KIND, 512, 20, OUT>(a, s);
    case 16: return hash_route_go<KIND, 512, 16, OUT>(a, s);
    case 12: return hash_route_go<KIND, 512, 12, OUT>(a, s);
    default: return hash_route_go<KIND, 512, 8, OUT>(a, s);
  }
}

template <uint32_t kT, uint32_t kPer, int IN, int OUT>
hipError_t split_go(const SplitArgs& a, uint32_t parents, hipStream_t s) {
  const size_t lds = route_lds(kT * kPer, OUT, a.fanout, IN == kFmtRaw && OUT == kFmtFinal);
  hipLaunchKernelGGL((k_split_level<kT, kPer, IN, OUT>), dim3(parents * a.slices), dim3(kT), lds, s, a);
  return hipGetLastError();
}

template <int IN, int OUT>
hipError_t split_shape(uint32_t threads, uint32_t per, const SplitArgs& a, uint32_t parents, hipStream_t s) {
  if (threads == 1024) {
    switch (per) {
      case 12: return split_go<1024, 12, IN, OUT>(a, parents, s);
      case 10: return split_go<1024, 10, IN, OUT>(a, parents, s);
      case 8: return split_go<1024, 8, IN, OUT>(a, parents, s);
      case 6: return split_go<1024, 6, IN, OUT>(a, parents, s);
      default: return split_go<1024, 4, IN, OUT>(a, parents, s);
    }
  }
  switch (per) {
    case 24: return split_go<512, 24, IN, OUT>(a, parents, s);
    case 20: return split_go<512, 20, IN, OUT>(a, parents, s);
    case 16: return split_go<512, 16, IN, OUT>(a, parents, s);
    case 12: return split_go<512, 12, IN, OUT>(a, parents, s);
    default: return split_go<512, 8, IN, OUT>(a, parents, s);
  }
}

}  // namespace

uint32_t split_tile_per(uint32_t threads, uint32_t per, int in_fmt, int out_fmt, uint32_t fanout) {
  return fitting_per(threads, per, out_fmt, fanout, in_fmt == kFmtRaw && out_fmt == kFmtFinal);
}

hipError_t launch_hash_route(int kind, int out_fmt, const HashRouteArgs& a, uint32_t threads, uint32_t per,
                             hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.fanout == 0 || a.fanout > kMaxFanout || (out_fmt == kFmtRaw && (kind != kKindU64 || a.sets)) ||
      !(threads == 512 || threads == 1024))
    return hipErrorInvalidValue;
  // A string tile waits on two dependent loads per key (offsets, then
  // bytes): tiles of 4096 records hide that better than longer runs per bin
  // pay (5.2 against 5.4 ms for 2e8 strings with 8192, 6.1 ms with 10,240).
  if (kind == kKindBytes) {
    threads = 512;
    per = 8;
  }
  per = fitting_per(threads, per, out_fmt, a.fanout, out_fmt != kFmtRaw);
  switch (kind) {
    case kKindU64:
      if (out_fmt == kFmtRaw) return hash_route_shape<kKindU64, kFmtRaw>(threads, per, a, s);
      if (out_fmt == kFmtRec) return hash_route_shape<kKindU64, kFmtRec>(threads, per, a, s);
      return hash_route_shape<kKindU64, kFmtFinal>(threads, per, a, s);
    case kKindBytes:
      if (out_fmt == kFmtRec) return hash_route_go<kKindBytes, 512, 8, kFmtRec>(a, s);
      return hash_route_go<kKindBytes, 512, 8, kFmtFinal>(a, s);
    default:
      if (out_fmt == kFmtRec) return hash_route_shape<kKindKey, kFmtRec>(threads, per, a, s);
      return hash_route_shape<kKindKey, kFmtFinal>(threads, per, a, s);
  }
}

hipError_t launch_split_level(int in_fmt, int out_fmt, const SplitArgs& a, uint32_t parents, uint32_t threads,
                              uint32_t per, hipStream_t s) {
  if (parents == 0) return hipSuccess;
  if (a.fanout == 0 || a.fanout > kMaxFanout || !(threads == 512 || threads == 1024)) return hipErrorInvalidValue;
  // `per` must be what split_tile_per chose: the slices were sized for it
  if (per != split_tile_per(threads, per, in_fmt, out_fmt, a.fanout)) return hipErrorInvalidValue;
  if (in_fmt == kFmtRaw && out_fmt == kFmtRaw) return split_shape<kFmtRaw, kFmtRaw>(threads, per, a, parents, s);
  if (in_fmt == kFmtRaw && out_fmt == kFmtFinal) return split_shape<kFmtRaw, kFmtFinal>(threads, per, a, parents, s);
  if (in_fmt == kFmtRec && out_fmt == kFmtRec) return split_shape<kFmtRec, kFmtRec>(threads, per, a, parents, s);
  if (in_fmt == kFmtRec && out_fmt == kFmtFinal) return split_shape<kFmtRec, kFmtFinal>(threads, per, a, parents, s);
  return hipErrorInvalidValue;
}

}  // namespace phobic
