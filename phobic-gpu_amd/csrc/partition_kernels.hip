// partition_kernels.hip -- hashing and multi-level radix partitioning of keys
// into fixed-capacity partition regions (gfx950).
//
// The reference counts bucket sizes and scatters keys with one global atomic
// per key (bucket_sizes.comp:21-35, redistribute_keys.comp:21-53).  Random
// per-key global atomics sustain only a few tens of G/s on MI355X, so here
// keys are routed to their partition in L levels of fan-out f <= 1024: every
// level loads a tile of threads * per records, sorts it by child bin in LDS,
// takes one global atomic per (tile, child bin) and writes coalesced runs.
// Plain u64 builds at level 1 and every split with at most 1024 parent bins
// use the software-pipelined persistent kernels (k_route_keys, k_split_pipe,
// described below); strings, pre-hashed keys, batched sets and larger splits
// use the per-tile kernels (k_hash_route, k_split_level).
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

#include <algorithm>
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
//   base[F] u32 | hist[F] u32 | lstart[F] u32 | wsum[kMaxWaves] u32 |
//   `tabs` fulcrum tables (2048 u32 each) when the level maps keys to
//   buckets (one, or one per lambda of a batch of several configurations)
struct TileLayout {
  uint32_t lo, meta, child, base, hist, lstart, wsum, fulcs, bytes;
};
__host__ __device__ inline TileLayout tile_layout(uint32_t tile, int fmt, uint32_t F, uint32_t tabs) {
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
  l.base = take(F * 4, 4);
  l.hist = take(F * 4, 4);
  l.lstart = take(F * 4, 4);
  l.wsum = take(kMaxWaves * 4, 4);
  l.fulcs = take(tabs * kFulcs * 4, 4);
  l.bytes = at;
  return l;
}

template <int FMT>
struct Tile {
  uint64_t* lo;
  typename Meta<FMT>::T* meta;
  uint16_t* child;
  uint32_t* base;  // per bin: its first slot for this tile minus its first tile record (mod 2^32)
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
  t.base = reinterpret_cast<uint32_t*>(lds + l.base);
  t.hist = reinterpret_cast<uint32_t*>(lds + l.hist);
  t.lstart = reinterpret_cast<uint32_t*>(lds + l.lstart);
  t.wsum = reinterpret_cast<uint32_t*>(lds + l.wsum);
  t.fulcs = reinterpret_cast<uint32_t*>(lds + l.fulcs);
  return t;
}

// Raw buffer loads (gfx9 resource word 3: 32-bit data format).  A buffer
// resource holds a wave-uniform base and a size in bytes; a lane reads at a
// 32-bit byte offset plus a scalar offset, and a read past the size returns
// 0 instead of touching memory.  The last, partial tile therefore needs no
// clamped indices, and no load sits in a per-lane branch (a load under a
// branch makes the compiler wait for every load in flight where the branch
// joins).
constexpr int kBufWord3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kBufWord3);
}
__device__ __forceinline__ uint64_t buf_load_u64(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
  return (uint64_t)v[0] | ((uint64_t)v[1] << 32);
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
  // the returned ranges are waited for where they are used, after the scan
  // (they stay 32-bit until then: a widening copy inside these branches
  // would wait for them here)
  uint32_t g0 = 0, g1 = 0;
  if (c0 < F) g0 = atomicAdd(&child_count[child_first + c0], v0);
  if (c1 < F) g1 = atomicAdd(&child_count[child_first + c1], v1);
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
  // while that is below child_cap, which the launchers keep below 2^32, so
  // the slot is base[c] + d in 32-bit arithmetic
  auto bin = [&](uint32_t c, uint32_t g, uint32_t v, uint32_t l) {
    s.base[c] = g - l;
    s.hist[c] = l + (uint32_t)min((uint64_t)v, child_cap > g ? child_cap - g : 0);
  };
  if (c0 < F) bin(c0, g0, v0, l0);
  if (c1 < F) bin(c1, g1, v1, l0 + v0);
  __syncthreads();
  // Records leave in chunks of four per thread: the LDS reads of a chunk
  // are unconditional (a slot past `total` holds a stale or unwritten child,
  // clamped into the bins), so they are in flight together, and only the
  // stores are predicated.
  // The thread index is hidden from the optimiser here: otherwise it hoists
  // the kPer addresses out_lo + 8 * (k * kT + tid) out of the tile loop and
  // keeps them live (24 or more registers, spilled).
  uint32_t tid_o = tid;
  asm volatile("" : "+v"(tid_o));
  bool over = false;
  constexpr uint32_t kChunk = kPer % 4 == 0 ? 4 : 2;
  static_assert(kPer % kChunk == 0, "records per thread come in whole chunks");
#pragma unroll
  for (uint32_t k0 = 0; k0 < kPer; k0 += kChunk) {
    uint32_t c[kChunk], lim[kChunk];
    uint64_t at[kChunk], v[kChunk];
    MT mv[kChunk];
#pragma unroll
    for (uint32_t j = 0; j < kChunk; j++) c[j] = min((uint32_t)s.child[(k0 + j) * kT + tid_o], F - 1);
#pragma unroll
    for (uint32_t j = 0; j < kChunk; j++) {
      const uint32_t d = (k0 + j) * kT + tid_o;
      lim[j] = min(s.hist[c[j]], total);
      at[j] = (child_first + c[j]) * child_cap + (uint32_t)(s.base[c[j]] + d);
      v[j] = s.lo[d];
      if constexpr (kMeta) mv[j] = s.meta[d];
    }
#pragma unroll
    for (uint32_t j = 0; j < kChunk; j++) {
      const uint32_t d = (k0 + j) * kT + tid_o;
      if (d < lim[j]) {
        out_lo[at[j]] = v[j];
        if constexpr (kMeta) static_cast<MT*>(out_meta)[at[j]] = mv[j];
      } else if (d < total) {
        over = true;
      }
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
//
// SETS (batched builds) is a template flag: the set lookup loads inside the
// key loop, and the wait for those loads would also wait for the next
// tile's keys, which are in flight by then (vector memory counters retire
// in order).  Plain builds compile it out.
template <int KIND, uint32_t kT, uint32_t kPer, int OUT, bool SETS>
__global__ __launch_bounds__(kT) void k_hash_route(HashRouteArgs a) {
  constexpr uint32_t kTile = kT * kPer;
  constexpr bool kBuckets = OUT != kFmtRaw;
  static_assert(OUT != kFmtRaw || (KIND == kKindU64 && !SETS), "only u64 keys of plain builds travel raw");
  extern __shared__ __align__(16) uint8_t lds[];
  // sets of several configurations bring a fulcrum table per lambda
  const uint32_t tabs = kBuckets ? (SETS && a.ntabs ? a.ntabs : 1u) : 0u;
  const TileLayout lay = tile_layout(kTile, OUT, a.fanout, tabs);
  const Tile<OUT> s = tile_at<OUT>(lds, lay);
  if constexpr (kBuckets) {
    const uint32_t* tsrc = SETS && a.ntabs ? a.ftabs : a.fulcs;
    for (uint32_t i = threadIdx.x; i < tabs * kFulcs; i += kT) s.fulcs[i] = tsrc[i];
    __syncthreads();
  }
  const uint64_t ntiles = (a.n + kTile - 1) / kTile;
  const float rG = 1.0f / (float)a.G;
  // u64 keys: the next tile's keys are loaded while this tile is routed,
  // through a buffer resource of the tile (keys past the end read as 0)
  uint64_t kv[kPer];
  auto load_u64 = [&](uint64_t t, uint64_t (&dst)[kPer]) {
    const uint64_t i0 = t * kTile;
    const uint32_t cnt = (uint32_t)min((uint64_t)kTile, a.n - i0);
    const uint64_t* src = nullptr;
    if (!(SETS && a.set_keys)) {
      src = a.src.u64 + i0;
    } else {
      // sets read in place: the set of the tile's first key (the last s
      // with set_key_off[s] <= i0, wave-uniform)
      uint32_t lo_s = 0, hi_s = a.sets;
      while (hi_s - lo_s > 1) {
        const uint32_t mid = (lo_s + hi_s) / 2;
        if (a.set_key_off[mid] <= i0) lo_s = mid;
        else hi_s = mid;
      }
      const uint64_t s_beg = a.set_key_off[lo_s], s_end = a.set_key_off[lo_s + 1];
      if (i0 + cnt > s_end) {  // the tile spans sets: each key finds its own
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
          const uint64_t i = i0 + k * kT + threadIdx.x;
          dst[k] = 0;
          if (i < a.n) {
            uint32_t sk = lo_s;
            while (a.set_key_off[sk + 1] <= i) sk++;
            dst[k] = a.set_keys[sk][i - a.set_key_off[sk]];
          }
        }
        return;
      }
      src = a.set_keys[lo_s] + (i0 - s_beg);
    }
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(src, cnt * 8);
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) dst[k] = buf_load_u64(r, threadIdx.x * 8, k * kT * 8);
  };
  if (KIND == kKindU64 && blockIdx.x < ntiles) load_u64(blockIdx.x, kv);
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    uint32_t set = 0;
    uint64_t set_end = ~0ull;
    uint32_t pb = 0, pn = a.np;
    uint64_t seed = a.seed;  // of the current set
    uint32_t ftab = 0;       // its fulcrum table, as an offset into s.fulcs
    if constexpr (SETS) {
      // set of this thread's first key: last s with set_key_off[s] <= i
      const uint64_t i0 = t * kTile + threadIdx.x;
      uint32_t lo_s = 0, hi_s = a.sets;
      while (hi_s - lo_s > 1) {
        const uint32_t mid = (lo_s + hi_s) / 2;
        if (a.set_key_off[mid] <= i0) lo_s = mid;
        else hi_s = mid;
      }
      set = lo_s;
      // first key offset of the next set, and the set's partition range
      if (set + 1 < a.sets) set_end = a.set_key_off[set + 1];
      pb = a.set_part_base[set];
      pn = a.set_part_base[set + 1] - pb;
      if (a.set_seed) {
        seed = a.set_seed[set];
        ftab = a.set_ftab[set] * kFulcs;
      }
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
        if (SETS && i >= set_end) {  // a later set (loads only at a boundary)
          do {
            set++;
            set_end = set + 1 < a.sets ? a.set_key_off[set + 1] : ~0ull;
          } while (i >= set_end);
          pb = a.set_part_base[set];
          pn = a.set_part_base[set + 1] - pb;
          if (a.set_seed) {
            seed = a.set_seed[set];
            ftab = a.set_ftab[set] * kFulcs;
          }
        }
        const U128 h = KIND == kKindU64 ? xxh3_128_u64(kv[k], seed) : hash_key<KIND>(a.src, i, seed);
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
          const uint32_t b = bucket_of((uint32_t)(h.lo >> 32), s.fulcs + ftab);
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

// ---------------------------------------------------------------------------
// Software-pipelined persistent routing (plain u64 keys at level 1, and every
// split level with at most kPipeParents parent bins).
//
// The kernels above route one tile after another in phases separated by
// barriers: load and hash (vector ALU), count / scan / scatter (LDS) and
// write-out (HBM), each with the other units idle.  One tile takes 135-150 KB
// of LDS, so a CU holds one workgroup and nothing overlaps the phases.
// Here each persistent workgroup keeps one routed tile staged in LDS while it
// hashes the next one: in phase A half of the waves of every SIMD first
// write the staged tile out and then decode the next one from registers, the
// other half do the same in the other order, so on every SIMD the HBM-bound
// write-out of some waves runs beside the ALU-bound hashing of the others.
// The tile after that is loaded (into registers, through a buffer resource)
// before the LDS phases of the current one, so its loads are in flight while
// the LDS works.  Per tile:
//   A  write-out of the staged tile  ||  decode of the current tile
//      issue the loads of the next tile                          barrier
//   B  count the current tile's records per child bin (LDS atomics)  barrier
//   C  scan the counts, reserve each bin's range (global atomics)     barrier
//   D  bin starts in the tile; clear the counts for the next tile     barrier
//   E  bin bases and limits (the reserved ranges arrive here); scatter the
//      current tile into LDS in bin order: it is now the staged tile  barrier
// Each child bin receives the same records as with the per-tile kernels, in
// the order its ranges are reserved (nondeterministic in both; no output
// depends on it: the search orders a partition's keys by bucket, and a
// bucket's pilot depends on the set of its keys only).
//
// LDS: lo[tile] u64 | meta[tile] | child[tile] u16 | hist[F] | lim[F] |
//      base[F] | lstart[F] | wsum[16] | tstart[parents + 1] and
//      pcnt[parents] (splits, parents <= kPipeParents) |
//      fulcrums (levels that map keys to buckets)
constexpr uint32_t kPipeParents = 1024;

struct PipeLayout {
  uint32_t lo, meta, child, hist, lim, base, lstart, wsum, tstart, fulcs, bytes;
};
__host__ __device__ inline PipeLayout pipe_layout(uint32_t tile, int fmt, uint32_t F, bool fulcs, bool split) {
  PipeLayout l;
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
  l.hist = take(F * 4, 4);
  l.lim = take(F * 4, 4);
  l.base = take(F * 4, 4);
  l.lstart = take(F * 4, 4);
  l.wsum = take(kMaxWaves * 4, 4);
  l.tstart = take(split ? (2 * kPipeParents + 1) * 4 : 0, 4);  // tstart[parents + 1] | pcnt[parents]
  l.fulcs = take(fulcs ? kFulcs * 4 : 0, 4);
  l.bytes = at;
  return l;
}

template <int FMT>
struct PTile {
  uint64_t* lo;
  typename Meta<FMT>::T* meta;
  uint16_t* child;
  uint32_t* hist;    // records of the current tile per bin
  uint32_t* lim;     // per bin of the staged tile: end of its records that fit the bin
  uint32_t* base;    // per bin of the staged tile: first slot minus first record (mod 2^32)
  uint32_t* lstart;  // per bin of the current tile: its first record in tile order
  uint32_t* wsum;
  uint32_t* tstart;
  uint32_t* fulcs;
};

template <int FMT>
__device__ __forceinline__ PTile<FMT> ptile_at(uint8_t* lds, const PipeLayout& l) {
  PTile<FMT> t;
  t.lo = reinterpret_cast<uint64_t*>(lds + l.lo);
  t.meta = reinterpret_cast<typename Meta<FMT>::T*>(lds + l.meta);
  t.child = reinterpret_cast<uint16_t*>(lds + l.child);
  t.hist = reinterpret_cast<uint32_t*>(lds + l.hist);
  t.lim = reinterpret_cast<uint32_t*>(lds + l.lim);
  t.base = reinterpret_cast<uint32_t*>(lds + l.base);
  t.lstart = reinterpret_cast<uint32_t*>(lds + l.lstart);
  t.wsum = reinterpret_cast<uint32_t*>(lds + l.wsum);
  t.tstart = reinterpret_cast<uint32_t*>(lds + l.tstart);
  t.fulcs = reinterpret_cast<uint32_t*>(lds + l.fulcs);
  return t;
}

// One tile, wave-uniform: its first input record, valid records (a prefix of
// the tile), first child bin and (splits) the parent's first partition.
struct TileRef {
  uint64_t in;
  uint64_t child_first;
  uint32_t total;
  uint32_t first_part;
};

// Phase A's write-out of the staged tile: records leave in chunks of four
// per thread (two when kPer is not a multiple of four) whose LDS reads are
// unconditional (a slot past `total` holds a stale or unwritten child,
// clamped into the bins), so a chunk's reads are in flight together and
// only the stores are predicated.  The thread index is hidden from the
// optimiser: otherwise it hoists the kPer addresses out + 8 * (k * kT + tid)
// out of the tile loop and keeps them live (spilled).
template <uint32_t kT, uint32_t kPer, int FMT>
__device__ __forceinline__ void pipe_write(const PTile<FMT>& s, const TileRef& p, uint32_t F, uint64_t child_cap,
                                           uint64_t* __restrict__ out_lo, void* __restrict__ out_meta, bool& over) {
  using MT = typename Meta<FMT>::T;
  constexpr bool kMeta = FMT != kFmtRaw;
  uint32_t tid_o = threadIdx.x;
  asm volatile("" : "+v"(tid_o));
  constexpr uint32_t kChunk = kPer % 4 == 0 ? 4 : 2;
  static_assert(kPer % kChunk == 0, "records per thread come in whole chunks");
#pragma unroll
  for (uint32_t k0 = 0; k0 < kPer; k0 += kChunk) {
    uint32_t c[kChunk], lim[kChunk];
    uint64_t at[kChunk], v[kChunk];
    MT mv[kChunk];
#pragma unroll
    for (uint32_t j = 0; j < kChunk; j++) c[j] = min((uint32_t)s.child[(k0 + j) * kT + tid_o], F - 1);
#pragma unroll
    for (uint32_t j = 0; j < kChunk; j++) {
      const uint32_t d = (k0 + j) * kT + tid_o;
      lim[j] = min(s.lim[c[j]], p.total);
      at[j] = (p.child_first + c[j]) * child_cap + (uint32_t)(s.base[c[j]] + d);
      v[j] = s.lo[d];
      if constexpr (kMeta) mv[j] = s.meta[d];
    }
#pragma unroll
    for (uint32_t j = 0; j < kChunk; j++) {
      const uint32_t d = (k0 + j) * kT + tid_o;
      if (d < lim[j]) {
        out_lo[at[j]] = v[j];
        if constexpr (kMeta) static_cast<MT*>(out_meta)[at[j]] = mv[j];
      } else if (d < p.total) {
        over = true;
      }
    }
  }
}

// Phases B-E for the current tile (registers lo / meta / child, record k of
// thread t is tile record k * kT + t, valid below cur.total); ends with the
// tile staged in LDS.  Expects hist zeroed and the staged tile written out.
template <uint32_t kT, uint32_t kPer, int FMT>
__device__ __forceinline__ void pipe_route(const PTile<FMT>& s, const TileRef& cur, const uint64_t (&lo)[kPer],
                                           const uint32_t (&meta)[kPer], const uint32_t (&child)[kPer], uint32_t F,
                                           uint64_t child_cap, uint32_t* __restrict__ child_count) {
  using MT = typename Meta<FMT>::T;
  constexpr bool kMeta = FMT != kFmtRaw;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // B: counts
  uint32_t rank[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++)
    if (k * kT + tid < cur.total) rank[k] = atomicAdd(&s.hist[child[k]], 1u);
  __syncthreads();
  // C: scan (two bins per thread) and range reservations; the returned
  // ranges are waited for in E (they stay 32-bit until then: a widening
  // copy inside these branches would wait for them here)
  const uint32_t c0 = 2 * tid, c1 = c0 + 1;
  const uint32_t v0 = c0 < F ? s.hist[c0] : 0, v1 = c1 < F ? s.hist[c1] : 0;
  const uint32_t pair = v0 + v1;
  const uint32_t incl = wave_incl_scan(pair);
  if (lane == 63) s.wsum[wid] = incl;
  uint32_t g0 = 0, g1 = 0;
  if (c0 < F) g0 = atomicAdd(&child_count[cur.child_first + c0], v0);
  if (c1 < F) g1 = atomicAdd(&child_count[cur.child_first + c1], v1);
  __syncthreads();
  // D: bin starts; the counts are in registers now, so they are cleared for
  // the next tile
  uint32_t before = 0;
  for (uint32_t w = 0; w < wid; w++) before += s.wsum[w];
  const uint32_t l0 = before + incl - pair;
  if (c0 < F) {
    s.lstart[c0] = l0;
    s.hist[c0] = 0;
  }
  if (c1 < F) {
    s.lstart[c1] = l0 + v0;
    s.hist[c1] = 0;
  }
  __syncthreads();
  // E: record d of bin c goes to bin slot g_c + (d - lstart_c) while that is
  // below child_cap (< 2^32, checked by the launcher); scatter
  auto bin = [&](uint32_t c, uint32_t g, uint32_t v, uint32_t l) {
    s.base[c] = g - l;
    s.lim[c] = l + (uint32_t)min((uint64_t)v, child_cap > g ? child_cap - g : 0);
  };
  if (c0 < F) bin(c0, g0, v0, l0);
  if (c1 < F) bin(c1, g1, v1, l0 + v0);
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    if (k * kT + tid < cur.total) {
      const uint32_t d = s.lstart[child[k]] + rank[k];
      s.lo[d] = lo[k];
      if constexpr (kMeta) s.meta[d] = (MT)meta[k];
      s.child[d] = (uint16_t)child[k];
    }
  }
  __syncthreads();
}

// The pipeline over this workgroup's tiles.  `Src` provides the tiles
// (first(ref), next(ref) -> false past the last), fetch(ref, in) issuing a
// tile's loads into registers, settle(in) waiting for them, and
// decode(ref, in, k, lo, meta, child).
template <uint32_t kT, uint32_t kPer, int FMT, class Src>
__device__ __forceinline__ void pipe_run(Src& src, const PTile<FMT>& s, uint32_t F, uint64_t child_cap,
                                         uint32_t* __restrict__ child_count, uint64_t* __restrict__ out_lo,
                                         void* __restrict__ out_meta, uint32_t* __restrict__ status) {
  using In = typename Src::In;
  TileRef cur, prev;
  if (!src.first(cur)) return;  // the whole workgroup leaves together
  In in;
  src.fetch(cur, in);
  // the first tile's loads are waited for here, so that the loop starts
  // with no loads in flight on every path into it: otherwise the waits in
  // the decode (for loads issued before the write-out's stores) would also
  // wait for those stores
  src.settle(in);
  for (uint32_t i = threadIdx.x; i < F; i += kT) s.hist[i] = 0;
  bool have_prev = false, over = false;
  // waves go to the four SIMDs in turn: bit 2 of the wave index splits each
  // SIMD's waves into the two orders
  const bool write_first = (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) & 4) != 0;
  for (;;) {
    uint64_t lo[kPer];
    uint32_t meta[kPer], child[kPer];
    auto decode = [&]() {
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) src.decode(cur, in, k, lo[k], meta[k], child[k]);
    };
    // A
    if (write_first) {
      if (have_prev) pipe_write<kT, kPer, FMT>(s, prev, F, child_cap, out_lo, out_meta, over);
      decode();
    } else {
      decode();
      if (have_prev) pipe_write<kT, kPer, FMT>(s, prev, F, child_cap, out_lo, out_meta, over);
    }
    TileRef nxt = cur;
    const bool more = src.next(nxt);
    if (more) src.fetch(nxt, in);
    __syncthreads();
    // B-E
    pipe_route<kT, kPer, FMT>(s, cur, lo, meta, child, F, child_cap, child_count);
    prev = cur;
    have_prev = true;
    if (!more) break;
    cur = nxt;
  }
  pipe_write<kT, kPer, FMT>(s, prev, F, child_cap, out_lo, out_meta, over);
  if (over) atomicOr(status, 1u << kErrPartitionOverflow);
}

// Level 1 of plain u64 builds: this workgroup's tiles are a contiguous run of
// the key array.
template <uint32_t kT, uint32_t kPer, int OUT>
struct KeyTiles {
  static constexpr uint32_t kTile = kT * kPer;
  struct In {
    uint64_t key[kPer];
  };
  const HashRouteArgs& a;
  const uint32_t* fulcs;
  uint64_t t, t_end;  // this workgroup's tiles: t, t + gridDim.x, ... below t_end
  float rG;

  __device__ bool first(TileRef& r) {
    t = blockIdx.x;
    t_end = (a.n + kTile - 1) / kTile;
    rG = 1.0f / (float)a.G;
    return t < t_end && at(r);
  }
  __device__ bool at(TileRef& r) const {
    r.in = t * kTile;
    r.total = (uint32_t)min((uint64_t)kTile, a.n - r.in);
    r.child_first = 0;
    r.first_part = 0;
    return true;
  }
  __device__ bool next(TileRef& r) {
    t += gridDim.x;
    if (t >= t_end) return false;
    return at(r);
  }
  __device__ void fetch(const TileRef& r, In& in) const {
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(a.src.u64 + r.in, r.total * 8);
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) in.key[k] = buf_load_u64(rs, threadIdx.x * 8, k * kT * 8);
  }
  __device__ void settle(In& in) const {
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) asm volatile("" : "+v"(in.key[k]));
  }
  __device__ void decode(const TileRef&, const In& in, uint32_t k, uint64_t& lo, uint32_t& meta,
                         uint32_t& child) const {
    const U128 h = xxh3_128_u64(in.key[k], a.seed);
    const uint32_t p = partition_of((uint32_t)h.lo, a.np);
    // level-1 bin c = p / G and r = p mod G: the f32 quotient estimate is
    // within one of c (c < 2^11, relative error below 2^-22), G < 2^16, so
    // q * G is a 24-bit product and one step corrects it
    uint32_t c = (uint32_t)(__uint2float_rn(p) * rG);
    int32_t rm = (int32_t)(p - __umul24(c, a.G));
    if (rm < 0) {
      c--;
      rm += (int32_t)a.G;
    } else if (rm >= (int32_t)a.G) {
      c++;
      rm -= (int32_t)a.G;
    }
    if constexpr (OUT == kFmtRaw) {
      lo = in.key[k];
      meta = 0;
    } else {
      lo = h.hi;
      const uint32_t b = bucket_of((uint32_t)(h.lo >> 32), fulcs);
      meta = OUT == kFmtRec ? ((uint32_t)rm << 16) | b : b;
    }
    child = min(c, a.fanout - 1);
  }
};

template <uint32_t kT, uint32_t kPer, int OUT>
__global__ __launch_bounds__(kT) void k_route_keys(HashRouteArgs a) {
  constexpr uint32_t kTile = kT * kPer;
  constexpr bool kBuckets = OUT != kFmtRaw;
  extern __shared__ __align__(16) uint8_t lds[];
  const PipeLayout lay = pipe_layout(kTile, OUT, a.fanout, kBuckets, false);
  const PTile<OUT> s = ptile_at<OUT>(lds, lay);
  if constexpr (kBuckets) {
    for (uint32_t i = threadIdx.x; i < kFulcs; i += kT) s.fulcs[i] = a.fulcs[i];
    __syncthreads();
  }
  KeyTiles<kT, kPer, OUT> src{a, s.fulcs};
  pipe_run<kT, kPer, OUT>(src, s, a.fanout, (uint64_t)a.G * a.cap, a.child_count, a.out_lo, a.out_meta, a.status);
}

// Split levels: the tiles of all parents in parent order (parent b has
// ceil(count_b / kTile) of them, tstart[] their exclusive scan, in LDS).
// Each XCD takes a contiguous eighth of them, which its workgroups share in
// turn (workgroup j of the XCD takes tiles j, j + W, ...): neighbouring tiles
// of a parent then run at the same time on CUs of one XCD, so their runs in
// each child bin are adjacent and their partial lines merge in its L2.
template <uint32_t kT, uint32_t kPer, int IN, int OUT>
struct SplitTiles {
  static constexpr uint32_t kTile = kT * kPer;
  struct In {
    uint64_t lo[kPer];
    uint32_t m[kPer];
  };
  const SplitArgs& a;
  const uint32_t* tstart;  // LDS: the parents' first tiles
  const uint32_t* pcnt;    // LDS: the parents' records (clamped to a bin)
  const uint32_t* fulcs;
  uint32_t parents;
  uint32_t t, t_end, step, b;  // tiles t, t + step, ... below t_end; b = parent of tile t
  uint32_t G, inv;

  __device__ bool first(TileRef& r) {
    const uint32_t ntiles = __builtin_amdgcn_readfirstlane(tstart[parents]);
    const uint32_t x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    step = (gridDim.x - x + kXcds - 1) / kXcds;  // workgroups on XCD x
    t = (uint32_t)((uint64_t)ntiles * x / kXcds) + j;
    t_end = (uint32_t)((uint64_t)ntiles * (x + 1) / kXcds);
    if (t >= t_end) return false;
    // b: last parent with tstart[b] <= t
    uint32_t lo_b = 0, hi_b = parents;
    while (hi_b - lo_b > 1) {
      const uint32_t mid = (lo_b + hi_b) / 2;
      if (__builtin_amdgcn_readfirstlane(tstart[mid]) <= t) lo_b = mid;
      else hi_b = mid;
    }
    b = lo_b;
    G = a.G_child;
    inv = G > 1 ? (uint32_t)(0xFFFFFFFFu / G) + 1 : 0;
    return at(r);
  }
  // (everything here is wave-uniform and read through readfirstlane, so the
  // tile's buffer resources are scalar)
  __device__ bool at(TileRef& r) {
    while (__builtin_amdgcn_readfirstlane(tstart[b + 1]) <= t) b++;
    const uint32_t cnt = __builtin_amdgcn_readfirstlane(pcnt[b]);
    const uint32_t tile0 = (t - __builtin_amdgcn_readfirstlane(tstart[b])) * kTile;
    r.in = (uint64_t)b * a.parent_cap + tile0;
    r.total = min(kTile, cnt - tile0);
    r.child_first = (uint64_t)b * a.fanout;
    r.first_part = b * a.G_parent;
    return true;
  }
  __device__ bool next(TileRef& r) {
    t += step;
    if (t >= t_end) return false;
    return at(r);
  }
  __device__ void fetch(const TileRef& r, In& in) const {
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(a.in_lo + r.in, r.total * 8);
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) in.lo[k] = buf_load_u64(rs, threadIdx.x * 8, k * kT * 8);
    if constexpr (IN == kFmtRec) {
      const __amdgpu_buffer_rsrc_t rm =
          buf_rsrc(static_cast<const uint32_t*>(a.in_meta) + r.in, r.total * 4);
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++)
        in.m[k] = __builtin_amdgcn_raw_buffer_load_b32(rm, (int)(threadIdx.x * 4), (int)(k * kT * 4), 0);
    }
  }
  __device__ void settle(In& in) const {
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      asm volatile("" : "+v"(in.lo[k]));
      if constexpr (IN == kFmtRec) asm volatile("" : "+v"(in.m[k]));
    }
  }
  __device__ void decode(const TileRef& r, const In& in, uint32_t k, uint64_t& lo, uint32_t& meta,
                         uint32_t& child) const {
    uint32_t pl;
    lo = in.lo[k];
    meta = 0;
    if constexpr (IN == kFmtRaw) {
      const U128 h = xxh3_128_u64(in.lo[k], a.seed);
      pl = partition_of((uint32_t)h.lo, a.np) - r.first_part;
      if constexpr (OUT == kFmtFinal) {
        meta = bucket_of((uint32_t)(h.lo >> 32), fulcs);
        lo = h.hi;
      }
    } else {
      pl = in.m[k] >> 16;
    }
    // child bin pl / G_child = umulhi(pl, floor(2^32 / G_child) + 1) for
    // pl < 2^16 and 2 <= G_child < 2^16 (the error stays below 1 / G_child)
    const uint32_t c = G > 1 ? __umulhi(pl, inv) : pl;
    if constexpr (IN == kFmtRec)
      meta = OUT == kFmtRec ? ((pl - c * G) << 16) | (in.m[k] & 0xFFFFu) : in.m[k] & 0xFFFFu;
    child = min(c, a.fanout - 1);
  }
};

template <uint32_t kT, uint32_t kPer, int IN, int OUT>
__global__ __launch_bounds__(kT) void k_split_pipe(SplitArgs a, uint32_t parents) {
  constexpr uint32_t kTile = kT * kPer;
  constexpr bool kBuckets = IN == kFmtRaw && OUT == kFmtFinal;
  static_assert(IN != kFmtRaw || OUT != kFmtRec, "raw keys go to raw or final records");
  static_assert(IN != kFmtRec || OUT != kFmtRaw, "hashed records cannot become keys again");
  extern __shared__ __align__(16) uint8_t lds[];
  const PipeLayout lay = pipe_layout(kTile, OUT, a.fanout, kBuckets, true);
  const PTile<OUT> s = ptile_at<OUT>(lds, lay);
  if constexpr (kBuckets)
    for (uint32_t i = threadIdx.x; i < kFulcs; i += kT) s.fulcs[i] = a.fulcs[i];
  // pcnt: the parents' record counts, clamped to a bin (records past it
  // were never written); tstart: exclusive scan of their tile counts
  // (parents <= 1024, two per thread)
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  uint32_t* pcnt = s.tstart + kPipeParents + 1;
  const uint32_t pcap = (uint32_t)min((uint64_t)0xFFFFFFFFu, a.parent_cap);
  auto tiles_of = [&](uint32_t b) {
    if (b >= parents) return 0u;
    const uint32_t c = min(a.parent_count[b], pcap);
    pcnt[b] = c;
    return (c + kTile - 1) / kTile;
  };
  const uint32_t b0 = 2 * tid;
  const uint32_t n0 = tiles_of(b0), n1 = tiles_of(b0 + 1);
  const uint32_t pair = n0 + n1;
  const uint32_t incl = wave_incl_scan(pair);
  if (lane == 63) s.wsum[wid] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wid; w++) before += s.wsum[w];
  const uint32_t e0 = before + incl - pair;
  if (b0 <= parents) s.tstart[b0] = e0;
  if (b0 + 1 <= parents) s.tstart[b0 + 1] = e0 + n0;
  __syncthreads();
  SplitTiles<kT, kPer, IN, OUT> src{a, s.tstart, pcnt, s.fulcs, parents};
  pipe_run<kT, kPer, OUT>(src, s, a.fanout, (uint64_t)a.G_child * a.cap, a.child_count, a.out_lo, a.out_meta,
                          a.status);
}

namespace {

// Fulcrum tables a level-1 tile holds: none for raw keys, one, or one per
// lambda of a batch of several configurations.
uint32_t fulc_tabs(int out_fmt, bool sets, const HashRouteArgs& a) {
  return out_fmt == kFmtRaw ? 0u : sets && a.ntabs ? a.ntabs : 1u;
}

// tabs: fulcrum tables the tile holds (0 when the level maps no buckets)
size_t route_lds(uint32_t tile, int out_fmt, uint32_t F, uint32_t tabs) {
  return tile_layout(tile, out_fmt, F, tabs).bytes;
}

// Tile shapes (threads, records per thread) a routing launch can take.
bool valid_shape(uint32_t threads, uint32_t per) {
  return threads == 512 ? (per == 8 || per == 12 || per == 16 || per == 20 || per == 24)
                        : threads == 1024 && (per == 4 || per == 6 || per == 8 || per == 10 || per == 12);
}

// The largest records per thread, at most `per`, whose tile of `threads`
// threads fits the workgroup's LDS.
uint32_t fitting_per(uint32_t threads, uint32_t per, int out_fmt, uint32_t F, uint32_t tabs) {
  for (uint32_t p : {24u, 20u, 16u, 12u, 10u, 8u, 6u, 4u})
    if (p <= per && valid_shape(threads, p) && route_lds(threads * p, out_fmt, F, tabs) <= kLdsLimit) return p;
  return threads == 512 ? 8 : 4;
}

template <int KIND, uint32_t kT, uint32_t kPer, int OUT, bool SETS>
hipError_t hash_route_go(const HashRouteArgs& a, hipStream_t s) {
  constexpr uint32_t kTile = kT * kPer;
  const uint64_t ntiles = (a.n + kTile - 1) / kTile;
  const uint32_t grid = (uint32_t)(ntiles < 4096 ? ntiles : 4096);
  const size_t lds = route_lds(kTile, OUT, a.fanout, fulc_tabs(OUT, SETS, a));
  hipLaunchKernelGGL((k_hash_route<KIND, kT, kPer, OUT, SETS>), dim3(grid), dim3(kT), lds, s, a);
  return hipGetLastError();
}

template <int KIND, int OUT, bool SETS>
hipError_t hash_route_shape(uint32_t threads, uint32_t per, const HashRouteArgs& a, hipStream_t s) {
  if (threads == 1024) {
    switch (per) {
      case 12: return hash_route_go<KIND, 1024, 12, OUT, SETS>(a, s);
      case 10: return hash_route_go<KIND, 1024, 10, OUT, SETS>(a, s);
      case 8: return hash_route_go<KIND, 1024, 8, OUT, SETS>(a, s);
      case 6: return hash_route_go<KIND, 1024, 6, OUT, SETS>(a, s);
      default: return hash_route_go<KIND, 1024, 4, OUT, SETS>(a, s);
    }
  }
  switch (per) {
    case 24: return hash_route_go<KIND, 512, 24, OUT, SETS>(a, s);
    case 20: return hash_route_go<KIND, 512, 20, OUT, SETS>(a, s);
    case 16: return hash_route_go<KIND, 512, 16, OUT, SETS>(a, s);
    case 12: return hash_route_go<KIND, 512, 12, OUT, SETS>(a, s);
    default: return hash_route_go<KIND, 512, 8, OUT, SETS>(a, s);
  }
}

// Pipelined kernels: the largest records per thread, at most `per`, whose
// tile fits the workgroup's LDS with the pipeline's layout.
uint32_t pipe_per(uint32_t threads, uint32_t per, int out_fmt, uint32_t F, bool fulcs, bool split) {
  for (uint32_t p : {24u, 20u, 16u, 12u, 10u, 8u, 6u, 4u})
    if (p <= per && valid_shape(threads, p) && pipe_layout(threads * p, out_fmt, F, fulcs, split).bytes <= kLdsLimit)
      return p;
  return threads == 512 ? 8 : 4;
}

// A persistent launch: as many workgroups as the device holds at once (at
// most `work`, at least one).  The device's answer for a (device, kernel,
// threads, LDS) is kept per host thread, so repeated builds ask once.
hipError_t persistent_grid(const void* fn, uint32_t threads, size_t lds, uint64_t work, uint32_t* grid) {
  struct Entry {
    const void* fn;
    int dev;
    uint32_t threads;
    size_t lds;
    uint64_t groups;
  };
  static thread_local Entry seen[16];
  static thread_local unsigned next_slot = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  uint64_t g = 0;
  for (const Entry& x : seen)
    if (x.fn == fn && x.dev == dev && x.threads == threads && x.lds == lds) g = x.groups;
  if (!g) {
    int cus = 0, per_cu = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, lds);
    if (e != hipSuccess) return e;
    g = (uint64_t)std::max(cus, 1) * (uint64_t)std::max(per_cu, 1);
    seen[next_slot++ % 16] = Entry{fn, dev, threads, lds, g};
  }
  *grid = (uint32_t)std::max<uint64_t>(1, std::min(g, work));
  return hipSuccess;
}

template <uint32_t kT, uint32_t kPer, int OUT>
hipError_t route_keys_go(const HashRouteArgs& a, hipStream_t s) {
  constexpr uint32_t kTile = kT * kPer;
  const size_t lds = pipe_layout(kTile, OUT, a.fanout, OUT != kFmtRaw, false).bytes;
  uint32_t grid = 0;
  const hipError_t e = persistent_grid(reinterpret_cast<const void*>(&k_route_keys<kT, kPer, OUT>), kT, lds,
                                       (a.n + kTile - 1) / kTile, &grid);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_route_keys<kT, kPer, OUT>), dim3(grid), dim3(kT), lds, s, a);
  return hipGetLastError();
}

template <int OUT>
hipError_t route_keys_shape(uint32_t threads, uint32_t per, const HashRouteArgs& a, hipStream_t s) {
  if (threads == 1024) {
    switch (per) {
      case 12: return route_keys_go<1024, 12, OUT>(a, s);
      case 10: return route_keys_go<1024, 10, OUT>(a, s);
      case 8: return route_keys_go<1024, 8, OUT>(a, s);
      case 6: return route_keys_go<1024, 6, OUT>(a, s);
      default: return route_keys_go<1024, 4, OUT>(a, s);
    }
  }
  switch (per) {
    case 24: return route_keys_go<512, 24, OUT>(a, s);
    case 20: return route_keys_go<512, 20, OUT>(a, s);
    case 16: return route_keys_go<512, 16, OUT>(a, s);
    case 12: return route_keys_go<512, 12, OUT>(a, s);
    default: return route_keys_go<512, 8, OUT>(a, s);
  }
}

template <uint32_t kT, uint32_t kPer, int IN, int OUT>
hipError_t split_pipe_go(const SplitArgs& a, uint32_t parents, hipStream_t s) {
  const size_t lds = pipe_layout(kT * kPer, OUT, a.fanout, IN == kFmtRaw && OUT == kFmtFinal, true).bytes;
  uint32_t grid = 0;
  // the tiles are known only on the device (the parents' counts): every
  // workgroup the device holds is launched, those without tiles leave
  const hipError_t e = persistent_grid(reinterpret_cast<const void*>(&k_split_pipe<kT, kPer, IN, OUT>), kT, lds,
                                       ~0ull, &grid);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_split_pipe<kT, kPer, IN, OUT>), dim3(grid), dim3(kT), lds, s, a, parents);
  return hipGetLastError();
}

template <int IN, int OUT>
hipError_t split_pipe_shape(uint32_t threads, uint32_t per, const SplitArgs& a, uint32_t parents, hipStream_t s) {
  if (threads == 1024) {
    switch (per) {
      case 12: return split_pipe_go<1024, 12, IN, OUT>(a, parents, s);
      case 10: return split_pipe_go<1024, 10, IN, OUT>(a, parents, s);
      case 8: return split_pipe_go<1024, 8, IN, OUT>(a, parents, s);
      case 6: return split_pipe_go<1024, 6, IN, OUT>(a, parents, s);
      default: return split_pipe_go<1024, 4, IN, OUT>(a, parents, s);
    }
  }
  switch (per) {
    case 24: return split_pipe_go<512, 24, IN, OUT>(a, parents, s);
    case 20: return split_pipe_go<512, 20, IN, OUT>(a, parents, s);
    case 16: return split_pipe_go<512, 16, IN, OUT>(a, parents, s);
    case 12: return split_pipe_go<512, 12, IN, OUT>(a, parents, s);
    default: return split_pipe_go<512, 8, IN, OUT>(a, parents, s);
  }
}

template <uint32_t kT, uint32_t kPer, int IN, int OUT>
hipError_t split_go(const SplitArgs& a, uint32_t parents, hipStream_t s) {
  const size_t lds = route_lds(kT * kPer, OUT, a.fanout, IN == kFmtRaw && OUT == kFmtFinal ? 1u : 0u);
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
  return fitting_per(threads, per, out_fmt, fanout, in_fmt == kFmtRaw && out_fmt == kFmtFinal ? 1u : 0u);
}

hipError_t launch_hash_route(int kind, int out_fmt, const HashRouteArgs& a, uint32_t threads, uint32_t per,
                             hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.fanout == 0 || a.fanout > kMaxFanout || (out_fmt == kFmtRaw && (kind != kKindU64 || a.sets)) ||
      !(threads == 512 || threads == 1024) || (uint64_t)a.G * a.cap > 0xFFFFFFFFull)
    return hipErrorInvalidValue;
  // A string tile waits on two dependent loads per key (offsets, then
  // bytes), which sixteen waves of a 1024-thread workgroup hide better than
  // eight: tiles of 1024 x 8 take 2e8 strings through level 1 in 4.5-4.6 ms
  // against 5.9-6.0 ms for 512 x 8, 4.8-5.0 ms for 1024 x 4 or 512 x 4 and
  // 5.9 ms for 512 x 12 (8192 records of 14 bytes, the bins and the
  // fulcrums fit the LDS at every fan-out).
  if (kind == kKindBytes) {
    threads = 1024;
    per = 8;
  }
  // batched builds (sets) take u64 keys or byte strings; sets of several
  // configurations bring a seed and a fulcrum table index per set
  if (a.sets && kind == kKindKey) return hipErrorInvalidValue;
  if (a.set_keys && (!a.sets || kind != kKindU64)) return hipErrorInvalidValue;
  if (a.ntabs > kMaxFulcTabs || (a.set_seed != nullptr) != (a.ntabs != 0) ||
      (a.ntabs && (!a.sets || !a.set_ftab || !a.ftabs)))
    return hipErrorInvalidValue;
  if (kind == kKindU64 && !a.sets) {  // the pipelined kernel: raw keys, or final records when L = 1
    const uint32_t pp = pipe_per(threads, per, out_fmt, a.fanout, out_fmt != kFmtRaw, false);
    if (out_fmt == kFmtRaw) return route_keys_shape<kFmtRaw>(threads, pp, a, s);
    if (out_fmt == kFmtFinal) return route_keys_shape<kFmtFinal>(threads, pp, a, s);
    return hipErrorInvalidValue;
  }
  per = fitting_per(threads, per, out_fmt, a.fanout, fulc_tabs(out_fmt, a.sets != 0, a));
  switch (kind) {
    case kKindU64:  // batched (plain builds took the pipelined kernel)
      if (out_fmt == kFmtRec) return hash_route_shape<kKindU64, kFmtRec, true>(threads, per, a, s);
      return hash_route_shape<kKindU64, kFmtFinal, true>(threads, per, a, s);
    case kKindBytes:
      // 8192 Rec records, 1024 bins and four fulcrum tables take 156 KiB
      if (route_lds(1024 * 8, out_fmt, a.fanout, fulc_tabs(out_fmt, a.sets != 0, a)) > kLdsLimit)
        return hipErrorInvalidValue;
      if (a.sets) {
        if (out_fmt == kFmtRec) return hash_route_go<kKindBytes, 1024, 8, kFmtRec, true>(a, s);
        return hash_route_go<kKindBytes, 1024, 8, kFmtFinal, true>(a, s);
      }
      if (out_fmt == kFmtRec) return hash_route_go<kKindBytes, 1024, 8, kFmtRec, false>(a, s);
      return hash_route_go<kKindBytes, 1024, 8, kFmtFinal, false>(a, s);
    default:
      if (out_fmt == kFmtRec) return hash_route_shape<kKindKey, kFmtRec, false>(threads, per, a, s);
      return hash_route_shape<kKindKey, kFmtFinal, false>(threads, per, a, s);
  }
}

hipError_t launch_split_level(int in_fmt, int out_fmt, const SplitArgs& a, uint32_t parents, uint32_t threads,
                              uint32_t per, hipStream_t s) {
  if (parents == 0) return hipSuccess;
  if (a.fanout == 0 || a.fanout > kMaxFanout || !(threads == 512 || threads == 1024) ||
      (uint64_t)a.G_child * a.cap > 0xFFFFFFFFull)
    return hipErrorInvalidValue;
  // `per` must be what split_tile_per chose: the slices were sized for it
  if (per != split_tile_per(threads, per, in_fmt, out_fmt, a.fanout)) return hipErrorInvalidValue;
  if (parents <= kPipeParents) {  // the pipelined kernel (slices unused)
    const uint32_t pp =
        pipe_per(threads, per, out_fmt, a.fanout, in_fmt == kFmtRaw && out_fmt == kFmtFinal, true);
    if (in_fmt == kFmtRaw && out_fmt == kFmtRaw) return split_pipe_shape<kFmtRaw, kFmtRaw>(threads, pp, a, parents, s);
    if (in_fmt == kFmtRaw && out_fmt == kFmtFinal)
      return split_pipe_shape<kFmtRaw, kFmtFinal>(threads, pp, a, parents, s);
    if (in_fmt == kFmtRec && out_fmt == kFmtRec) return split_pipe_shape<kFmtRec, kFmtRec>(threads, pp, a, parents, s);
    if (in_fmt == kFmtRec && out_fmt == kFmtFinal)
      return split_pipe_shape<kFmtRec, kFmtFinal>(threads, pp, a, parents, s);
    return hipErrorInvalidValue;
  }
  if (in_fmt == kFmtRaw && out_fmt == kFmtRaw) return split_shape<kFmtRaw, kFmtRaw>(threads, per, a, parents, s);
  if (in_fmt == kFmtRaw && out_fmt == kFmtFinal) return split_shape<kFmtRaw, kFmtFinal>(threads, per, a, parents, s);
  if (in_fmt == kFmtRec && out_fmt == kFmtRec) return split_shape<kFmtRec, kFmtRec>(threads, per, a, parents, s);
  if (in_fmt == kFmtRec && out_fmt == kFmtFinal) return split_shape<kFmtRec, kFmtFinal>(threads, per, a, parents, s);
  return hipErrorInvalidValue;
}

}  // namespace phobic
