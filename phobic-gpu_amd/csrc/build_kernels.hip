// build_kernels.hip -- gfx950 kernels of the PHOBIC construction path.
//
// Stage split (this library's own, not the reference's six dispatches):
//   k_hash_scatter    streaming: hash every key on the device (XXH3-128),
//                     map it to (partition, bucket) and append its 64 lower
//                     hash bits + bucket id to the partition's fixed-capacity
//                     region.  Replaces initial_hash (mphf_builder.h:201-216),
//                     bucket_sizes.comp and redistribute_keys.comp's global
//                     scatter.
//   k_partition_offsets  exclusive scan of partition key counts
//                     (prefix_sum.comp + prefix_sum_offset.comp).
//   k_partition_search   one wave64 per partition: bucket counts, bucket
//                     sort (size desc, index desc), in-LDS redistribution and
//                     the pilot search (sort_sum.comp, apply_partition_offset
//                     .comp, redistribute_keys.comp, search.comp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_hash.hpp"
#include "kernels.hpp"

namespace phobic {

// ---------------------------------------------------------------- utils --

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = v;
}

__device__ __forceinline__ void flag_error(uint32_t* status, uint32_t code) {
  atomicOr(status, 1u << code);
}

// --------------------------------------------------- hash + scatter -------

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

template <int KIND>
__global__ __launch_bounds__(256) void k_hash_scatter(ScatterArgs a) {
  __shared__ uint32_t fl[kFulcs];
  for (uint32_t i = threadIdx.x; i < kFulcs; i += blockDim.x) fl[i] = a.fulcs[i];
  __syncthreads();
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < a.n; i += stride) {
    const U128 h = hash_key<KIND>(a.src, i, a.seed);
    const uint32_t p = partition_of((uint32_t)h.lo, a.np);
    const uint32_t b = bucket_of((uint32_t)(h.lo >> 32), fl);
    const uint32_t r = atomicAdd(&a.pcount[p], 1u);
    if (r < a.cap) {
      const uint64_t at = (uint64_t)p * a.cap + r;
      a.rec_lower[at] = h.hi;
      a.rec_bkt[at] = (uint16_t)b;
    } else {
      flag_error(a.status, kErrPartitionOverflow);
    }
  }
}

// ---------------------------------------------- partition offsets scan ----
// One block of 1024 threads, 16 counts per thread per chunk, carried across
// chunks.  offsets[0] = 0, offsets[p+1] = sum of counts[0..p].

__global__ __launch_bounds__(1024) void k_partition_offsets(const uint32_t* __restrict__ counts, uint32_t np,
                                                            uint32_t* __restrict__ offsets) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry_s;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) {
    offsets[0] = 0;
    carry_s = 0;
  }
  __syncthreads();
  constexpr uint32_t kPer = 16, kChunk = 1024 * kPer;
  for (uint64_t c0 = 0; c0 < np; c0 += kChunk) {
    uint32_t v[kPer];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      const uint64_t idx = c0 + (uint64_t)tid * kPer + j;
      v[j] = idx < np ? counts[idx] : 0;
      local += v[j];
    }
    // wave inclusive scan of per-thread totals
    uint32_t incl = local;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t wbase = 0, chunk_total = 0;
    for (uint32_t w = 0; w < 16; w++) {
      const uint32_t ws = wsum[w];
      if (w < wid) wbase += ws;
      chunk_total += ws;
    }
    const uint32_t carry = carry_s;
    uint32_t run = carry + wbase + incl - local;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      const uint64_t idx = c0 + (uint64_t)tid * kPer + j;
      run += v[j];
      if (idx < np) offsets[idx + 1] = run;
    }
    __syncthreads();
    if (tid == 0) carry_s = carry + chunk_total;
    __syncthreads();
  }
}

// ------------------------------------------------- per-partition search ---

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// readlane of register `r[k >> 6]` from lane k & 63, k wave-uniform.
__device__ __forceinline__ uint32_t pick4(const uint32_t (&r)[4], uint32_t k) {
  const uint32_t c = k >> 6;
  const uint32_t src = c == 0 ? r[0] : c == 1 ? r[1] : c == 2 ? r[2] : r[3];
  return __builtin_amdgcn_readlane(src, k & 63);
}

__global__ __launch_bounds__(64) void k_partition_search(SearchArgs a) {
  extern __shared__ __align__(16) uint8_t smem[];
  const uint32_t lane = threadIdx.x;
  const uint32_t cap = a.cap, B = a.B, tw = (cap + 31) / 32;
  uint64_t* skeys = (uint64_t*)smem;          // cap
  uint32_t* taken = (uint32_t*)(skeys + cap); // tw
  uint32_t* coll = taken + tw;                // tw
  uint32_t* bcnt = coll + tw;                 // B   key count per bucket
  uint32_t* bcur = bcnt + B;                  // B   placement cursor
  uint32_t* bstart = bcur + B;                // B   start inside the partition
  uint32_t* order = bstart + B;               // B   bucket index by processing rank
  uint32_t* pil = order + B;                  // B   pilot per bucket index
  uint32_t* hist = pil + B;                   // 256 bin t = 256 - size

  const uint32_t p = blockIdx.x;
  const uint32_t ps = a.pcount[p];
  const uint64_t base = (uint64_t)p * cap;
  if (ps > cap) {  // k_hash_scatter already raised kErrPartitionOverflow
    for (uint32_t b = lane; b < B; b += 64) a.pilots[p + (uint64_t)b * a.np] = 0;
    return;
  }
  for (uint32_t i = lane; i < tw; i += 64) {
    taken[i] = 0;
    coll[i] = 0;
  }
  for (uint32_t i = lane; i < B; i += 64) {
    bcnt[i] = 0;
    bcur[i] = 0;
    pil[i] = 0;
  }
  for (uint32_t i = lane; i < 256; i += 64) hist[i] = 0;
  __syncthreads();

  // bucket sizes (bucket_sizes.comp:21-35, here with LDS atomics)
  for (uint32_t i = lane; i < ps; i += 64) atomicAdd(&bcnt[a.rec_bkt[base + i]], 1u);
  __syncthreads();
  bool over = false;
  for (uint32_t b = lane; b < B; b += 64) {
    const uint32_t s = bcnt[b];
    if (s > 256) over = true;
    else if (s) atomicAdd(&hist[256 - s], 1u);
  }
  if (__any(over)) {
    if (lane == 0) flag_error(a.status, kErrBucketOverflow);
    for (uint32_t b = lane; b < B; b += 64) a.pilots[p + (uint64_t)b * a.np] = 0;
    return;
  }
  __syncthreads();

  // bin starts: exclusive scan over t = 256 - size (largest size first),
  // as the per-partition histogram scan of sort_sum.comp:60-139,184-199.
  uint32_t m;
  {
    uint32_t v[4], loc = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      v[j] = hist[4 * lane + j];
      loc += v[j];
    }
    const uint32_t incl = wave_incl_scan(loc);
    uint32_t run = incl - loc;
    m = __shfl(incl, 63, 64);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      hist[4 * lane + j] = run;
      run += v[j];
    }
  }
  __syncthreads();

  // ranks: within one size, larger bucket index first (the serial loop of
  // sort_sum.comp:156-172 runs index B-1 .. 0).
  for (uint32_t c = 0; c < B; c += 64) {
    const int idx = (int)B - 1 - (int)(c + lane);
    const uint32_t s = idx >= 0 ? bcnt[idx] : 0;
    bool act = s > 0;
    uint64_t am = __ballot(act);
    while (am) {
      const int leader = __ffsll((unsigned long long)am) - 1;
      const uint32_t v = __shfl(s, leader, 64);
      const bool mine = act && s == v;
      const uint64_t same = __ballot(mine);
      const uint32_t t = 256 - v;
      const uint32_t bv = hist[t];
      if (mine) {
        order[bv + __popcll(same & lanemask_lt())] = (uint32_t)idx;
        act = false;
      }
      __syncthreads();
      if (lane == (uint32_t)leader) hist[t] = bv + __popcll(same);
      __syncthreads();
      am &= ~same;
    }
  }
  __syncthreads();

  // bucket start offsets inside the partition (sort_sum.comp:203-225)
  {
    uint32_t carry = 0;
    for (uint32_t r0 = 0; r0 < m; r0 += 64) {
      const uint32_t r = r0 + lane;
      const uint32_t b = r < m ? order[r] : 0;
      const uint32_t s = r < m ? bcnt[b] : 0;
      const uint32_t incl = wave_incl_scan(s);
      if (r < m) bstart[b] = carry + incl - s;
      carry += __shfl(incl, 63, 64);
    }
  }
  __syncthreads();

  // redistribution into LDS (redistribute_keys.comp:21-53; order inside a
  // bucket is irrelevant to the pilots)
  for (uint32_t i = lane; i < ps; i += 64) {
    const uint32_t b = a.rec_bkt[base + i];
    const uint32_t pos = bstart[b] + atomicAdd(&bcur[b], 1u);
    skeys[pos] = a.rec_lower[base + i];
  }
  __syncthreads();

  // pilot search (search.comp:105-175): for each bucket in order the
  // smallest V = d + ps*pilot such that (init_pos + d) mod ps are pairwise
  // distinct and free.
  uint32_t err = 0;
  for (uint32_t r = 0; r < m && !err; r++) {
    const uint32_t b = order[r];
    const uint32_t s = bcnt[b];
    const uint32_t st = bstart[b];
    uint64_t key[4];
    uint32_t pos[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t k = lane + 64 * j;
      key[j] = k < s ? skeys[st + k] : 0;
      pos[j] = 0;
    }
    uint32_t pilot = 0, dsel = 0;
    for (;;) {
      if ((uint64_t)ps * pilot > 0x7FFFFFFFull) {
        err = kErrPilotOverflow;
        break;
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t k = lane + 64 * j;
        if (k < s) pos[j] = (mix32((uint32_t)key[j], mix32((uint32_t)(key[j] >> 32), pilot)) >> 1) % ps;
      }
      if (s > 1) {
        bool c = false;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t k = lane + 64 * j;
          if (k < s) {
            const uint32_t msk = 1u << (pos[j] & 31);
            const uint32_t old = atomicOr(&coll[pos[j] >> 5], msk);
            c |= (old & msk) != 0;
          }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t k = lane + 64 * j;
          if (k < s) coll[pos[j] >> 5] = 0;
        }
        __syncthreads();
        if (__any(c)) {
          if (pilot == 0) {
            bool dup = false;
            for (uint32_t k2 = 0; k2 < s; k2++) {
              const uint64_t kv = skeys[st + k2];
#pragma unroll
              for (int j = 0; j < 4; j++) {
                const uint32_t k = lane + 64 * j;
                if (k < s && k != k2 && key[j] == kv) dup = true;
              }
            }
            if (__any(dup)) {
              err = kErrDuplicateKeys;
              break;
            }
          }
          pilot++;
          continue;
        }
      }
      bool found = false;
      for (uint32_t D = 0; D < ps; D += 64) {
        const uint32_t d = D + lane;
        bool ok = d < ps && ((uint64_t)d + (uint64_t)ps * pilot) <= 0x7FFFFFFFull;
        const uint32_t dd = d < ps ? d : 0;
        for (uint32_t k = 0; k < s; k++) {
          const uint32_t pk = pick4(pos, k);
          uint32_t q = pk + dd;
          if (q >= ps) q -= ps;
          ok = ok && !((taken[q >> 5] >> (q & 31)) & 1u);
          if ((k & 7) == 7 && !__any(ok)) break;
        }
        const uint64_t mk = __ballot(ok);
        if (mk) {
          dsel = D + (uint32_t)(__ffsll((unsigned long long)mk) - 1);
          found = true;
          break;
        }
      }
      if (found) break;
      pilot++;
    }
    if (err) break;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t k = lane + 64 * j;
      if (k < s) {
        uint32_t q = pos[j] + dsel;
        if (q >= ps) q -= ps;
        atomicOr(&taken[q >> 5], 1u << (q & 31));
      }
    }
    if (lane == 0) pil[b] = dsel + ps * pilot;
    __syncthreads();
  }
  if (err && lane == 0) flag_error(a.status, err);
  __syncthreads();
  for (uint32_t b = lane; b < B; b += 64) a.pilots[p + (uint64_t)b * a.np] = pil[b];
}

// ------------------------------------------------------------- evaluate ---
// MPHF::operator() (mphf.hpp:32-40,77-86) over a device batch of u64 keys.

__global__ __launch_bounds__(256) void k_eval_u64(EvalArgs a) {
  __shared__ uint32_t fl[kFulcs];
  for (uint32_t i = threadIdx.x; i < kFulcs; i += blockDim.x) fl[i] = a.fulcs[i];
  __syncthreads();
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < a.count; i += stride) {
    const U128 h = xxh3_128_u64(a.keys[i], a.seed);
    const uint32_t p = partition_of((uint32_t)h.lo, a.np);
    const uint32_t b = bucket_of((uint32_t)(h.lo >> 32), fl);
    const uint32_t V = a.pilots[p + (uint64_t)b * a.np];
    const uint32_t off = a.offsets[p];
    const uint32_t ps = a.offsets[p + 1] - off;
    const uint32_t hp = V / ps;
    const uint32_t hv = mix32((uint32_t)h.hi, mix32((uint32_t)(h.hi >> 32), hp)) >> 1;
    a.out[i] = off + (hv + V) % ps;
  }
}

// Bijection check: marks positions in a bitmap of n bits; counts positions
// out of range or seen twice.
__global__ __launch_bounds__(256) void k_check_bijection(const uint32_t* __restrict__ pos, uint64_t count,
                                                         uint64_t n, uint32_t* bitmap, unsigned long long* bad) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t local = 0;
  for (; i < count; i += stride) {
    const uint32_t q = pos[i];
    if (q >= n) {
      local++;
      continue;
    }
    const uint32_t m = 1u << (q & 31);
    if (atomicOr(&bitmap[q >> 5], m) & m) local++;
  }
  if (local) atomicAdd(bad, (unsigned long long)local);
}

// splitmix64 (Steele, Lea, Flood 2014): output i uses state seed + (i+1)*gamma.
__global__ __launch_bounds__(256) void k_splitmix64(uint64_t seed, uint64_t n, uint64_t* out) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    out[i] = z ^ (z >> 31);
  }
}

// ------------------------------------------------------------ launchers ---

static inline uint32_t stream_grid(uint64_t n, uint32_t block, uint32_t max_blocks) {
  uint64_t g = (n + block - 1) / block;
  if (g > max_blocks) g = max_blocks;
  return g ? (uint32_t)g : 1u;
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(stream_grid(n, 256, 4096)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_hash_scatter(int kind, const ScatterArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const dim3 grid(stream_grid(a.n, 256, 256 * 32)), block(256);
  switch (kind) {
    case kKindU64: hipLaunchKernelGGL(k_hash_scatter<kKindU64>, grid, block, 0, s, a); break;
    case kKindBytes: hipLaunchKernelGGL(k_hash_scatter<kKindBytes>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_hash_scatter<kKindKey>, grid, block, 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_partition_offsets(const uint32_t* counts, uint32_t np, uint32_t* offsets, hipStream_t s) {
  hipLaunchKernelGGL(k_partition_offsets, dim3(1), dim3(1024), 0, s, counts, np, offsets);
  return hipGetLastError();
}

size_t search_lds_bytes(uint32_t cap, uint32_t B) {
  const size_t tw = (cap + 31) / 32;
  return (size_t)cap * 8 + 2 * tw * 4 + 5 * (size_t)B * 4 + 256 * 4;
}

hipError_t launch_partition_search(const SearchArgs& a, hipStream_t s) {
  if (a.np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_partition_search, dim3(a.np), dim3(64), search_lds_bytes(a.cap, a.B), s, a);
  return hipGetLastError();
}

hipError_t launch_eval_u64(const EvalArgs& a, hipStream_t s) {
  if (a.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_eval_u64, dim3(stream_grid(a.count, 256, 256 * 32)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_check_bijection(const uint32_t* pos, uint64_t count, uint64_t n, uint32_t* bitmap,
                                  unsigned long long* bad, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_bijection, dim3(stream_grid(count, 256, 256 * 32)), dim3(256), 0, s, pos, count, n,
                     bitmap, bad);
  return hipGetLastError();
}

hipError_t launch_splitmix64(uint64_t seed, uint64_t n, uint64_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_splitmix64, dim3(stream_grid(n, 256, 256 * 32)), dim3(256), 0, s, seed, n, out);
  return hipGetLastError();
}

}  // namespace phobic
