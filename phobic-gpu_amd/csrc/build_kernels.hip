// build_kernels.hip -- gfx950 kernels of the PHOBIC construction path.
//
// Stage split (this library's own, not the reference's six dispatches):
//   k_hash_route / k_split_level  (partition_kernels.hip) streaming: hash
//                     every key on the device (XXH3-128), map it to
//                     (partition, bucket) and route its 64 lower hash bits +
//                     bucket id to the partition's fixed-capacity region in a
//                     few LDS-staged radix levels.  Replaces initial_hash
//                     (mphf_builder.h:201-216), bucket_sizes.comp and
//                     redistribute_keys.comp's global scatter.
//   k_scan_totals/write  exclusive scans of partition key counts and remap
//                     sizes (prefix_sum.comp + prefix_sum_offset.comp).
//   k_partition_search   (search_kernel.hip) one wave64 per partition:
//                     bucket counts, bucket sort (size desc, index desc),
//                     redistribution and the pilot search (sort_sum.comp,
//                     apply_partition_offset.comp, redistribute_keys.comp,
//                     search.comp).
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

// ---------------------------------------------- partition offsets scan ----
// Exclusive scans of the partition key counts (offsets, the reference's
// prefix_sum.comp + prefix_sum_offset.comp with the host's leading 0,
// mphf_builder.h:254-257) and, for alpha < 1, of the remap entries per
// partition (table_size(count) - count).  Two launches over tiles of
// kScanTile counts: tile totals, then every tile adds the totals before it
// (at most 2^21 / kScanTile = 512 of them) and scans itself.

constexpr uint32_t kScanThreads = 256, kScanPer = 16, kScanTile = kScanThreads * kScanPer;

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* wred) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (uint32_t o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if (lane == 0) wred[wid] = v;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (uint32_t w = 0; w < kScanThreads / 64; w++) t += wred[w];
  return t;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_totals(const uint32_t* __restrict__ counts, uint32_t np,
                                                             uint32_t a_ppm, uint32_t* __restrict__ totals) {
  __shared__ uint32_t wred[kScanThreads / 64];
  uint32_t c_sum = 0, r_sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < kScanPer; j++) {
    const uint64_t idx = (uint64_t)blockIdx.x * kScanTile + j * kScanThreads + threadIdx.x;
    const uint32_t c = idx < np ? counts[idx] : 0;
    c_sum += c;
    if (a_ppm) r_sum += table_size(c, a_ppm) - c;
  }
  c_sum = block_sum(c_sum, wred);
  if (a_ppm) r_sum = block_sum(r_sum, wred);
  if (threadIdx.x == 0) {
    totals[2 * blockIdx.x] = c_sum;
    totals[2 * blockIdx.x + 1] = r_sum;
  }
}

// Exclusive block scan of one value per thread; returns the thread's prefix.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wred) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  __syncthreads();
  if (lane == 63) wred[wid] = incl;
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (uint32_t w = 0; w < kScanThreads / 64; w++)
    if (w < wid) before += wred[w];
  return before + incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_write(const uint32_t* __restrict__ counts, uint32_t np,
                                                            uint32_t a_ppm, const uint32_t* __restrict__ totals,
                                                            uint32_t* __restrict__ offsets,
                                                            uint32_t* __restrict__ remap_offsets) {
  __shared__ uint32_t wred[kScanThreads / 64];
  // totals of the tiles before this one
  uint32_t c_base = 0, r_base = 0;
  for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kScanThreads) {
    c_base += totals[2 * t];
    r_base += totals[2 * t + 1];
  }
  c_base = block_sum(c_base, wred);
  if (a_ppm) r_base = block_sum(r_base, wred);
  // thread t scans counts [tile0 + t*kScanPer, +kScanPer)
  const uint64_t i0 = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
  uint32_t c[kScanPer];
  uint32_t c_loc = 0, r_loc = 0;
#pragma unroll
  for (uint32_t j = 0; j < kScanPer; j++) {
    c[j] = i0 + j < np ? counts[i0 + j] : 0;
    c_loc += c[j];
    if (a_ppm) r_loc += table_size(c[j], a_ppm) - c[j];
  }
  uint32_t c_run = c_base + block_excl_scan(c_loc, wred);
  uint32_t r_run = a_ppm ? r_base + block_excl_scan(r_loc, wred) : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    offsets[0] = 0;
    if (a_ppm) remap_offsets[0] = 0;
  }
#pragma unroll
  for (uint32_t j = 0; j < kScanPer; j++) {
    c_run += c[j];
    if (a_ppm) r_run += table_size(c[j], a_ppm) - c[j];
    if (i0 + j < np) {
      offsets[i0 + j + 1] = c_run;
      if (a_ppm) remap_offsets[i0 + j + 1] = r_run;
    }
  }
}

// ------------------------------------------------------------- evaluate ---
// MPHF::operator() (mphf.hpp:77-86, hashPos mphf.hpp:32-40) over a device
// batch of keys: the initial hash of the key type (hasher.hpp:22-66), its
// partition and bucket, then offsets[p] + ((h >> 1) + V) mod m.  A key in a
// partition without keys (a key outside the build) gets offsets[p].

template <int KIND>
__device__ __forceinline__ U128 eval_hash(const KeySource& src, uint64_t i, uint64_t seed) {
  if constexpr (KIND == kKindU64) {
    return xxh3_128_u64(src.u64[i], seed);
  } else if constexpr (KIND == kKindBytes) {
    const uint64_t a = src.offs[i], b = src.offs[i + 1];
    return xxh3_128_bytes(src.bytes + a, b - a, seed);
  } else if constexpr (KIND == kKindPair) {
    return xxh3_128_bytes(reinterpret_cast<const uint8_t*>(src.keys + i), 16, seed);
  } else {
    const uint4 k = src.keys[i];
    return U128{(uint64_t)k.x | ((uint64_t)k.y << 32), (uint64_t)k.z | ((uint64_t)k.w << 32)};
  }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_eval(EvalArgs a) {
  __shared__ uint32_t fl[kFulcs];
  for (uint32_t i = threadIdx.x; i < kFulcs; i += blockDim.x) fl[i] = a.fulcs[i];
  __syncthreads();
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const bool remap = a.a_ppm < 1000000u;
  for (; i < a.count; i += stride) {
    const U128 h = eval_hash<KIND>(a.src, i, a.seed);
    const uint32_t p = partition_of((uint32_t)h.lo, a.np);
    const uint32_t b = bucket_of((uint32_t)(h.lo >> 32), fl);
    const uint32_t off = a.offsets[p];
    const uint32_t ps = a.offsets[p + 1] - off;
    if (ps == 0) {
      a.out[i] = off;
      continue;
    }
    const uint32_t V = a.pilots[p + (uint64_t)b * a.np];
    const uint32_t tm = table_size(ps, a.a_ppm);
    const uint32_t hp = V / tm;
    const uint32_t hv = mix32((uint32_t)h.hi, mix32((uint32_t)(h.hi >> 32), hp)) >> 1;
    uint32_t pos = (hv + V) % tm;
    if (remap && pos >= ps) pos = a.remap[a.remap_offsets[p] + pos - ps];
    a.out[i] = off + pos;
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

// Batched builds: set s owns global partitions [pb[s], pb[s+1]) and its
// np_s + 1 exclusive offsets start at out[pb[s] + s].
__global__ __launch_bounds__(256) void k_batch_offsets(const uint32_t* __restrict__ goff,
                                                       const uint32_t* __restrict__ pb, uint32_t sets,
                                                       uint32_t total, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  // set s with pb[s] + s <= t < pb[s+1] + s + 1
  uint32_t lo = 0, hi = sets;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (pb[mid] + mid <= t) lo = mid;
    else hi = mid;
  }
  const uint32_t j = t - (pb[lo] + lo);
  out[t] = goff[pb[lo] + j] - goff[pb[lo]];
}

// Plain streaming copy of 16-byte words (the known-size reference for the
// HBM traffic counters): every word read once and written once.
__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_batch_remap_bases(const uint32_t* __restrict__ goff,
                                                           const uint32_t* __restrict__ pb,
                                                           const uint32_t* __restrict__ rbase, uint32_t sets,
                                                           uint32_t np, uint32_t* __restrict__ out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  // the set s with pb[s] <= p < pb[s+1] (empty sets share their base)
  uint32_t lo = 0, hi = sets;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (pb[mid] <= p) lo = mid;
    else hi = mid;
  }
  while (lo + 1 < sets && pb[lo + 1] <= p) lo++;
  out[p] = rbase[lo] + goff[p] - goff[pb[lo]];
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

// Synthetic printable strings: string i fills [offs[i], offs[i+1]) with bytes
// 0x20..0x7E drawn from a splitmix64 stream of (seed, i), eight per output.
// Nothing makes two strings differ; at lengths of 8 and more, duplicates in
// a few 1e8 strings are unlikely, and a build reports any it meets.
__global__ __launch_bounds__(256) void k_gen_strings(uint64_t seed, uint64_t n, const uint64_t* __restrict__ offs,
                                                    uint8_t* __restrict__ bytes) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint64_t a = offs[i], len = offs[i + 1] - a;
    uint64_t state = seed ^ (i * 0xD1B54A32D192ED03ull);
    uint64_t r = 0;
    for (uint64_t j = 0; j < len; j++) {
      if (j % 8 == 0) {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        r = z ^ (z >> 31);
      }
      bytes[a + j] = (uint8_t)(0x20 + (uint32_t)((r & 0xFF) % 95u));
      r >>= 8;
    }
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

uint32_t partition_scan_tiles(uint32_t np) { return (np + kScanTile - 1) / kScanTile; }

hipError_t launch_partition_offsets(const uint32_t* counts, uint32_t np, uint32_t a_ppm, uint32_t* offsets,
                                    uint32_t* remap_offsets, uint32_t* totals, hipStream_t s) {
  const uint32_t tiles = partition_scan_tiles(np);
  if (tiles == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, offsets, 1ull, 0u);
    if (a_ppm) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, remap_offsets, 1ull, 0u);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_totals, dim3(tiles), dim3(kScanThreads), 0, s, counts, np, a_ppm, totals);
  hipLaunchKernelGGL(k_scan_write, dim3(tiles), dim3(kScanThreads), 0, s, counts, np, a_ppm, totals, offsets,
                     remap_offsets);
  return hipGetLastError();
}

hipError_t launch_eval(int kind, const EvalArgs& a, hipStream_t s) {
  if (a.count == 0) return hipSuccess;
  const dim3 grid(stream_grid(a.count, 256, 256 * 32)), block(256);
  switch (kind) {
    case kKindU64: hipLaunchKernelGGL(k_eval<kKindU64>, grid, block, 0, s, a); break;
    case kKindBytes: hipLaunchKernelGGL(k_eval<kKindBytes>, grid, block, 0, s, a); break;
    case kKindPair: hipLaunchKernelGGL(k_eval<kKindPair>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_eval<kKindKey>, grid, block, 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_batch_offsets(const uint32_t* global_offsets, const uint32_t* set_part_base, uint32_t sets,
                                uint32_t total_entries, uint32_t* out, hipStream_t s) {
  if (total_entries == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_offsets, dim3((total_entries + 255) / 256), dim3(256), 0, s, global_offsets,
                     set_part_base, sets, total_entries, out);
  return hipGetLastError();
}

hipError_t launch_batch_remap_bases(const uint32_t* goff, const uint32_t* set_part_base,
                                    const uint32_t* set_remap_base, uint32_t sets, uint32_t np, uint32_t* out,
                                    hipStream_t s) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_remap_bases, dim3((np + 255) / 256), dim3(256), 0, s, goff, set_part_base,
                     set_remap_base, sets, np, out);
  return hipGetLastError();
}

hipError_t launch_check_bijection(const uint32_t* pos, uint64_t count, uint64_t n, uint32_t* bitmap,
                                  unsigned long long* bad, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_bijection, dim3(stream_grid(count, 256, 256 * 32)), dim3(256), 0, s, pos, count, n,
                     bitmap, bad);
  return hipGetLastError();
}

hipError_t launch_copy16(void* dst, const void* src, uint64_t words, hipStream_t s) {
  if (words == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy16, dim3(stream_grid(words, 256, 256 * 64)), dim3(256), 0, s, static_cast<uint4*>(dst),
                     static_cast<const uint4*>(src), words);
  return hipGetLastError();
}

hipError_t launch_splitmix64(uint64_t seed, uint64_t n, uint64_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_splitmix64, dim3(stream_grid(n, 256, 256 * 32)), dim3(256), 0, s, seed, n, out);
  return hipGetLastError();
}

hipError_t launch_gen_strings(uint64_t seed, uint64_t n, const uint64_t* offs, uint8_t* bytes, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gen_strings, dim3(stream_grid(n, 256, 256 * 8)), dim3(256), 0, s, seed, n, offs, bytes);
  return hipGetLastError();
}

}  // namespace phobic
