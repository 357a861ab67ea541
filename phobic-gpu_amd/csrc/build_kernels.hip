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
//   k_partition_offsets  exclusive scan of partition key counts
//                     (prefix_sum.comp + prefix_sum_offset.comp).
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
// One block of 1024 threads, 16 counts per thread per chunk, carried across
// chunks.  offsets[0] = 0, offsets[p+1] = sum of counts[0..p].

__global__ __launch_bounds__(1024) void k_partition_offsets(const uint32_t* __restrict__ counts, uint32_t np,
                                                            uint32_t a_ppm, uint32_t* __restrict__ offsets) {
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
      const uint32_t c = idx < np ? counts[idx] : 0;
      v[j] = a_ppm ? table_size(c, a_ppm) - c : c;
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
    const uint32_t tm = table_size(ps, a.a_ppm);
    const uint32_t hp = V / tm;
    const uint32_t hv = mix32((uint32_t)h.hi, mix32((uint32_t)(h.hi >> 32), hp)) >> 1;
    uint32_t pos = (hv + V) % tm;
    if (pos >= ps) pos = a.remap[a.remap_offsets[p] + pos - ps];
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

hipError_t launch_partition_offsets(const uint32_t* counts, uint32_t np, uint32_t a_ppm, uint32_t* offsets,
                                    hipStream_t s) {
  hipLaunchKernelGGL(k_partition_offsets, dim3(1), dim3(1024), 0, s, counts, np, a_ppm, offsets);
  return hipGetLastError();
}

hipError_t launch_eval_u64(const EvalArgs& a, hipStream_t s) {
  if (a.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_eval_u64, dim3(stream_grid(a.count, 256, 256 * 32)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_batch_offsets(const uint32_t* global_offsets, const uint32_t* set_part_base, uint32_t sets,
                                uint32_t total_entries, uint32_t* out, hipStream_t s) {
  if (total_entries == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_offsets, dim3((total_entries + 255) / 256), dim3(256), 0, s, global_offsets,
                     set_part_base, sets, total_entries, out);
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
