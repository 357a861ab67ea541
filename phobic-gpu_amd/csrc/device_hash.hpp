// device_hash.hpp -- device-side hashing and key mapping for the PHOBIC
// construction path (gfx950).
//
//  * XXH3-128 (xxHash 0.8.x algorithm) of a u64 key or a byte string; the
//    reference hashes on the host with XXH128(ptr, len, 0)
//    (include/phobicGpu/hasher.hpp:33-61); here it runs inside the first
//    streaming kernel.
//  * Key split {partitioner, bucketer, lower1, lower2} = {lo32(low64),
//    hi32(low64), lo32(high64), hi32(high64)} (hasher.hpp:36-37).
//  * partition / bucket mapping (shaders/key_mapping.glsl:9-49) and the
//    pilot hash (shaders/hashing.glsl:11-17).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phobic {

constexpr uint32_t kFulcs = 2048;  // shaders/constants.glsl:1

// XXH3 primes and the 192-byte default secret (format constants of XXH3).
constexpr uint64_t kP32_1 = 0x9E3779B1u, kP32_2 = 0x85EBCA77u, kP32_3 = 0xC2B2AE3Du;
constexpr uint64_t kP64_1 = 0x9E3779B185EBCA87ull, kP64_2 = 0xC2B2AE3D27D4EB4Full,
                   kP64_3 = 0x165667B19E3779F9ull, kP64_4 = 0x85EBCA77C2B2AE63ull,
                   kP64_5 = 0x27D4EB2F165667C5ull;
constexpr uint64_t kMX1 = 0x165667919E3779F9ull, kMX2 = 0x9FB21C651E98DF25ull;

__device__ __constant__ static const uint64_t kSecret64[24] = {
    0xbe4ba423396cfeb8ull, 0x1cad21f72c81017cull, 0xdb979083e96dd4deull, 0x1f67b3b7a4a44072ull,
    0x78e5c0cc4ee679cbull, 0x2172ffcc7dd05a82ull, 0x8e2443f7744608b8ull, 0x4c263a81e69035e0ull,
    0xcb00c391bb52283cull, 0xa32e531b8b65d088ull, 0x4ef90da297486471ull, 0xd8acdea946ef1938ull,
    0x3f349ce33f76faa8ull, 0x1d4f0bc7c7bbdcf9ull, 0x3159b4cd4be0518aull, 0x647378d9c97e9fc8ull,
    0xc3ebd33483acc5eaull, 0xeb6313faffa081c5ull, 0x49daf0b751dd0d17ull, 0x9e68d429265516d3ull,
    0xfca1477d58be162bull, 0xce31d07ad1b8f88full, 0x280416958f3acb45ull, 0x7e404bbbcafbd7afull};

// Secret word at an arbitrary byte offset (the 129..240 and long paths read
// the secret at odd offsets).
__device__ __forceinline__ uint64_t sec64(uint32_t off) {
  uint32_t w = off >> 3, sh = (off & 7) * 8;
  uint64_t a = kSecret64[w];
  if (sh == 0) return a;
  uint64_t b = kSecret64[w + 1];
  return (a >> sh) | (b << (64 - sh));
}
__device__ __forceinline__ uint32_t sec32(uint32_t off) { return (uint32_t)sec64(off); }

struct U128 {
  uint64_t lo, hi;
};

// The 128-bit product from four 32x32->64 multiply-adds (v_mad_u64_u32),
// each partial product formed once.  No sum overflows: t and u are at most
// (2^32-1)^2 + 2^32-1 < 2^64, and the high word is the exact quotient.
__device__ __forceinline__ U128 mul128(uint64_t a, uint64_t b) {
  const uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32), bl = (uint32_t)b, bh = (uint32_t)(b >> 32);
  const uint64_t p0 = (uint64_t)al * bl;
  const uint64_t t = (uint64_t)ah * bl + (p0 >> 32);
  const uint64_t u = (uint64_t)al * bh + (uint32_t)t;
  return {(u << 32) | (uint32_t)p0, (uint64_t)ah * bh + (t >> 32) + (u >> 32)};
}
__device__ __forceinline__ uint64_t fold64(uint64_t a, uint64_t b) {
  const U128 m = mul128(a, b);
  return m.lo ^ m.hi;
}
__device__ __forceinline__ uint64_t avalanche3(uint64_t h) {
  h ^= h >> 37;
  h *= kMX1;
  return h ^ (h >> 32);
}
__device__ __forceinline__ uint64_t avalanche64(uint64_t h) {
  h ^= h >> 33;
  h *= kP64_2;
  h ^= h >> 29;
  h *= kP64_3;
  return h ^ (h >> 32);
}
__device__ __forceinline__ uint32_t bswap32d(uint32_t x) { return __builtin_bswap32(x); }
__device__ __forceinline__ uint64_t bswap64d(uint64_t x) { return __builtin_bswap64(x); }

// XXH3-128 of an 8-byte little-endian key: the 4..8 byte path with len = 8,
// where (lo32 | hi32 << 32) is the key value itself.
__device__ __forceinline__ U128 xxh3_128_u64(uint64_t key, uint64_t seed) {
  uint64_t s = seed ^ ((uint64_t)bswap32d((uint32_t)seed) << 32);
  const uint64_t flip = (0xdb979083e96dd4deull ^ 0x1f67b3b7a4a44072ull);  // secret[16..32)
  U128 h = mul128(key ^ (flip + s), kP64_1 + (8ull << 2));
  h.hi += h.lo << 1;
  h.lo ^= h.hi >> 3;
  h.lo ^= h.lo >> 35;
  h.lo *= kMX2;
  h.lo ^= h.lo >> 28;
  h.hi = avalanche3(h.hi);
  return h;
}

// Byte reader over a global string.  Reads are byte loads combined in
// registers, so any alignment is valid and nothing past the string is read.
struct ByteSrc {
  const uint8_t* p;
  __device__ __forceinline__ uint32_t b(uint64_t i) const { return p[i]; }
  __device__ __forceinline__ uint32_t r32(uint64_t i) const {
    return b(i) | (b(i + 1) << 8) | (b(i + 2) << 16) | (b(i + 3) << 24);
  }
  __device__ __forceinline__ uint64_t r64(uint64_t i) const {
    return (uint64_t)r32(i) | ((uint64_t)r32(i + 4) << 32);
  }
};

__device__ __forceinline__ uint64_t mix16(const ByteSrc& in, uint64_t at, uint32_t so, uint64_t seed) {
  return fold64(in.r64(at) ^ (sec64(so) + seed), in.r64(at + 8) ^ (sec64(so + 8) - seed));
}

__device__ __forceinline__ U128 mix32(U128 acc, const ByteSrc& in, uint64_t a, uint64_t b, uint32_t so,
                                      uint64_t seed) {
  acc.lo += mix16(in, a, so, seed);
  acc.lo ^= in.r64(b) + in.r64(b + 8);
  acc.hi += mix16(in, b, so + 16, seed);
  acc.hi ^= in.r64(a) + in.r64(a + 8);
  return acc;
}

// Long inputs (> 240 bytes): the striped accumulator.  The secret for a
// non-zero seed is derived word by word (custom secret: +seed / -seed on
// alternating 64-bit words).
__device__ __forceinline__ uint64_t lsec64(uint32_t off, uint64_t seed) {
  // custom secret word at an arbitrary byte offset
  uint32_t w = off >> 3, sh = (off & 7) * 8;
  auto word = [&](uint32_t i) -> uint64_t {
    uint64_t v = kSecret64[i];
    return (i & 1) ? v - seed : v + seed;
  };
  uint64_t a = word(w);
  if (sh == 0) return a;
  return (a >> sh) | (word(w + 1) << (64 - sh));
}

__device__ __noinline__ U128 xxh3_128_long(const ByteSrc& in, uint64_t len, uint64_t seed) {
  uint64_t acc[8] = {kP32_3, kP64_1, kP64_2, kP64_3, kP64_4, kP32_2, kP64_5, kP32_1};
  const uint64_t spb = (192 - 64) / 8, block = 64 * spb;
  const uint64_t nb = (len - 1) / block;
  auto stripe = [&](uint64_t at, uint32_t so) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint64_t v = in.r64(at + 8 * i);
      uint64_t k = v ^ lsec64(so + 8 * i, seed);
      acc[i ^ 1] += v;
      acc[i] += (uint64_t)(uint32_t)k * (k >> 32);
    }
  };
  for (uint64_t b = 0; b < nb; b++) {
    for (uint64_t k = 0; k < spb; k++) stripe(b * block + 64 * k, (uint32_t)(8 * k));
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint64_t a = acc[i];
      a ^= a >> 47;
      a ^= lsec64(192 - 64 + 8 * i, seed);
      acc[i] = a * kP32_1;
    }
  }
  const uint64_t last = ((len - 1) - block * nb) / 64;
  for (uint64_t k = 0; k < last; k++) stripe(nb * block + 64 * k, (uint32_t)(8 * k));
  stripe(len - 64, 192 - 64 - 7);
  auto merge = [&](uint32_t so, uint64_t start) {
    uint64_t r = start;
#pragma unroll
    for (int i = 0; i < 4; i++)
      r += fold64(acc[2 * i] ^ lsec64(so + 16 * i, seed), acc[2 * i + 1] ^ lsec64(so + 16 * i + 8, seed));
    return avalanche3(r);
  };
  U128 h;
  h.lo = merge(11, len * kP64_1);
  h.hi = merge(192 - 64 - 11, ~(len * kP64_2));
  return h;
}

// XXH3-128 of an arbitrary byte string.
__device__ __forceinline__ U128 xxh3_128_bytes(const uint8_t* p, uint64_t len, uint64_t seed) {
  ByteSrc in{p};
  U128 h;
  if (len == 0) {
    h.lo = avalanche64(seed ^ (sec64(64) ^ sec64(72)));
    h.hi = avalanche64(seed ^ (sec64(80) ^ sec64(88)));
  } else if (len <= 3) {
    uint32_t c1 = in.b(0), c2 = in.b(len >> 1), c3 = in.b(len - 1);
    uint32_t cl = (c1 << 16) | (c2 << 24) | c3 | ((uint32_t)len << 8);
    uint32_t ch = bswap32d(cl);
    ch = (ch << 13) | (ch >> 19);
    uint64_t fl = (uint64_t)(sec32(0) ^ sec32(4)) + seed;
    uint64_t fh = (uint64_t)(sec32(8) ^ sec32(12)) - seed;
    h.lo = avalanche64((uint64_t)cl ^ fl);
    h.hi = avalanche64((uint64_t)ch ^ fh);
  } else if (len <= 8) {
    uint64_t s = seed ^ ((uint64_t)bswap32d((uint32_t)seed) << 32);
    uint64_t x = (uint64_t)in.r32(0) + ((uint64_t)in.r32(len - 4) << 32);
    uint64_t flip = (sec64(16) ^ sec64(24)) + s;
    h = mul128(x ^ flip, kP64_1 + (len << 2));
    h.hi += h.lo << 1;
    h.lo ^= h.hi >> 3;
    h.lo ^= h.lo >> 35;
    h.lo *= kMX2;
    h.lo ^= h.lo >> 28;
    h.hi = avalanche3(h.hi);
  } else if (len <= 16) {
    uint64_t fl = (sec64(32) ^ sec64(40)) - seed;
    uint64_t fh = (sec64(48) ^ sec64(56)) + seed;
    uint64_t a = in.r64(0), b = in.r64(len - 8);
    U128 m = mul128(a ^ b ^ fl, kP64_1);
    m.lo += (len - 1) << 54;
    b ^= fh;
    m.hi += b + (uint64_t)(uint32_t)b * (kP32_2 - 1);
    m.lo ^= bswap64d(m.hi);
    h = mul128(m.lo, kP64_2);
    h.hi += m.hi * kP64_2;
    h.lo = avalanche3(h.lo);
    h.hi = avalanche3(h.hi);
  } else if (len <= 128) {
    U128 acc{len * kP64_1, 0};
    if (len > 32) {
      if (len > 64) {
        if (len > 96) acc = mix32(acc, in, 48, len - 64, 96, seed);
        acc = mix32(acc, in, 32, len - 48, 64, seed);
      }
      acc = mix32(acc, in, 16, len - 32, 32, seed);
    }
    acc = mix32(acc, in, 0, len - 16, 0, seed);
    h.lo = avalanche3(acc.lo + acc.hi);
    h.hi = 0 - avalanche3(acc.lo * kP64_1 + acc.hi * kP64_4 + (len - seed) * kP64_2);
  } else if (len <= 240) {
    U128 acc{len * kP64_1, 0};
    const uint32_t rounds = (uint32_t)(len / 32);
    for (uint32_t i = 0; i < 4; i++) acc = mix32(acc, in, 32 * i, 32 * i + 16, 32 * i, seed);
    acc.lo = avalanche3(acc.lo);
    acc.hi = avalanche3(acc.hi);
    for (uint32_t i = 4; i < rounds; i++) acc = mix32(acc, in, 32 * i, 32 * i + 16, 3 + 32 * (i - 4), seed);
    acc = mix32(acc, in, len - 16, len - 32, 136 - 17 - 16, 0 - seed);
    h.lo = avalanche3(acc.lo + acc.hi);
    h.hi = 0 - avalanche3(acc.lo * kP64_1 + acc.hi * kP64_4 + (len - seed) * kP64_2);
  } else {
    h = xxh3_128_long(in, len, seed);
  }
  return h;
}

// ------------------------------------------------------------ mapping ----

// shaders/key_mapping.glsl:43-45
__device__ __forceinline__ uint32_t partition_of(uint32_t partitioner, uint32_t np) {
  return __umulhi(partitioner, np);
}

// shaders/key_mapping.glsl:34-41.  fulcs[i] is weighted by the fractional
// part and fulcs[i+1] by its complement, as in the reference.
__device__ __forceinline__ uint32_t bucket_of(uint32_t bucketer, const uint32_t* __restrict__ fulcs) {
  uint32_t part = bucketer * (kFulcs - 1);
  uint32_t idx = __umulhi(bucketer, kFulcs - 1);
  uint32_t v1 = __umulhi(fulcs[idx], part);
  uint32_t v2 = __umulhi(fulcs[idx + 1], 0xFFFFFFFFu - part);
  return (v1 + v2) >> 16;
}

// shaders/hashing.glsl:11-17
__device__ __forceinline__ uint32_t mix32(uint32_t key, uint32_t pilot) {
  key ^= pilot;
  key = ((key >> 16) ^ key) * 0x45d9f3bu;
  key = ((key >> 16) ^ key) * 0x45d9f3bu;
  return (key >> 16) ^ key;
}

}  // namespace phobic
