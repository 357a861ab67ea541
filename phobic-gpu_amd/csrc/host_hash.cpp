// host_hash.cpp -- host XXH3-128 for the query side of the C++ API.
//
// MPHF::operator() (include/phobicGpu/mphf.hpp:77-86) hashes each queried
// key on the host with XXH128(ptr, len, seed) (hasher.hpp:33-61); builds hash
// on the device (device_hash.hpp).  Restated from the xxHash 0.8.x XXH3
// algorithm; the default secret is the algorithm's format constant.
#include <cstdint>
#include <cstring>

#include "../../include/phobic_hip.h"

namespace {

using u64 = uint64_t;
using u32 = uint32_t;
using u128 = unsigned __int128;

constexpr u64 P32_1 = 0x9E3779B1u, P32_2 = 0x85EBCA77u, P32_3 = 0xC2B2AE3Du;
constexpr u64 P64_1 = 0x9E3779B185EBCA87ull, P64_2 = 0xC2B2AE3D27D4EB4Full, P64_3 = 0x165667B19E3779F9ull,
              P64_4 = 0x85EBCA77C2B2AE63ull, P64_5 = 0x27D4EB2F165667C5ull;
constexpr u64 MX1 = 0x165667919E3779F9ull, MX2 = 0x9FB21C651E98DF25ull;

constexpr u64 kSecretWords[24] = {
    0xbe4ba423396cfeb8ull, 0x1cad21f72c81017cull, 0xdb979083e96dd4deull, 0x1f67b3b7a4a44072ull,
    0x78e5c0cc4ee679cbull, 0x2172ffcc7dd05a82ull, 0x8e2443f7744608b8ull, 0x4c263a81e69035e0ull,
    0xcb00c391bb52283cull, 0xa32e531b8b65d088ull, 0x4ef90da297486471ull, 0xd8acdea946ef1938ull,
    0x3f349ce33f76faa8ull, 0x1d4f0bc7c7bbdcf9ull, 0x3159b4cd4be0518aull, 0x647378d9c97e9fc8ull,
    0xc3ebd33483acc5eaull, 0xeb6313faffa081c5ull, 0x49daf0b751dd0d17ull, 0x9e68d429265516d3ull,
    0xfca1477d58be162bull, 0xce31d07ad1b8f88full, 0x280416958f3acb45ull, 0x7e404bbbcafbd7afull};

struct Secret {
  uint8_t b[192];
  explicit Secret(u64 seed) {  // custom secret for a seed (+seed / -seed per 64-bit word)
    for (int i = 0; i < 24; i++) {
      const u64 w = kSecretWords[i] + ((i & 1) ? (u64)0 - seed : seed);
      std::memcpy(b + 8 * i, &w, 8);
    }
  }
};

inline u64 r64(const uint8_t* p) {
  u64 v;
  std::memcpy(&v, p, 8);
  return v;
}
inline u32 r32(const uint8_t* p) {
  u32 v;
  std::memcpy(&v, p, 4);
  return v;
}
inline u64 fold(u64 a, u64 b) {
  const u128 r = (u128)a * b;
  return (u64)r ^ (u64)(r >> 64);
}
inline u64 av3(u64 h) {
  h ^= h >> 37;
  h *= MX1;
  return h ^ (h >> 32);
}
inline u64 av64(u64 h) {
  h ^= h >> 33;
  h *= P64_2;
  h ^= h >> 29;
  h *= P64_3;
  return h ^ (h >> 32);
}

struct H {
  u64 lo, hi;
};

inline u64 mix16(const uint8_t* in, const uint8_t* sec, u64 seed) {
  return fold(r64(in) ^ (r64(sec) + seed), r64(in + 8) ^ (r64(sec + 8) - seed));
}
inline H mix32(H acc, const uint8_t* a, const uint8_t* b, const uint8_t* sec, u64 seed) {
  acc.lo += mix16(a, sec, seed);
  acc.lo ^= r64(b) + r64(b + 8);
  acc.hi += mix16(b, sec + 16, seed);
  acc.hi ^= r64(a) + r64(a + 8);
  return acc;
}

H xxh3_128(const uint8_t* in, u64 len, u64 seed) {
  static const Secret def(0);
  const uint8_t* sec = def.b;
  H h;
  if (len == 0) {
    h.lo = av64(seed ^ (r64(sec + 64) ^ r64(sec + 72)));
    h.hi = av64(seed ^ (r64(sec + 80) ^ r64(sec + 88)));
  } else if (len <= 3) {
    const u32 cl = ((u32)in[0] << 16) | ((u32)in[len >> 1] << 24) | in[len - 1] | ((u32)len << 8);
    u32 ch = __builtin_bswap32(cl);
    ch = (ch << 13) | (ch >> 19);
    h.lo = av64((u64)cl ^ ((u64)(r32(sec) ^ r32(sec + 4)) + seed));
    h.hi = av64((u64)ch ^ ((u64)(r32(sec + 8) ^ r32(sec + 12)) - seed));
  } else if (len <= 8) {
    const u64 s = seed ^ ((u64)__builtin_bswap32((u32)seed) << 32);
    const u64 x = (u64)r32(in) + ((u64)r32(in + len - 4) << 32);
    const u128 m = (u128)(x ^ ((r64(sec + 16) ^ r64(sec + 24)) + s)) * (P64_1 + (len << 2));
    h.lo = (u64)m;
    h.hi = (u64)(m >> 64);
    h.hi += h.lo << 1;
    h.lo ^= h.hi >> 3;
    h.lo ^= h.lo >> 35;
    h.lo *= MX2;
    h.lo ^= h.lo >> 28;
    h.hi = av3(h.hi);
  } else if (len <= 16) {
    const u64 fl = (r64(sec + 32) ^ r64(sec + 40)) - seed, fh = (r64(sec + 48) ^ r64(sec + 56)) + seed;
    u64 a = r64(in), b = r64(in + len - 8);
    u128 m = (u128)(a ^ b ^ fl) * P64_1;
    u64 mlo = (u64)m + ((len - 1) << 54), mhi = (u64)(m >> 64);
    b ^= fh;
    mhi += b + (u64)(u32)b * (P32_2 - 1);
    mlo ^= __builtin_bswap64(mhi);
    const u128 r = (u128)mlo * P64_2;
    h.lo = av3((u64)r);
    h.hi = av3((u64)(r >> 64) + mhi * P64_2);
  } else if (len <= 240) {
    H acc{len * P64_1, 0};
    if (len <= 128) {
      if (len > 32) {
        if (len > 64) {
          if (len > 96) acc = mix32(acc, in + 48, in + len - 64, sec + 96, seed);
          acc = mix32(acc, in + 32, in + len - 48, sec + 64, seed);
        }
        acc = mix32(acc, in + 16, in + len - 32, sec + 32, seed);
      }
      acc = mix32(acc, in, in + len - 16, sec, seed);
    } else {
      const u64 rounds = len / 32;
      for (u64 i = 0; i < 4; i++) acc = mix32(acc, in + 32 * i, in + 32 * i + 16, sec + 32 * i, seed);
      acc.lo = av3(acc.lo);
      acc.hi = av3(acc.hi);
      for (u64 i = 4; i < rounds; i++) acc = mix32(acc, in + 32 * i, in + 32 * i + 16, sec + 3 + 32 * (i - 4), seed);
      acc = mix32(acc, in + len - 16, in + len - 32, sec + 136 - 17 - 16, (u64)0 - seed);
    }
    h.lo = av3(acc.lo + acc.hi);
    h.hi = (u64)0 - av3(acc.lo * P64_1 + acc.hi * P64_4 + (len - seed) * P64_2);
  } else {
    const Secret custom(seed);
    const uint8_t* s = seed ? custom.b : sec;
    u64 acc[8] = {P32_3, P64_1, P64_2, P64_3, P64_4, P32_2, P64_5, P32_1};
    auto stripe = [&](const uint8_t* p, const uint8_t* k) {
      for (int i = 0; i < 8; i++) {
        const u64 v = r64(p + 8 * i), x = v ^ r64(k + 8 * i);
        acc[i ^ 1] += v;
        acc[i] += (u64)(u32)x * (x >> 32);
      }
    };
    const u64 spb = (192 - 64) / 8, block = 64 * spb, nb = (len - 1) / block;
    for (u64 b = 0; b < nb; b++) {
      for (u64 k = 0; k < spb; k++) stripe(in + b * block + 64 * k, s + 8 * k);
      for (int i = 0; i < 8; i++) {
        u64 a = acc[i];
        a ^= a >> 47;
        a ^= r64(s + 192 - 64 + 8 * i);
        acc[i] = a * P32_1;
      }
    }
    const u64 last = ((len - 1) - block * nb) / 64;
    for (u64 k = 0; k < last; k++) stripe(in + nb * block + 64 * k, s + 8 * k);
    stripe(in + len - 64, s + 192 - 64 - 7);
    auto merge = [&](const uint8_t* k, u64 start) {
      u64 r = start;
      for (int i = 0; i < 4; i++) r += fold(acc[2 * i] ^ r64(k + 16 * i), acc[2 * i + 1] ^ r64(k + 16 * i + 8));
      return av3(r);
    };
    h.lo = merge(s + 11, len * P64_1);
    h.hi = merge(s + 192 - 64 - 11, ~(len * P64_2));
  }
  return h;
}

}  // namespace

extern "C" int phobic_xxh3_128(const void* data, uint64_t len, uint64_t seed, uint64_t* lo64, uint64_t* hi64) {
  if ((len && !data) || !lo64 || !hi64) return PHOBIC_E_INVALID_CONFIG;
  const H h = xxh3_128(static_cast<const uint8_t*>(data), len, seed);
  *lo64 = h.lo;
  *hi64 = h.hi;
  return PHOBIC_OK;
}
