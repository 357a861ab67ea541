// hasher.hpp -- Key and the initial hashers (reference include/phobicGpu/hasher.hpp:7-67).
// xxhash runs XXH3-128 on the host for queries (builds hash on the device).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "phobic_hip.h"

namespace phobicgpu {

struct Key {
  uint32_t partitioner, bucketer, lower1, lower2;
  Key() {}
  Key(uint32_t partitioner, uint32_t bucketer, uint32_t lower1, uint32_t lower2)
      : partitioner(partitioner), bucketer(bucketer), lower1(lower1), lower2(lower2) {}
};

struct nohash {
  static inline const Key& hash(const Key& val, uint64_t = 0) { return val; }
};

struct xxhash {
  // {lo32(low64), hi32(low64), lo32(high64), hi32(high64)} (hasher.hpp:36-37)
  static inline Key of(const void* p, size_t len, uint64_t seed) {
    uint64_t lo, hi;
    if (phobic_xxh3_128(p, len, seed, &lo, &hi) != PHOBIC_OK) throw std::runtime_error("phobic_xxh3_128 failed");
    return Key((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
  }
  static inline Key hash(const std::string& val, uint64_t seed = 0) { return of(val.data(), val.size(), seed); }
  static inline Key hash(uint64_t val, uint64_t seed = 0) { return of(&val, sizeof(val), seed); }
  static inline Key hash(std::pair<uint64_t, uint64_t> val, uint64_t seed = 0) {
    uint64_t b[2] = {val.first, val.second};
    return of(b, sizeof(b), seed);
  }
  static inline Key hash(const Key& val, uint64_t seed = 0) { return of(&val, sizeof(val), seed); }
};

}  // namespace phobicgpu
