// mphf.hpp -- the minimal perfect hash function: partition offsets, fulcrums
// and the pilot encoder (reference include/phobicGpu/mphf.hpp:15-98).
// Evaluation: partition = hi32(partitioner * np), bucket from the fulcrum
// table, pilot V, position = offset + ((hash(l1, hash(l2, V/ps)) >> 1) + V) mod ps.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../encoders/encoders.hpp"
#include "hasher.hpp"
#include "mphf_config.h"

namespace phobicgpu {

// Writes the fields an encoder visits as little-endian u64 words; vectors
// as their length followed by their words.
struct ByteSaver {
  std::vector<uint8_t>& out;
  void put(uint64_t v) {
    for (int i = 0; i < 8; i++) out.push_back(uint8_t(v >> (8 * i)));
  }
  void visit(uint64_t& v) { put(v); }
  void visit(std::vector<uint64_t>& v) {
    put(v.size());
    for (uint64_t w : v) put(w);
  }
  template <typename T>
  void visit(T& t) {
    t.visit(*this);
  }
};

template <typename PilotEncoder, typename PartitionOffsetEncoder, typename Hasher>
class MPHF {
 private:
  std::vector<uint32_t> fulcs;
  uint32_t partitions = 0;
  uint32_t partitionSize = 0;
  uint64_t seed = 0;
  uint64_t keys = 0;
  PilotEncoder pilots;
  PartitionOffsetEncoder partitionOffsets;
  // alpha < 1 (extension): table sizes and slot remapping
  MPHFconfig cfg;
  compact remap, remapOffsets;
  bool hasRemap = false;

  // hashing.glsl:11-17 / mphf.hpp:24-30
  static inline uint32_t hash(uint32_t pilot, uint32_t key) {
    key ^= pilot;
    key = ((key >> 16) ^ key) * 0x45d9f3bu;
    key = ((key >> 16) ^ key) * 0x45d9f3bu;
    return (key >> 16) ^ key;
  }

  static inline uint32_t hashPos(uint32_t pilot, uint32_t lower1, uint32_t lower2, uint32_t tableSize) {
    const uint32_t hp = pilot / tableSize;
    const uint32_t hv = hash(lower1, hash(lower2, hp)) >> 1;
    return (hv + pilot) % tableSize;
  }

  // mphf.hpp:42-49
  inline uint64_t getBucket(uint64_t bucketBits) const {
    const uint64_t z = bucketBits * uint64_t(FULCS_INTER - 1);
    const uint64_t index = z >> 32, part = z & 0xFFFFFFFFu;
    const uint64_t v1 = (uint64_t(fulcs[index]) * part) >> 32;
    const uint64_t v2 = (uint64_t(fulcs[index + 1]) * (0xFFFFFFFFu - part)) >> 32;
    return (v1 + v2) >> 16;
  }

 public:
  PilotEncoder& getPilotEncoder() { return pilots; }

  constexpr static bool noHash() { return std::is_same_v<Hasher, nohash>; }

  template <typename keyType>
  inline Key initialHash(const keyType& keyRaw) const {
    return Hasher::hash(keyRaw, seed);
  }

  // pilots: interleaved p + b*np; offsets: np+1 exclusive (mphf.hpp:65-75)
  void setData(const std::vector<uint32_t>& pilotsIn, std::vector<uint32_t>& offsetsIn, uint32_t parts,
               MPHFconfig config) {
    fulcs = config.getFulcs();
    partitions = parts;
    partitionSize = config.partitionSize;
    seed = config.seed;
    cfg = config;
    keys = offsetsIn.empty() ? 0 : offsetsIn.back();
    partitionOffsets.encode(offsetsIn.begin(), config.partitionSize, parts + 1);
    if (parts) pilots.encode(pilotsIn.begin(), parts, config.bucketCountPerPartition);
  }

  template <typename keyType>
  inline uint32_t operator()(const keyType& keyRaw) const {
    const Key key = initialHash(keyRaw);
    const uint64_t partition = (uint64_t(key.partitioner) * uint64_t(partitions)) >> 32;
    const uint64_t bucket = getBucket(key.bucketer);
    const uint32_t pilot = (uint32_t)pilots.access(partition, bucket);
    const uint32_t off = (uint32_t)partitionOffsets.access(partition);
    const uint32_t ps = (uint32_t)partitionOffsets.access(partition + 1) - off;
    if (!hasRemap) return off + hashPos(pilot, key.lower1, key.lower2, ps);
    uint32_t pos = hashPos(pilot, key.lower1, key.lower2, cfg.tableSize(ps));
    if (pos >= ps) pos = (uint32_t)remap.access(remapOffsets.access(partition) + pos - ps);
    return off + pos;
  }

  // alpha < 1 (extension): slots [ps, m) of partition p stand for
  // remap[remapOffsets[p] + q - ps]
  void setRemap(const std::vector<uint32_t>& entries, const std::vector<uint32_t>& offsets) {
    remap.encode(entries.begin(), entries.size());
    remapOffsets.encode(offsets.begin(), offsets.size());
    hasRemap = true;
  }

  float getBitsPerKey() const {
    const uint64_t rbits = hasRemap ? remap.num_bits() + remapOffsets.num_bits() : 0;
    return float(pilots.num_bits() + partitionOffsets.num_bits() + rbits + 32 * fulcs.size() +
                 32 * uint64_t(partitions)) /
           float(keys);
  }

  std::string getResultLine() {
    const double n = double(keys);
    return "total_bits=" + std::to_string(getBitsPerKey()) + " pilot_bits=" + std::to_string(pilots.num_bits() / n) +
           " offsets_bits=" + std::to_string(partitionOffsets.num_bits() / n);
  }

  uint64_t size() const { return keys; }

  // Encoded MPHF bytes: magic "PHOBMPH1", n, P, np, B, the 2048 fulcrums
  // (u32 pairs per u64 word), the pilot encoder's fields, the partition
  // offset encoder's fields (all little-endian u64).
  std::vector<uint8_t> serialize() {
    std::vector<uint8_t> out;
    ByteSaver s{out};
    s.put(0x3148504d424f4850ull);
    s.put(keys);
    s.put(partitionSize);
    s.put(partitions);
    const uint64_t B = fulcs.empty() ? 0 : (fulcs.back() >> 16);
    s.put(B);
    for (size_t i = 0; i + 1 < fulcs.size(); i += 2) s.put(uint64_t(fulcs[i]) | (uint64_t(fulcs[i + 1]) << 32));
    if (partitions) pilots.visit(s);
    partitionOffsets.visit(s);
    if (hasRemap) {  // extension section, present only for alpha < 1
      s.put(cfg.alphaPpm());
      remapOffsets.visit(s);
      remap.visit(s);
    }
    return out;
  }
};

}  // namespace phobicgpu
