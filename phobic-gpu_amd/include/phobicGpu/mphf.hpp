// mphf.hpp -- the minimal perfect hash function: partition offsets, fulcrums
// and the pilot encoder (reference include/phobicGpu/mphf.hpp:15-98).
// Evaluation: partition = hi32(partitioner * np), bucket from the fulcrum
// table, pilot V, position = offset + ((hash(l1, hash(l2, V/ps)) >> 1) + V) mod ps.
#pragma once
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../encoders/encoders.hpp"
#include "hasher.hpp"
#include "mphf_config.h"

namespace phobicgpu {

// Writes the fields an encoder visits as little-endian u64 words; a vector
// as its length followed by its elements (u64 / i64 one per word, u16 four
// per word, the first in the low bits).
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
  void visit(std::vector<int64_t>& v) {
    put(v.size());
    for (int64_t w : v) put(uint64_t(w));
  }
  void visit(std::vector<uint16_t>& v) {
    put(v.size());
    for (size_t i = 0; i < v.size(); i += 4) {
      uint64_t w = 0;
      for (size_t k = 0; k < 4 && i + k < v.size(); k++) w |= uint64_t(v[i + k]) << (16 * k);
      put(w);
    }
  }
  template <typename T>
  void visit(T& t) {
    t.visit(*this);
  }
};

// Reads what ByteSaver wrote; throws std::runtime_error on truncated input.
struct ByteLoader {
  const uint8_t* p;
  size_t left;
  uint64_t get() {
    if (left < 8) throw std::runtime_error("truncated MPHF bytes");
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= uint64_t(p[i]) << (8 * i);
    p += 8;
    left -= 8;
    return v;
  }
  uint64_t length() {
    const uint64_t n = get();
    if (n > left) throw std::runtime_error("corrupt MPHF bytes");
    return n;
  }
  void visit(uint64_t& v) { v = get(); }
  void visit(std::vector<uint64_t>& v) {
    v.resize(length());
    for (auto& w : v) w = get();
  }
  void visit(std::vector<int64_t>& v) {
    v.resize(length());
    for (auto& w : v) w = int64_t(get());
  }
  void visit(std::vector<uint16_t>& v) {
    v.resize(length());
    for (size_t i = 0; i < v.size(); i += 4) {
      const uint64_t w = get();
      for (size_t k = 0; k < 4 && i + k < v.size(); k++) v[i + k] = uint16_t(w >> (16 * k));
    }
  }
  template <typename T>
  void visit(T& t) {
    t.visit(*this);
  }
};

template <typename PilotEncoder, typename PartitionOffsetEncoder, typename Hasher>
class MPHF {
 private:
  static constexpr uint64_t kMagic = 0x3248504d424f4850ull;  // "PHOBMPH2"

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

  // static as in the reference (mphf.hpp:61); the seed is this build's
  // extension, 0 being the reference's hash
  template <typename keyType>
  static inline Key initialHash(const keyType& keyRaw, uint64_t hashSeed = 0) {
    return Hasher::hash(keyRaw, hashSeed);
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
    const Key key = initialHash(keyRaw, seed);
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

  // Encoding field of the serialised header: the pilot encoder's id in the
  // low byte (interleaved columns: 0 compact, 1 rice; one sequence: 2
  // compact, 3 rice; dual columns: 4 + 2*first + second, so 6 for
  // rice-then-compact), the offset encoder's id in the next byte (differences
  // to i*P: 0 compact, 1 rice; direct: 2 compact, 3 rice).
  static constexpr uint64_t layoutId() { return PilotEncoder::encoding_id | (PartitionOffsetEncoder::encoding_id << 8); }

  // Encoded MPHF bytes (little-endian u64 words): magic "PHOBMPH2", n, P,
  // np, B, the XXH3 seed, the encoding (layoutId()), the load factor in ppm
  // (10^6 = alpha 1), the 2048 fulcrums as u32 pairs, the pilots, the
  // partition offsets, and for alpha < 1 the remap offsets and entries.  The
  // reference has no on-disk format.
  std::vector<uint8_t> serialize() {
    std::vector<uint8_t> out;
    ByteSaver s{out};
    s.put(kMagic);
    s.put(keys);
    s.put(partitionSize);
    s.put(partitions);
    const uint64_t B = fulcs.empty() ? 0 : (fulcs.back() >> 16);
    s.put(B);
    s.put(seed);
    s.put(layoutId());
    s.put(hasRemap ? cfg.alphaPpm() : 1000000u);
    for (size_t i = 0; i + 1 < fulcs.size(); i += 2) s.put(uint64_t(fulcs[i]) | (uint64_t(fulcs[i + 1]) << 32));
    if (partitions) pilots.visit(s);
    partitionOffsets.visit(s);
    if (hasRemap) {
      remapOffsets.visit(s);
      remap.visit(s);
    }
    return out;
  }

  // The inverse of serialize(); throws std::runtime_error on bytes of
  // another format, pilot encoding or length.
  void deserialize(const uint8_t* data, size_t len) {
    ByteLoader l{data, len};
    if (l.get() != kMagic) throw std::runtime_error("not PHOBIC MPHF bytes");
    keys = l.get();
    partitionSize = (uint32_t)l.get();
    partitions = (uint32_t)l.get();
    const uint64_t B = l.get();
    seed = l.get();
    if (l.get() != layoutId()) throw std::runtime_error("MPHF bytes use another pilot or offset encoding");
    const uint64_t a_ppm = l.get();
    fulcs.assign(FULCS_INTER, 0);
    for (size_t i = 0; i < fulcs.size(); i += 2) {
      const uint64_t w = l.get();
      fulcs[i] = uint32_t(w);
      fulcs[i + 1] = uint32_t(w >> 32);
    }
    cfg = MPHFconfig();
    cfg.partitionSize = partitionSize;
    cfg.bucketCountPerPartition = (uint32_t)B;
    cfg.seed = seed;
    cfg.alpha = double(a_ppm) / 1e6;
    pilots = PilotEncoder();
    if (partitions) {
      pilots.resize(B);
      pilots.visit(l);
    }
    partitionOffsets.visit(l);
    hasRemap = a_ppm < 1000000u;
    if (hasRemap) {
      remapOffsets.visit(l);
      remap.visit(l);
    }
    if (l.left) throw std::runtime_error("trailing MPHF bytes");
  }

  void save(const std::string& path) {
    const std::vector<uint8_t> bytes = serialize();
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(bytes.data()), (std::streamsize)bytes.size());
    if (!out) throw std::runtime_error("cannot write " + path);
  }

  void load(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read " + path);
    const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    deserialize(bytes.data(), bytes.size());
  }
};

}  // namespace phobicgpu
