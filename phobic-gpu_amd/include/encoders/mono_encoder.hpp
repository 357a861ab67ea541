// mono_encoder.hpp -- all pilots in one base-encoded sequence, in the
// interleaved build order p + b*np (reference
// include/encoders/pilotEncoders/mono_encoders.hpp:7-41).
#pragma once
#include <string>

namespace phobicgpu {

template <typename BaseEncoder>
struct mono_encoder {
  static constexpr uint64_t encoding_id = 2 + BaseEncoder::encoding_id;

  template <typename Iterator>
  void encode(Iterator begin, uint64_t parts, uint64_t buckets) {
    partitions = parts;
    enc.encode(begin, parts * buckets);
  }
  std::string name() { return BaseEncoder::name(); }
  size_t size() const { return enc.size(); }
  size_t num_bits() const { return enc.num_bits(); }
  inline uint64_t access(uint64_t partition, uint64_t bucket) const { return enc.access(partitions * bucket + partition); }
  void resize(uint64_t) {}
  // the partition count travels with the sequence (the reference's visit
  // has only the sequence)
  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(partitions);
    visitor.visit(enc);
  }

 private:
  uint64_t partitions = 0;
  BaseEncoder enc;
};

}  // namespace phobicgpu
