// direct_partition_offset_encoder.hpp -- partition offsets stored as they
// are, one base-encoded sequence of np+1 values (reference
// include/encoders/partitionOffsetEnocders/direct_partition_offset_encoder.hpp:7-38).
#pragma once
#include <cstdint>
#include <string>

namespace phobicgpu {

template <typename BaseEncoder>
struct direct_partition_encoder {
  static constexpr uint64_t encoding_id = 2 + BaseEncoder::encoding_id;
  template <typename Iterator>
  void encode(Iterator begin, uint64_t /*partitionSize*/, uint64_t partitions) {
    enc.encode(begin, partitions);
  }
  static std::string name() { return "Direct<" + BaseEncoder::name() + ">"; }
  size_t size() const { return enc.size(); }
  size_t num_bits() const { return enc.num_bits(); }
  inline uint64_t access(uint64_t partition) const { return enc.access(partition); }
  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(enc);
  }

 private:
  BaseEncoder enc;
};

}  // namespace phobicgpu
