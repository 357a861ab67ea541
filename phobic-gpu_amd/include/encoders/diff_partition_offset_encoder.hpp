// diff_partition_offset_encoder.hpp -- partition offsets stored as
// differences to the expected offset i*P (reference
// include/encoders/base/linear_diff_encoder.hpp:7-53 and
// partitionOffsetEnocders/diff_partition_offset_encoder.hpp:7-38):
// v_i = offsets[i] - i*P is stored as (|v| << 1) | (v > 0).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace phobicgpu {

template <typename BaseEncoder>
struct linear_diff_encoder {
  template <typename Iterator>
  void encode(Iterator begin, uint64_t size, uint64_t inc) {
    increment = (int64_t)inc;
    std::vector<uint64_t> diff;
    diff.reserve(size);
    int64_t expected = 0;
    for (uint64_t i = 0; i < size; ++i, ++begin) {
      const int64_t v = (int64_t)*begin - expected;
      const uint64_t a = (uint64_t)(v < 0 ? -v : v);
      diff.push_back((a << 1) | uint64_t(v > 0));
      expected += increment;
    }
    enc.encode(diff.begin(), size);
  }
  static std::string name() { return BaseEncoder::name(); }
  size_t size() const { return enc.size(); }
  size_t num_bits() const { return enc.num_bits(); }
  inline uint64_t access(uint64_t i) const {
    const uint64_t value = enc.access(i);
    const int64_t s = (int64_t(value & 1) * 2 - 1) * int64_t(value >> 1);
    return uint64_t(s + int64_t(i) * increment);
  }
  int64_t step() const { return increment; }
  template <typename Visitor>
  void visit(Visitor& visitor) {
    uint64_t inc = (uint64_t)increment;
    visitor.visit(inc);  // a loading visitor overwrites it
    increment = (int64_t)inc;
    visitor.visit(enc);
  }

 private:
  int64_t increment = 0;
  BaseEncoder enc;
};

template <typename BaseEncoder>
struct diff_partition_encoder {
  static constexpr uint64_t encoding_id = BaseEncoder::encoding_id;  // in the serialised MPHF header
  template <typename Iterator>
  void encode(Iterator begin, uint64_t partitionSize, uint64_t partitions) {
    enc.encode(begin, partitions, partitionSize);
  }
  static std::string name() { return "DiffToExpected<" + BaseEncoder::name() + ">"; }
  size_t size() const { return enc.size(); }
  size_t num_bits() const { return enc.num_bits(); }
  inline uint64_t access(uint64_t partition) const { return enc.access(partition); }
  template <typename Visitor>
  void visit(Visitor& visitor) {
    enc.visit(visitor);
  }

 private:
  linear_diff_encoder<BaseEncoder> enc;
};

}  // namespace phobicgpu
