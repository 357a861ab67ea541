// interleaved_encoder.hpp -- one base encoder per bucket column: column j
// holds the pilots of bucket j of every partition, i.e. pilots[j*np .. j*np+np)
// of the interleaved build output (reference
// include/encoders/pilotEncoders/interleaved_encoder.hpp:5-43).
#pragma once
#include <string>
#include <vector>

namespace phobicgpu {

template <typename BaseEncoder>
struct interleaved_encoder {
  static constexpr uint64_t encoding_id = BaseEncoder::encoding_id;
  template <typename Iterator>
  void encode(Iterator begin, uint64_t partitions, uint64_t buckets) {
    encoders.resize(buckets);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t j = 0; j < (int64_t)buckets; j++) encoders[j].encode(begin + j * partitions, partitions);
  }
  inline uint64_t access(uint64_t partition, uint64_t bucket) const { return encoders[bucket].access(partition); }
  std::string name() { return "InterEncoder<" + BaseEncoder::name() + ">"; }
  uint64_t num_bits() const {
    uint64_t s = 0;
    for (auto& e : encoders) s += e.num_bits();
    return s;
  }
  // one encoder per bucket column (before a loading visitor fills them)
  void resize(uint64_t buckets) { encoders.resize(buckets); }
  template <typename Visitor>
  void visit(Visitor& visitor) {
    for (auto& e : encoders) e.visit(visitor);
  }

 private:
  std::vector<BaseEncoder> encoders;
};

}  // namespace phobicgpu
