// interleaved_encoder_dual.hpp -- the first floor(B * tradeoff) bucket
// columns with one base encoder, the rest with another (reference
// include/encoders/pilotEncoders/interleaved_encoder_dual.hpp:7-55).  The
// reference benchmark's default pilot encoder is
// interleaved_encoder_dual<rice, compact> with tradeoff 0.5: the first
// columns hold the large buckets, placed first, whose pilots are small and
// Rice-code well; the late columns hold small buckets with large pilots.
#pragma once
#include <stdexcept>
#include <string>

#include "interleaved_encoder.hpp"

namespace phobicgpu {

template <typename BaseEncoder1, typename BaseEncoder2>
struct interleaved_encoder_dual {
  static constexpr uint64_t encoding_id = 4 + 2 * BaseEncoder1::encoding_id + BaseEncoder2::encoding_id;

  template <typename Iterator>
  void encode(Iterator begin, uint64_t partitions, uint64_t buckets) {
    initialized = true;
    total = buckets;
    // the reference multiplies in float and truncates
    buckets1 = uint64_t(float(buckets) * tradeoff);
    encoder1.encode(begin, partitions, buckets1);
    encoder2.encode(begin + buckets1 * partitions, partitions, buckets - buckets1);
  }

  // share of bucket columns that use BaseEncoder1; set before encode()
  void setEncoderTradeoff(float enc_tradeoff) {
    if (!(enc_tradeoff >= 0.0f && enc_tradeoff <= 1.0f)) throw std::invalid_argument("tradeoff outside [0, 1]");
    if (initialized) throw std::logic_error("tradeoff set after encode");
    tradeoff = enc_tradeoff;
  }

  inline uint64_t access(uint64_t partition, uint64_t bucket) const {
    return bucket < buckets1 ? encoder1.access(partition, bucket) : encoder2.access(partition, bucket - buckets1);
  }

  std::string name() {
    return "InterEncoderDual<" + BaseEncoder1::name() + "," + BaseEncoder2::name() + "," + std::to_string(tradeoff) +
           ">";
  }

  uint64_t num_bits() const { return encoder1.num_bits() + encoder2.num_bits(); }

  // the bucket count, before a loading visitor fills the columns
  void resize(uint64_t buckets) { total = buckets; }

  // the split point travels with the columns, so a loaded copy needs no
  // tradeoff
  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(buckets1);
    if (buckets1 > total) throw std::runtime_error("dual encoder split beyond the bucket count");
    encoder1.resize(buckets1);
    encoder2.resize(total - buckets1);
    encoder1.visit(visitor);
    encoder2.visit(visitor);
  }

 private:
  bool initialized = false;
  uint64_t buckets1 = 0, total = 0;
  float tradeoff = 0.5f;
  interleaved_encoder<BaseEncoder1> encoder1;
  interleaved_encoder<BaseEncoder2> encoder2;
};

}  // namespace phobicgpu
