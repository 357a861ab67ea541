// encoders.hpp -- base value encoders (reference include/encoders/base/encoders.hpp).
// `compact` is the one FastQueryMphf uses (encoders.hpp:45-74), `rice` the
// one SmallSpaceMphf uses (encoders.hpp:14-43).
#pragma once
#include <string>

#include "compact_vector.hpp"
#include "rice_sequence.hpp"

namespace phobicgpu {

struct compact {
  static constexpr uint64_t encoding_id = 0;  // in the serialised MPHF header
  template <typename Iterator>
  void encode(Iterator begin, uint64_t n) {
    m_values.build(begin, n);
  }
  static std::string name() { return "compact"; }
  size_t size() const { return m_values.size(); }
  size_t num_bits() const { return m_values.bytes() * 8; }
  uint64_t access(uint64_t i) const { return m_values.access(i); }
  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(m_values);
  }

 private:
  compact_vector m_values;
};

struct rice {
  static constexpr uint64_t encoding_id = 1;
  template <typename Iterator>
  void encode(Iterator begin, uint64_t n) {
    m_values.encode(begin, n);
  }
  static std::string name() { return "rice"; }
  size_t size() const { return m_values.size(); }
  size_t num_bits() const { return m_values.bytes() * 8; }
  uint64_t access(uint64_t i) const { return m_values.access(i); }
  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(m_values);
  }

 private:
  rice_sequence m_values;
};

}  // namespace phobicgpu
