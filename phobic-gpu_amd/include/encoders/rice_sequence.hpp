// rice_sequence.hpp -- Rice (Golomb power-of-two) coded integer sequence,
// the base encoder of SmallSpaceMphf's pilot columns (reference
// include/encoders/base/rice_sequence.hpp:12-76).  Value v is split into its
// low l bits, kept in a compact_vector of width l, and v >> l, kept in unary
// (that many zeros, then a one) in a bit vector with a select1 index; value i
// is (select(i) - select(i-1) - 1) << l | low[i].
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "bit_vector.hpp"
#include "compact_vector.hpp"

namespace phobicgpu {

struct rice_sequence {
  // The Rice parameter for a geometric source with the values' mean
  // (Kiely's rule: l = ceil(log2(log(phi - 1) / log(1 - p))), p = 1/(mean+1)),
  // evaluated in double precision as the reference does; 0 when every value
  // is 0.
  template <typename Iterator>
  static uint64_t optimal_l(Iterator begin, uint64_t n) {
    uint64_t sum = 0;
    for (uint64_t i = 0; i < n; ++i, ++begin) sum += uint64_t(*begin);
    if (sum == 0) return 0;
    const double p = double(n) / (double(sum) + double(n));
    const double gold = std::log2(2.0 / (std::sqrt(5.0) + 1.0));
    const double l = std::ceil(std::log2(gold / std::log2(1.0 - p)));
    return l > 0 ? uint64_t(l) : 0;
  }

  template <typename Iterator>
  void encode(Iterator begin, uint64_t n) {
    if (n == 0) return;
    const uint64_t l = optimal_l(begin, n);
    const uint64_t mask = l >= 64 ? ~uint64_t(0) : (uint64_t(1) << l) - 1;
    std::vector<uint64_t> low(n);
    Iterator it = begin;
    for (uint64_t i = 0; i < n; ++i, ++it) {
      const uint64_t v = uint64_t(*it);
      low[i] = v & mask;
      m_high.push_unary(l >= 64 ? 0 : v >> l);
    }
    m_low.build_width(low.begin(), n, l);
    m_index.build(m_high);
  }

  uint64_t access(uint64_t i) const {
    const int64_t start = i ? int64_t(m_index.select(m_high, i - 1)) : -1;
    const int64_t end = int64_t(m_index.select(m_high, i));
    return (uint64_t(end - start - 1) << m_low.width()) | m_low.access(i);
  }

  uint64_t size() const { return m_low.size(); }
  uint64_t bytes() const { return m_high.bytes() + m_index.bytes() + m_low.bytes(); }

  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(m_high);
    visitor.visit(m_index);
    visitor.visit(m_low);
  }

 private:
  bit_vector m_high;
  darray1 m_index;
  compact_vector m_low;
};

}  // namespace phobicgpu
