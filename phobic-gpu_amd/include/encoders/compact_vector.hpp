// compact_vector.hpp -- fixed-width bit-packed integer vector (same layout as
// the reference's include/encoders/base/compact_vector.hpp:73-336, which the
// encoded MPHF bytes depend on): width w = max == 0 ? 1 : bit length of max,
// values packed LSB-first into 64-bit words, ceil(n*w/64) + 1 words.
#pragma once
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace phobicgpu {

struct compact_vector {
  compact_vector() : m_size(0), m_width(0), m_mask(0) {}

  template <typename Iterator>
  void build(Iterator begin, uint64_t n) {
    uint64_t mx = 0;
    Iterator it = begin;
    for (uint64_t i = 0; i < n; ++i, ++it) mx = std::max<uint64_t>(mx, *it);
    uint64_t w = 0;
    while (w < 64 && (mx >> w) != 0) w++;
    build(begin, n, w ? w : 1);
  }

  template <typename Iterator>
  void build(Iterator begin, uint64_t n, uint64_t w) {
    if (!w) throw std::runtime_error("width must be greater than 0");
    m_size = n;
    m_width = w;
    m_mask = w == 64 ? ~uint64_t(0) : ((uint64_t(1) << w) - 1);
    m_bits.assign((n * w + 63) / 64 + 1, 0);  // one guard word
    for (uint64_t i = 0; i < n; ++i, ++begin) {
      const uint64_t v = uint64_t(*begin) & m_mask, pos = i * w, blk = pos >> 6, sh = pos & 63;
      m_bits[blk] |= v << sh;
      if (sh + w > 64) m_bits[blk + 1] |= v >> (64 - sh);
    }
  }

  // Fixed width w, 0 allowed (then every value reads as 0 and the vector is
  // one guard word): the reference's compact_vector::builder(n, w) path.
  template <typename Iterator>
  void build_width(Iterator begin, uint64_t n, uint64_t w) {
    if (w == 0) {
      m_size = n;
      m_width = 0;
      m_mask = 0;
      m_bits.assign(1, 0);
      return;
    }
    build(begin, n, w);
  }

  inline uint64_t access(uint64_t i) const {
    if (!m_width) return 0;
    const uint64_t pos = i * m_width, blk = pos >> 6, sh = pos & 63;
    uint64_t v = m_bits[blk] >> sh;
    if (sh + m_width > 64) v |= m_bits[blk + 1] << (64 - sh);
    return v & m_mask;
  }
  inline uint64_t operator[](uint64_t i) const { return access(i); }

  uint64_t size() const { return m_size; }
  uint64_t width() const { return m_width; }
  const std::vector<uint64_t>& bits() const { return m_bits; }
  // size, width, mask + the vector (length word + words)
  size_t bytes() const { return 3 * sizeof(uint64_t) + sizeof(uint64_t) + m_bits.size() * sizeof(uint64_t); }

  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(m_size);
    visitor.visit(m_width);
    visitor.visit(m_mask);
    visitor.visit(m_bits);
  }

 private:
  uint64_t m_size, m_width, m_mask;
  std::vector<uint64_t> m_bits;
};

}  // namespace phobicgpu
