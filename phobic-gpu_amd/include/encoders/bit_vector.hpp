// bit_vector.hpp -- a growable bit vector and a select1 index over it, the
// parts of the Rice-coded pilot sequence (reference
// include/encoders/base/bit_vector.hpp and darray.hpp).  The index keeps the
// reference's inventory layout because the bits-per-key figure counts it:
// the position of every 4096th one (a block), the offset of every 64th one
// from its block's first one (a sub-block, 16 bits), and, for a block whose
// ones span 2^16 bits or more, all of its positions.
#pragma once
#include <cstdint>
#include <vector>

namespace phobicgpu {

struct bit_vector {
  void push_back(bool b) {
    if ((m_size & 63) == 0) m_bits.push_back(0);
    m_bits.back() |= uint64_t(b) << (m_size & 63);
    m_size++;
  }
  // n zeros then a one
  void push_unary(uint64_t n) {
    for (uint64_t i = 0; i < n; i++) push_back(false);
    push_back(true);
  }
  uint64_t size() const { return m_size; }
  const std::vector<uint64_t>& words() const { return m_bits; }
  // size + the word vector (its length and words)
  uint64_t bytes() const { return 8 + 8 + 8 * m_bits.size(); }

  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(m_size);
    visitor.visit(m_bits);
  }

 private:
  uint64_t m_size = 0;
  std::vector<uint64_t> m_bits;
};

// Position of the k-th (0-based) set bit of x; x has more than k set bits.
inline uint64_t select_in_word(uint64_t x, uint64_t k) {
  for (uint64_t i = 0; i < k; i++) x &= x - 1;
  return (uint64_t)__builtin_ctzll(x);
}

struct darray1 {
  static constexpr uint64_t kBlock = 4096, kSub = 64, kDense = uint64_t(1) << 16;

  void build(const bit_vector& bv) {
    m_ones = 0;
    m_blocks.clear();
    m_subs.clear();
    m_sparse.clear();
    std::vector<uint64_t> cur;
    cur.reserve(kBlock);
    const std::vector<uint64_t>& w = bv.words();
    for (uint64_t i = 0; i < w.size(); i++) {
      for (uint64_t x = w[i]; x; x &= x - 1) {
        const uint64_t pos = 64 * i + (uint64_t)__builtin_ctzll(x);
        if (pos >= bv.size()) break;
        cur.push_back(pos);
        m_ones++;
        if (cur.size() == kBlock) flush(cur);
      }
    }
    if (!cur.empty()) flush(cur);
  }

  // position of the one with rank idx (idx < number of ones)
  uint64_t select(const bit_vector& bv, uint64_t idx) const {
    const int64_t b = m_blocks[idx / kBlock];
    if (b < 0) return m_sparse[uint64_t(-b - 1) + idx % kBlock];
    const uint64_t start = uint64_t(b) + m_subs[idx / kSub];
    uint64_t rest = idx % kSub;
    if (!rest) return start;
    const std::vector<uint64_t>& w = bv.words();
    uint64_t wi = start >> 6;
    uint64_t word = w[wi] & (~uint64_t(0) << (start & 63));
    for (;;) {
      const uint64_t c = (uint64_t)__builtin_popcountll(word);
      if (rest < c) break;
      rest -= c;
      word = w[++wi];
    }
    return 64 * wi + select_in_word(word, rest);
  }

  uint64_t ones() const { return m_ones; }
  // the count + three vectors (each its length and elements)
  uint64_t bytes() const { return 8 + (8 + 8 * m_blocks.size()) + (8 + 2 * m_subs.size()) + (8 + 8 * m_sparse.size()); }

  template <typename Visitor>
  void visit(Visitor& visitor) {
    visitor.visit(m_ones);
    visitor.visit(m_blocks);
    visitor.visit(m_subs);
    visitor.visit(m_sparse);
  }

 private:
  void flush(std::vector<uint64_t>& cur) {
    if (cur.back() - cur.front() < kDense) {
      m_blocks.push_back(int64_t(cur.front()));
      for (uint64_t i = 0; i < cur.size(); i += kSub) m_subs.push_back(uint16_t(cur[i] - cur.front()));
    } else {  // sparse block: every position, sub-blocks unused
      m_blocks.push_back(-int64_t(m_sparse.size()) - 1);
      m_sparse.insert(m_sparse.end(), cur.begin(), cur.end());
      for (uint64_t i = 0; i < cur.size(); i += kSub) m_subs.push_back(uint16_t(0xFFFF));
    }
    cur.clear();
  }

  uint64_t m_ones = 0;
  std::vector<int64_t> m_blocks;
  std::vector<uint16_t> m_subs;
  std::vector<uint64_t> m_sparse;
};

}  // namespace phobicgpu
