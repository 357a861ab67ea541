// xcd.hpp -- workgroup renumbering for the 8 XCDs of an MI355X.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phobic {

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b mod 8).  xcd_block renumbers them so that each XCD takes a contiguous
// range of [0, n): neighbouring work items then share an L2, which merges
// their writes to neighbouring cache lines instead of sending partial lines
// from several L2s.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
  const uint32_t q = n / kXcds, rem = n % kXcds, x = b % kXcds, i = b / kXcds;
  return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + i;
}

}  // namespace phobic
