// search_layout.hpp -- LDS layout of one partition in the search kernel
// (search_kernel.hip), shared by the kernel, its launcher and a host-side
// test that checks it for every table size and bucket count the library
// accepts.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define PHOBIC_HD __host__ __device__
#else
#define PHOBIC_HD
#endif

namespace phobic {

PHOBIC_HD inline uint32_t layout_max(uint32_t a, uint32_t b) { return a > b ? a : b; }

// Sub-windows per lane of a scan pass: lane l owns offsets
// o0 + 2048 h + [32 l, 32 l + 32) for h < subwindows(tm).
PHOBIC_HD inline uint32_t subwindows(uint32_t tm) { return tm <= 2048 ? 1 : tm <= 4096 ? 2 : 4; }

// E words: marks reach bit 2*tm - 1; a scan pass reads up to word
// (2*tm - 2) / 32 + 64 * subwindows(tm).
PHOBIC_HD inline uint32_t e_words(uint32_t tcap) { return (2 * tcap) / 32 + 64 * subwindows(tcap) + 8; }

// LDS of one partition, in 32-bit words:
//   rt[B]   processing rank r -> bucket | size << 16 | lane-group log2 << 25
//           (a bucket's start in the slice is the sum of the sizes before
//           it: the keys placed so far)
//   cp[B]   steps 1-3: keys per bucket (LDS atomics); steps 5-6: pilot per bucket
//   U       steps 1-2: hist[256], bucket count per t = 256 - size
//           steps 3-4: bpos[B], next free index of each bucket in the slice
//           steps 5-6: E[e_words], coll[tcap / 32 + 1] (collision check)
// Word offsets of the regions inside one partition's LDS, used both by the
// kernel's pointers and by the size the launch reserves.
struct SearchLayout {
  uint32_t rt, cp, U, coll, words;
};
PHOBIC_HD inline SearchLayout search_layout(uint32_t tcap, uint32_t B) {
  SearchLayout l;
  l.rt = 0;
  l.cp = l.rt + B;
  l.U = l.cp + B;
  l.coll = l.U + e_words(tcap);
  // the union holds E + coll, or hist, or bpos
  l.words = l.U + layout_max(e_words(tcap) + (tcap + 31) / 32 + 1, layout_max(256u, B));
  return l;
}
PHOBIC_HD inline uint32_t lds_words(uint32_t tcap, uint32_t B) { return search_layout(tcap, B).words; }

// The kernel forms its region pointers as rt, rt + B, rt + 2B, rt + 2B +
// e_words(tcap) (this form compiles to the faster code); this checks that
// they are search_layout's.
inline bool layout_matches_kernel(uint32_t tcap, uint32_t B) {
  const SearchLayout l = search_layout(tcap, B);
  return l.rt == 0 && l.cp == B && l.U == 2 * B && l.coll == 2 * B + e_words(tcap) &&
         l.words >= l.coll + (tcap + 31) / 32 + 1 && l.words >= l.U + layout_max(256u, B);
}

}  // namespace phobic
