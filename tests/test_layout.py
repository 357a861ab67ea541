"""The search kernel's LDS layout (csrc/search_layout.hpp), checked on the
host for every table size up to 20,000 slots (and sampled above) and every
bucket count whose partition fits the 64 KiB a search workgroup may use,
which is what the library's configuration check accepts: the kernel's
pointer formation matches the layout, the union region holds the slot
bitmap with its collision words, the size histogram and the bucket cursors,
and every word a scan pass reads lies inside the bitmap.  The header is
compiled with g++ into a small checker."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "phobic-gpu_amd", "csrc")

CHECKER = r"""
#include <cstdio>
#include "search_layout.hpp"
using namespace phobic;
int main() {
  unsigned long long checked = 0;
  for (uint32_t tcap = 1; tcap <= 131072; tcap += tcap < 20000 ? 1 : 97) {
    for (uint32_t B = 1; B <= 65535; B++) {
      const uint32_t words = lds_words(tcap, B);
      if (words * 4ull > 64 * 1024) break;  // larger B only grows the layout
      const SearchLayout l = search_layout(tcap, B);
      if (!layout_matches_kernel(tcap, B)) { std::printf("mismatch tcap=%u B=%u\n", tcap, B); return 1; }
      // marks reach bit 2 * tcap + 128; a scan pass (H sub-windows, or one
      // sub-window per pass in the narrow mode) reads up to word
      // (2 * tcap - 2) / 32 + 64 * H
      const uint32_t ew = l.coll - l.U;
      if (ew * 32u <= 2 * tcap + 128) { std::printf("marks tcap=%u\n", tcap); return 1; }
      if ((2 * tcap - 2) / 32 + 64 * subwindows(tcap) >= ew) { std::printf("scan tcap=%u\n", tcap); return 1; }
      if (l.words < l.U + 256 || l.words < l.U + B || l.words < l.coll + (tcap + 31) / 32 + 1) {
        std::printf("union tcap=%u B=%u\n", tcap, B);
        return 1;
      }
      checked++;
    }
  }
  std::printf("checked %llu\n", checked);
  return 0;
}
"""


def test_search_layout_for_every_accepted_size():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "check.cpp"), os.path.join(d, "check")
        open(src, "w").write(CHECKER)
        subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", "-o", exe, src], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert int(r.stdout.split()[-1]) > 10_000_000
