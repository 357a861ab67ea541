// key_types.cpp -- the drop-in API over every key type the reference's
// hasher takes (include/phobicGpu/hasher.hpp:28-66): uint64_t,
// std::pair<uint64_t, uint64_t> and Key through xxhash, and pre-hashed Key
// values through nohash.  Each MPHF must be a bijection onto [0, n) (the
// validity contract of the reference's example.cpp:43-59).
// Usage: key_types [n]
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include <phobic_gpu_mphf.hpp>

using namespace phobicgpu;

// splitmix64 (distinct outputs for distinct counters)
static uint64_t mix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename Mphf, typename K>
static bool check(const char* name, const std::vector<K>& keys) {
  MPHFbuilder builder(MPHFconfig(7.0, 2048));
  Mphf f;
  builder.build(keys, f);
  std::vector<bool> taken(keys.size(), false);
  for (size_t i = 0; i < keys.size(); i++) {
    const size_t h = f(keys[i]);
    if (h >= keys.size() || taken[h]) {
      std::cerr << name << ": key " << i << (h >= keys.size() ? " out of range" : " collides") << std::endl;
      return false;
    }
    taken[h] = true;
  }
  std::cout << name << " valid for " << keys.size() << std::endl;
  return true;
}

int main(int argc, char** argv) {
  const size_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 200000;
  uint64_t s = 5;
  std::vector<uint64_t> u(n);
  std::vector<std::pair<uint64_t, uint64_t>> pr(n);
  std::vector<Key> ks(n);
  for (size_t i = 0; i < n; i++) {
    u[i] = mix(s);
    pr[i] = {mix(s), mix(s)};
    const uint64_t a = mix(s), b = mix(s);  // distinct 128-bit values
    ks[i] = Key((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
  }
  using NoHashMphf = MPHF<interleaved_encoder<compact>, diff_partition_encoder<compact>, nohash>;
  // the reference's static initial hash (mphf.hpp:61): XXH3-128 of the 8
  // bytes of 1 with seed 0 is 0xbdc94bce2eda264d_bc08dc21994df8a2
  const Key h = FastQueryMphf::initialHash(uint64_t(1));
  bool ok = h.partitioner == 0x994df8a2u && h.bucketer == 0xbc08dc21u && h.lower1 == 0x2eda264du &&
            h.lower2 == 0xbdc94bceu;
  std::cout << "initialHash " << (ok ? "matches" : "differs from") << " the known answer" << std::endl;
  ok = check<FastQueryMphf>("uint64_t", u) && ok;
  ok = check<FastQueryMphf>("pair<uint64_t,uint64_t>", pr) && ok;
  ok = check<FastQueryMphf>("Key (xxhash)", ks) && ok;
  ok = check<NoHashMphf>("Key (nohash)", ks) && ok;
  return ok ? 0 : 1;
}
