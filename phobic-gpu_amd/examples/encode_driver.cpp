// encode_driver.cpp -- runs the host half of the build (MPHF::setData with
// a chosen pilot and offset encoder, serialisation, loading and queries) on
// pilots / offsets read from a file.  Needs no GPU.
//   encode_driver [--small] [--pilots NAME] [--offsets NAME] IN OUT [QUERY_KEYS QUERY_OUT]
// --pilots: compact (FastQueryMphf, the default), rice (SmallSpaceMphf, also
// --small), mono_compact, mono_rice, dual_rice_compact (the reference
// benchmark's default, split 0.5); --offsets: diff_compact (the default),
// diff_rice, direct_compact, direct_rice.
// IN:  u64 n, u64 np, u64 B, u64 P, f64 lambda, u64 seed,
//      u32 pilots[np*B] (interleaved), u32 offsets[np+1]
// OUT: the serialised MPHF bytes.  QUERY_KEYS: u64 count + u64 keys;
// QUERY_OUT: u32 positions, answered by the MPHF loaded back from OUT's bytes
// (the driver fails if that copy serialises differently or answers
// differently from the built one).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <phobic_gpu_mphf.hpp>

using namespace phobicgpu;

template <typename T>
static void rd(std::ifstream& in, T* p, size_t count) {
  in.read(reinterpret_cast<char*>(p), sizeof(T) * count);
  if (!in) throw std::runtime_error("short read");
}

template <typename Mphf>
static int run(int argc, char** argv) {
  std::ifstream in(argv[1], std::ios::binary);
  uint64_t n, np, B, P, seed;
  double lambda;
  rd(in, &n, 1);
  rd(in, &np, 1);
  rd(in, &B, 1);
  rd(in, &P, 1);
  rd(in, &lambda, 1);
  rd(in, &seed, 1);
  std::vector<uint32_t> pilots(np * B), offsets(np + 1);
  rd(in, pilots.data(), pilots.size());
  rd(in, offsets.data(), offsets.size());
  MPHFconfig cfg(lambda, (uint32_t)P);
  cfg.seed = seed;
  if (cfg.bucketCountPerPartition != B) {
    std::cerr << "bucket count mismatch" << std::endl;
    return 1;
  }
  Mphf f;
  f.setData(pilots, offsets, (uint32_t)np, cfg);
  const std::vector<uint8_t> bytes = f.serialize();
  std::ofstream out(argv[2], std::ios::binary);
  out.write(reinterpret_cast<const char*>(bytes.data()), bytes.size());
  out.close();
  Mphf g;  // the same MPHF loaded back
  g.load(argv[2]);
  if (g.serialize() != bytes) {
    std::cerr << "loaded MPHF serialises differently" << std::endl;
    return 1;
  }
  if (argc == 5) {
    std::ifstream qin(argv[3], std::ios::binary);
    uint64_t count;
    rd(qin, &count, 1);
    std::vector<uint64_t> keys(count);
    rd(qin, keys.data(), count);
    std::vector<uint32_t> pos(count);
    for (uint64_t i = 0; i < count; i++) {
      pos[i] = g(keys[i]);
      if (pos[i] != f(keys[i])) {
        std::cerr << "loaded MPHF answers differently for key " << i << std::endl;
        return 1;
      }
    }
    std::ofstream qout(argv[4], std::ios::binary);
    qout.write(reinterpret_cast<const char*>(pos.data()), pos.size() * 4);
  }
  std::printf("bits_per_key=%.4f bytes=%zu\n", n ? f.getBitsPerKey() : 0.0f, bytes.size());
  return 0;
}

template <typename Pilots>
static int with_offsets(const std::string& offsets, int argc, char** argv) {
  if (offsets == "diff_compact") return run<MPHF<Pilots, diff_partition_encoder<compact>, xxhash>>(argc, argv);
  if (offsets == "diff_rice") return run<MPHF<Pilots, diff_partition_encoder<rice>, xxhash>>(argc, argv);
  if (offsets == "direct_compact") return run<MPHF<Pilots, direct_partition_encoder<compact>, xxhash>>(argc, argv);
  if (offsets == "direct_rice") return run<MPHF<Pilots, direct_partition_encoder<rice>, xxhash>>(argc, argv);
  std::cerr << "unknown offset encoder " << offsets << std::endl;
  return 2;
}

int main(int argc, char** argv) {
  std::string pilots = "compact", offsets = "diff_compact";
  while (argc > 1 && std::string(argv[1]).rfind("--", 0) == 0) {
    const std::string flag = argv[1];
    if (flag == "--small") {
      pilots = "rice";
      argv++;
      argc--;
    } else if ((flag == "--pilots" || flag == "--offsets") && argc > 2) {
      (flag == "--pilots" ? pilots : offsets) = argv[2];
      argv += 2;
      argc -= 2;
    } else {
      std::cerr << "unknown flag " << flag << std::endl;
      return 2;
    }
  }
  if (argc != 3 && argc != 5) {
    std::cerr << "usage: encode_driver [--small] [--pilots NAME] [--offsets NAME] IN OUT [QUERY_KEYS QUERY_OUT]"
              << std::endl;
    return 2;
  }
  if (pilots == "compact") return with_offsets<interleaved_encoder<compact>>(offsets, argc, argv);
  if (pilots == "rice") return with_offsets<interleaved_encoder<rice>>(offsets, argc, argv);
  if (pilots == "mono_compact") return with_offsets<mono_encoder<compact>>(offsets, argc, argv);
  if (pilots == "mono_rice") return with_offsets<mono_encoder<rice>>(offsets, argc, argv);
  if (pilots == "dual_rice_compact") return with_offsets<interleaved_encoder_dual<rice, compact>>(offsets, argc, argv);
  std::cerr << "unknown pilot encoder " << pilots << std::endl;
  return 2;
}
