// Prints the reference MPHFconfig for (lambda, P) given on the command line:
//   line 1: bucketCountPerPartition partitionMaxSize
//   line 2: the 2048 fulcrums
// Compiled against the reference's include/phobicGpu/mphf_config.h by
// oracle/Makefile (target "ref"); the output binary lives in oracle/_ref/.
#include <cstdio>
#include <cstdlib>
#include "phobicGpu/mphf_config.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: ref_config LAMBDA P\n");
    return 2;
  }
  phobicgpu::MPHFconfig cfg(std::atof(argv[1]), (uint32_t)std::strtoul(argv[2], nullptr, 10));
  std::printf("%u %u\n", cfg.bucketCountPerPartition, cfg.partitionMaxSize());
  const auto& f = cfg.getFulcs();
  for (size_t i = 0; i < f.size(); i++) std::printf("%u%c", f[i], i + 1 == f.size() ? '\n' : ' ');
  return 0;
}
