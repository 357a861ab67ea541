// Prints the reference MPHFconfig for (lambda, P) given on the command line
// (default constructor, or MPHFconfig(Bucketer*, lambda, P) with the
// reference's UniformBucketer / SkewBucketer when a third argument names it):
//   line 1: bucketCountPerPartition partitionMaxSize
//   line 2: the 2048 fulcrums
// Compiled against the reference's include/phobicGpu/mphf_config.h by
// oracle/Makefile (target "ref"); the output binary lives in oracle/_ref/.
#include <cstdio>
#include <cstdlib>
#include <string>
#include "phobicGpu/mphf_config.h"

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: ref_config LAMBDA P [uniform|skew]\n");
    return 2;
  }
  const double lambda = std::atof(argv[1]);
  const uint32_t P = (uint32_t)std::strtoul(argv[2], nullptr, 10);
  phobicgpu::UniformBucketer uniform;
  phobicgpu::SkewBucketer skew;
  phobicgpu::Bucketer* custom = nullptr;
  if (argc == 4) custom = std::string(argv[3]) == "skew" ? (phobicgpu::Bucketer*)&skew : (phobicgpu::Bucketer*)&uniform;
  phobicgpu::MPHFconfig cfg = custom ? phobicgpu::MPHFconfig(custom, lambda, P) : phobicgpu::MPHFconfig(lambda, P);
  std::printf("%u %u\n", cfg.bucketCountPerPartition, cfg.partitionMaxSize());
  const auto& f = cfg.getFulcs();
  for (size_t i = 0; i < f.size(); i++) std::printf("%u%c", f[i], i + 1 == f.size() ? '\n' : ' ');
  return 0;
}
