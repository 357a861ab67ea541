// mphf_config.h -- MPHFconfig, the build parameters (reference:
// include/phobicGpu/mphf_config.h:16-65), plus the XXH3 seed extension.
#pragma once
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

#include "bucketers.h"

namespace phobicgpu {

constexpr uint32_t FULCS_INTER = 2048;  // shaders/constants.glsl:1

struct MPHFconfig {
 private:
  std::vector<uint32_t> fulcrums;
  bool customBucketer = false;

 public:
  uint32_t partitionSize;
  uint32_t sortingBins;
  uint32_t bucketCountPerPartition;
  double m_averageBucketSize;
  uint64_t seed = 0;   // extension: XXH3 seed (the reference hashes with seed 0)
  double alpha = 1.0;  // extension: load factor in [0.5, 1]; 1 = the reference

  // table size of a partition with ps keys: ceil(ps / alpha) in ppm arithmetic
  uint32_t tableSize(uint32_t ps) const {
    const uint32_t a = alphaPpm();
    return a >= 1000000u ? ps : (uint32_t)(((uint64_t)ps * 1000000u + a - 1) / a);
  }
  uint32_t alphaPpm() const { return (uint32_t)std::llround((alpha == 0.0 ? 1.0 : alpha) * 1e6); }

  // B = ceil(P / lambda) with the OptBucketer table (mphf_config.h:26-32)
  MPHFconfig(double averageBucketSize = 8.0, uint32_t partitionSize = 2048)
      : partitionSize(partitionSize),
        sortingBins(256),
        bucketCountPerPartition(uint32_t(std::ceil(partitionSize / averageBucketSize))),
        m_averageBucketSize(averageBucketSize) {
    OptBucketer b;
    setBucketer(&b);
  }

  // custom bucket function, B = round(P / lambda) (mphf_config.h:34-42)
  MPHFconfig(Bucketer* bucketer, double averageBucketSize = 8.0, uint32_t partitionSize = 2048)
      : partitionSize(partitionSize),
        sortingBins(256),
        bucketCountPerPartition(uint32_t(std::round(partitionSize / averageBucketSize))),
        m_averageBucketSize(averageBucketSize) {
    setBucketer(bucketer);
    customBucketer = true;
  }

  // fulcrum table (mphf_config.h:44-55)
  void setBucketer(Bucketer* bucketer) {
    bucketer->init(m_averageBucketSize, partitionSize);
    fulcrums.clear();
    fulcrums.push_back(0);
    for (uint32_t xi = 1; xi < FULCS_INTER - 1; xi++) {
      const double x = double(xi) / double(FULCS_INTER - 1);
      fulcrums.push_back(uint32_t(bucketer->getBucketRel(x) * double(bucketCountPerPartition << 16)));
    }
    fulcrums.push_back(bucketCountPerPartition << 16);
  }

  const std::vector<uint32_t>& getFulcs() const { return fulcrums; }
  bool hasCustomBucketer() const { return customBucketer; }

  uint32_t partitionMaxSize() const { return partitionSize + partitionSize / 2; }
};

}  // namespace phobicgpu
