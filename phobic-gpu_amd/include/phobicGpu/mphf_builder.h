// mphf_builder.h -- MPHFbuilder::build (reference
// include/phobicGpu/mphf_builder.h:18-321).  The whole construction runs on the
// MI355X through the C ABI of phobic_hip.h (hash, partitioning, bucket sort,
// pilot search); the host stages the keys, reads back pilots and offsets and
// hands them to MPHF::setData, as BuildInvocation::run does.
#pragma once
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "host_timer.h"
#include "mphf.hpp"
#include "mphf_config.h"
#include "phobic_hip.h"

namespace phobicgpu {

class MPHFbuilder {
 private:
  MPHFconfig config;
  phobic_ctx* ctx = nullptr;
  int device;

  static void check(int st, const char* what) {
    if (st != PHOBIC_OK) throw std::runtime_error(std::string(what) + ": " + phobic_status_string(st));
  }

  phobic_config abiConfig() const {
    phobic_config c{};
    c.lambda = config.m_averageBucketSize;
    c.partition_size = config.partitionSize;
    c.seed = config.seed;
    c.alpha = config.alpha;
    if (config.hasCustomBucketer()) {
      c.buckets = config.bucketCountPerPartition;
      c.fulcrums = config.getFulcs().data();
    }
    return c;
  }

 public:
  // implicit like the reference's (mphf_builder.h:36)
  MPHFbuilder(MPHFconfig config = MPHFconfig(), int device = 0) : config(config), device(device) {}
  // copies share the configuration; each opens its own device context on
  // its first build
  MPHFbuilder(const MPHFbuilder& o) : config(o.config), device(o.device) {}
  MPHFbuilder& operator=(const MPHFbuilder& o) {
    if (this != &o) {
      if (ctx) phobic_ctx_destroy(ctx);
      ctx = nullptr;
      config = o.config;
      device = o.device;
    }
    return *this;
  }
  ~MPHFbuilder() {
    if (ctx) phobic_ctx_destroy(ctx);
  }

  const MPHFconfig& getConfig() const { return config; }

  template <typename Mphf, typename keyType>
  HostTimer build(const std::vector<keyType>& keys, Mphf& f) {
    HostTimer timer;
    timer.addLabel("initial_hash");  // on the device, inside the build (below)
    if (!ctx) check(phobic_ctx_create(device, &ctx), "phobic_ctx_create");
    timer.addLabel("allocation");
    const phobic_config cfg = abiConfig();
    const uint64_t n = keys.size();
    uint32_t np = 0, B = 0;
    check(phobic_query_sizes(&cfg, n, &np, &B), "phobic_query_sizes");
    std::vector<uint32_t> pilots((size_t)np * B), offsets((size_t)np + 1);
    uint64_t rcap = 0;
    check(phobic_remap_capacity(&cfg, n, &rcap), "phobic_remap_capacity");
    std::vector<uint32_t> remap(rcap ? rcap : 1), remapOffsets(rcap ? (size_t)np + 1 : 1);
    phobic_outputs out{pilots.data(), offsets.data(), rcap ? remap.data() : nullptr,
                       rcap ? remapOffsets.data() : nullptr};
    phobic_stage_times t{};
    int st;
    if constexpr (Mphf::noHash()) {
      static_assert(std::is_same_v<keyType, Key>, "nohash MPHFs are built from Key values");
      st = phobic_build_keys(ctx, reinterpret_cast<const phobic_key*>(keys.data()), n, &cfg, &out, &t);
    } else if constexpr (std::is_same_v<keyType, uint64_t>) {
      st = phobic_build_u64(ctx, keys.data(), n, &cfg, &out, &t);
    } else if constexpr (std::is_same_v<keyType, std::string>) {
      std::vector<uint64_t> offs(n + 1, 0);
      for (uint64_t i = 0; i < n; i++) offs[i + 1] = offs[i] + keys[i].size();
      std::vector<uint8_t> bytes(offs[n]);
      for (uint64_t i = 0; i < n; i++)
        if (!keys[i].empty()) std::memcpy(bytes.data() + offs[i], keys[i].data(), keys[i].size());
      st = phobic_build_bytes(ctx, bytes.data(), offs.data(), n, &cfg, &out, &t);
    } else if constexpr (std::is_same_v<keyType, std::pair<uint64_t, uint64_t>> || std::is_same_v<keyType, Key>) {
      // 16 bytes per key, hashed as the reference hashes a pair or a Key
      // (XXH128 of the value's bytes, hasher.hpp:51-66)
      static_assert(sizeof(Key) == 16, "Key is four 32-bit words");
      std::vector<uint64_t> offs(n + 1), words(2 * n);
      for (uint64_t i = 0; i < n; i++) {
        offs[i] = 16 * i;
        if constexpr (std::is_same_v<keyType, Key>) {
          std::memcpy(&words[2 * i], &keys[i], 16);
        } else {
          words[2 * i] = keys[i].first;
          words[2 * i + 1] = keys[i].second;
        }
      }
      offs[n] = 16 * n;
      st = phobic_build_bytes(ctx, reinterpret_cast<const uint8_t*>(words.data()), offs.data(), n, &cfg, &out, &t);
    } else {
      static_assert(sizeof(keyType) == 0, "supported key types: uint64_t, std::string, std::pair<u64,u64>, Key");
    }
    check(st, "MPHFbuilder::build");
    // The reference's labels (mphf_builder.h:138-146, 229-290), each the
    // end of its split: keys are hashed on the device (initial_hash is
    // empty: it marks the start), kernels are enqueued inside the build call (setup_commands is
    // empty), bucket sizes are counted by the routing pass that also hashes
    // and maps every key, bucket sorting and the in-partition
    // redistribution are fused into the search, partition offsets need no
    // separate application, and results are copied, not mapped.
    auto stage = [&](const char* name) {
      for (uint32_t i = 0; i < t.count; i++)
        if (std::strcmp(t.names[i], name) == 0) return (double)t.ms[i] * 1e6;
      return 0.0;
    };
    // input_transfer: host-side packing of the keys and their copy in
    double acc = timer.elapsed() - (t.total_ms + t.result_transfer_ms) * 1e6;
    timer.addLabelManually("input_transfer", acc);
    timer.addLabelManually("setup_commands", acc);
    acc += stage("hash_route");
    timer.addLabelManually("GPU_bucket_sizes", acc);
    timer.addLabelManually("GPU_bucket_sorting", acc);
    acc += stage("partition_offsets");
    timer.addLabelManually("GPU_partition_offsets", acc);
    timer.addLabelManually("GPU_apply_partition_offsets", acc);
    acc += stage("partition_split");
    timer.addLabelManually("GPU_key_redistribution", acc);
    acc += stage("partition_search");
    timer.addLabelManually("GPU_search", acc);
    timer.addLabelManually("GPU_memory_map", acc);
    timer.addLabel("result_transfer");
    if (rcap) {
      remap.resize(remapOffsets[np]);
      f.setRemap(remap, remapOffsets);
    }
    f.setData(pilots, offsets, np, config);
    timer.addLabel("encoding");
    return timer;
  }
};

}  // namespace phobicgpu
