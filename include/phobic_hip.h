/*
 * phobic_hip.h -- C ABI of the MI355X (gfx950) PHOBIC construction path.
 *
 * This is the boundary under the C++ API of include/phobic_gpu_mphf.hpp
 * (reference: stefanfred/PHOBIC-GPU).  In the reference, MPHFbuilder::build
 * (include/phobicGpu/mphf_builder.h:315-321) runs BuildInvocation::run
 * (mphf_builder.h:225-292), which hashes on the host, uploads the 16-byte
 * Keys, records the Vulkan stage sequence of six *_stage.cpp wrappers
 * (src/bucket_sizes_stage.cpp:27-45, src/bucket_sort_stage.cpp:5-47,
 * src/prefix_sum_stage.cpp:42-86, src/apply_partition_offset.cpp:29-38,
 * src/redisistribute_keys_stage.cpp:28-45, src/search_stage.cpp:6-59) and
 * reads back pilots u32[np*B] and partition offsets u32[np+1]
 * (mphf_builder.h:254-260) for MPHF::setData (include/phobicGpu/mphf.hpp:65-75).
 * Those stage wrappers have Vulkan-typed signatures and no C ABI; the
 * functions below are what a foreign binding of that path binds instead.
 *
 * All functions return a phobic_status.  Plain pointers and sizes only.
 * Output layouts are the reference's:
 *   pilots[p + b*np]  for partition p < np and bucket b < B (search.comp:168-174)
 *   offsets[0..np]    exclusive partition offsets, offsets[np] = n
 * Empty buckets get pilot 0 (the reference leaves them uninitialised).
 */
#ifndef PHOBIC_HIP_H
#define PHOBIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PHOBIC_OK = 0,
  PHOBIC_E_INVALID_CONFIG = 1,   /* bad lambda / partition size / argument */
  PHOBIC_E_HIP = 2,              /* a HIP runtime call failed */
  PHOBIC_E_OOM = 3,              /* device memory exhausted / over budget */
  PHOBIC_E_NO_DEVICE = 4,        /* no usable gfx950 device */
  PHOBIC_E_PARTITION_OVERFLOW = 5, /* a partition holds > P + P/2 keys */
  PHOBIC_E_BUCKET_OVERFLOW = 6,  /* a bucket holds > 256 keys */
  PHOBIC_E_DUPLICATE_KEYS = 7,   /* two keys hash to the same 128 bits */
  PHOBIC_E_PILOT_OVERFLOW = 8,   /* no pilot value below 2^31 fits */
  PHOBIC_E_BUSY = 9,             /* serving: request rejected (queue full) */
  PHOBIC_E_NOT_FOUND = 10        /* serving: unknown request id */
} phobic_status;

/* MPHFconfig (include/phobicGpu/mphf_config.h:16-65).  With buckets = 0 and
 * fulcrums = NULL this is the default constructor (mphf_config.h:26-32):
 * B = ceil(P / lambda) and the OptBucketer fulcrum table.  A custom Bucketer
 * (mphf_config.h:34-42) passes its B and its 2048-entry table.  `seed` is an
 * extension; 0 reproduces the reference's XXH128(ptr, len, 0). */
typedef struct {
  double lambda;              /* averageBucketSize */
  uint32_t partition_size;    /* P; partitions hold at most P + P/2 keys */
  uint32_t buckets;           /* 0 = ceil(P / lambda); else B for `fulcrums` */
  uint64_t seed;              /* XXH3 seed */
  const uint32_t* fulcrums;   /* NULL = OptBucketer; else 2048 entries */
} phobic_config;

/* The reference's pre-hashed Key (include/phobicGpu/hasher.hpp:7-20). */
typedef struct {
  uint32_t partitioner, bucketer, lower1, lower2;
} phobic_key;

/* Device time of each stage in milliseconds (HIP events around each
 * launch), playing the role of the reference's GPU_* timestamp labels
 * (mphf_builder.h:138-146).  names[i] labels ms[i]; the stage split is this
 * library's own (hashing moved on device, bucket sort + redistribution +
 * search fused per partition), so the labels differ from the reference's. */
#define PHOBIC_MAX_STAGES 8
typedef struct {
  uint32_t count;
  float ms[PHOBIC_MAX_STAGES];
  char names[PHOBIC_MAX_STAGES][32];
  float total_ms; /* first launch start to last launch end */
} phobic_stage_times;

typedef struct phobic_ctx phobic_ctx;

/* Context: one per thread (or externally locked).  Owns a device memory
 * pool that grows to the largest build seen and is reused afterwards. */
int phobic_ctx_create(int device, phobic_ctx** out);
int phobic_ctx_destroy(phobic_ctx* ctx);
const char* phobic_status_string(int status);

/* np = ceil(n / P), B = ceil(P / lambda)  (mphf_builder.h:90-91, mphf_config.h:29). */
int phobic_query_sizes(const phobic_config* cfg, uint64_t n, uint32_t* np, uint32_t* buckets);

/* The 2048-entry fulcrum table (mphf_config.h:44-55, bucketers.h:28-39). */
int phobic_fulcrums(const phobic_config* cfg, uint32_t* fulcs_out);

/* Host XXH3-128 of a byte string (the query-side initial hash,
 * hasher.hpp:33-61); lo64/hi64 receive the low and high halves. */
int phobic_xxh3_128(const void* data, uint64_t len, uint64_t seed, uint64_t* lo64, uint64_t* hi64);

/* ---- builds from host buffers (synchronous, like MPHFbuilder::build) ---- */

/* u64 keys, hashed on the device (hasher.hpp:42-48 semantics). */
int phobic_build_u64(phobic_ctx* ctx, const uint64_t* keys, uint64_t n, const phobic_config* cfg,
                     uint32_t* pilots_out, uint32_t* offsets_out, phobic_stage_times* times);

/* Byte strings: key i = bytes[offsets[i] .. offsets[i+1]) (hasher.hpp:33-39). */
int phobic_build_bytes(phobic_ctx* ctx, const uint8_t* bytes, const uint64_t* offsets, uint64_t n,
                       const phobic_config* cfg, uint32_t* pilots_out, uint32_t* offsets_out,
                       phobic_stage_times* times);

/* Pre-hashed keys: the reference's nohash path (mphf_builder.h:202-207). */
int phobic_build_keys(phobic_ctx* ctx, const phobic_key* keys, uint64_t n, const phobic_config* cfg,
                      uint32_t* pilots_out, uint32_t* offsets_out, phobic_stage_times* times);

/* ---- builds on device-resident buffers (asynchronous on `stream`) ----
 * Inputs and outputs are device pointers.  The call enqueues the whole
 * build on `stream` (a hipStream_t; NULL = the null stream) and returns;
 * the build status is written to *d_status (device u32) by the kernels and
 * can be read after the stream completes.  Library scratch comes from the
 * context pool, so builds on one context must not overlap in time unless
 * they use distinct contexts. */
int phobic_build_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n,
                            const phobic_config* cfg, uint32_t* d_pilots, uint32_t* d_offsets,
                            uint32_t* d_status, void* stream);
int phobic_build_bytes_device(phobic_ctx* ctx, const uint8_t* d_bytes, const uint64_t* d_offsets,
                              uint64_t n, const phobic_config* cfg, uint32_t* d_pilots,
                              uint32_t* d_part_offsets, uint32_t* d_status, void* stream);
int phobic_build_keys_device(phobic_ctx* ctx, const phobic_key* d_keys, uint64_t n,
                             const phobic_config* cfg, uint32_t* d_pilots, uint32_t* d_offsets,
                             uint32_t* d_status, void* stream);

/* Stage timings of the last build on this context (device events). */
int phobic_last_stage_times(phobic_ctx* ctx, phobic_stage_times* times);

/* ---- evaluation (MPHF::operator(), include/phobicGpu/mphf.hpp:77-86) ----
 * Device-side batched query of u64 keys against (pilots, offsets) in HBM:
 * d_out[i] = position of d_keys[i] in [0, n). */
int phobic_eval_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t count,
                           const phobic_config* cfg, uint32_t np, const uint32_t* d_pilots,
                           const uint32_t* d_offsets, uint32_t* d_out, void* stream);

/* The validity contract of example.cpp:43-59 / benchmark.cpp:45-62 on the
 * device: d_bad (device u64) receives the number of positions in d_pos that
 * are >= n or repeat an earlier one; d_bitmap is caller scratch of
 * ceil(n/32) u32 words.  Asynchronous on `stream`. */
int phobic_check_bijection_device(phobic_ctx* ctx, const uint32_t* d_pos, uint64_t count, uint64_t n,
                                  uint32_t* d_bitmap, unsigned long long* d_bad, void* stream);

/* ---- synthetic inputs (device-side generators used by benchmarks) ----
 * d_keys[i] = splitmix64 output number i+1 from state `seed`
 * (state += 0x9E3779B97F4A7C15, then the 30/27/31 mix).  Distinct by
 * construction. */
int phobic_gen_splitmix64_device(uint64_t seed, uint64_t n, uint64_t* d_keys, void* stream);

#ifdef __cplusplus
}
#endif
#endif
