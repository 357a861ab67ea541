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
  PHOBIC_E_NOT_FOUND = 10,       /* serving: unknown request id */
  PHOBIC_E_SCRATCH_LIMIT = 11    /* allocation-free build: needs more than its fixed scratch */
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
  double alpha;               /* load factor in [0.5, 1]; 0 or 1 = the reference */
} phobic_config;

/* Build outputs (host or device pointers, as the entry point says).
 * alpha = 1: pilots and offsets only.  alpha < 1 (an extension with no
 * reference counterpart): partition p with ps keys searches a table of
 * m = ceil(ps / alpha) slots (ppm arithmetic); its used slots q in [ps, m)
 * are remapped, in increasing order, onto the free slots of [0, ps):
 * remap[remap_offsets[p] + q - ps] (0 for unused q); remap_offsets[np] is
 * the number of entries used (at most phobic_remap_capacity). */
typedef struct {
  uint32_t* pilots;         /* np*B, pilots[p + b*np] */
  uint32_t* offsets;        /* np+1 */
  uint32_t* remap;          /* alpha < 1 only */
  uint32_t* remap_offsets;  /* alpha < 1 only, np+1 */
} phobic_outputs;

/* The reference's pre-hashed Key (include/phobicGpu/hasher.hpp:7-20). */
typedef struct {
  uint32_t partitioner, bucketer, lower1, lower2;
} phobic_key;

/* Device time of each stage in milliseconds (HIP events around each
 * launch), playing the role of the reference's GPU_* timestamp labels
 * (mphf_builder.h:138-146, 229-290).  names[i] labels ms[i] in launch order:
 *   hash_route         hash, partition and bucket of every key, level-1 routing
 *   partition_split    routing levels 2..L: keys to their partitions
 *   partition_offsets  exclusive scan of the partition key counts
 *   partition_search   per partition: bucket sizes, bucket sort, keys in
 *                      bucket order, pilot search (and the alpha < 1 remap)
 * The reference's labels map onto these as: GPU_bucket_sizes = hash_route,
 * GPU_key_redistribution = partition_split, GPU_partition_offsets =
 * partition_offsets, GPU_search = partition_search, and GPU_bucket_sorting,
 * GPU_apply_partition_offsets and GPU_memory_map = 0 (fused into
 * partition_search, not needed, outputs stay in device memory); the C++
 * MPHFbuilder::build emits the reference's labels.  The host-buffer entry
 * points also time their copies: input_transfer_ms from the first input
 * copy to the first kernel, result_transfer_ms from the last kernel's end to
 * the outputs in host memory (0 for the device entry points). */
#define PHOBIC_MAX_STAGES 8
typedef struct {
  uint32_t count;
  float ms[PHOBIC_MAX_STAGES];
  char names[PHOBIC_MAX_STAGES][32];
  float total_ms; /* first launch start to last launch end */
  float input_transfer_ms;
  float result_transfer_ms;
} phobic_stage_times;

typedef struct phobic_ctx phobic_ctx;

/* Context: one per thread (or externally locked).  Owns a device memory
 * pool that grows to the largest build seen and is reused afterwards. */
int phobic_ctx_create(int device, phobic_ctx** out);
int phobic_ctx_destroy(phobic_ctx* ctx);
const char* phobic_status_string(int status);

/* np = ceil(n / P), B = ceil(P / lambda)  (mphf_builder.h:90-91, mphf_config.h:29). */
int phobic_query_sizes(const phobic_config* cfg, uint64_t n, uint32_t* np, uint32_t* buckets);

/* Upper bound of remap entries for n keys (0 when alpha = 1). */
int phobic_remap_capacity(const phobic_config* cfg, uint64_t n, uint64_t* entries);

/* The 2048-entry fulcrum table (mphf_config.h:44-55, bucketers.h:28-39). */
int phobic_fulcrums(const phobic_config* cfg, uint32_t* fulcs_out);

/* Host XXH3-128 of a byte string (the query-side initial hash,
 * hasher.hpp:33-61); lo64/hi64 receive the low and high halves. */
int phobic_xxh3_128(const void* data, uint64_t len, uint64_t seed, uint64_t* lo64, uint64_t* hi64);

/* ---- builds from host buffers (synchronous, like MPHFbuilder::build) ---- */

/* u64 keys, hashed on the device (hasher.hpp:42-48 semantics). */
int phobic_build_u64(phobic_ctx* ctx, const uint64_t* keys, uint64_t n, const phobic_config* cfg,
                     const phobic_outputs* out, phobic_stage_times* times);

/* Byte strings: key i = bytes[offsets[i] .. offsets[i+1]) (hasher.hpp:33-39). */
int phobic_build_bytes(phobic_ctx* ctx, const uint8_t* bytes, const uint64_t* offsets, uint64_t n,
                       const phobic_config* cfg, const phobic_outputs* out, phobic_stage_times* times);

/* Pre-hashed keys: the reference's nohash path (mphf_builder.h:202-207). */
int phobic_build_keys(phobic_ctx* ctx, const phobic_key* keys, uint64_t n, const phobic_config* cfg,
                      const phobic_outputs* out, phobic_stage_times* times);

/* ---- builds on device-resident buffers (asynchronous on `stream`) ----
 * Inputs and outputs are device pointers.  The call enqueues the whole
 * build on `stream` (a hipStream_t; NULL = the null stream) and returns;
 * the build status is written to *d_status (device u32) by the kernels and
 * can be read after the stream completes.  Library scratch comes from the
 * context pool, so builds on one context must not overlap in time unless
 * they use distinct contexts. */
int phobic_build_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n,
                            const phobic_config* cfg, const phobic_outputs* d_out, uint32_t* d_status,
                            void* stream);
int phobic_build_bytes_device(phobic_ctx* ctx, const uint8_t* d_bytes, const uint64_t* d_offsets,
                              uint64_t n, const phobic_config* cfg, const phobic_outputs* d_out,
                              uint32_t* d_status, void* stream);
int phobic_build_keys_device(phobic_ctx* ctx, const phobic_key* d_keys, uint64_t n,
                             const phobic_config* cfg, const phobic_outputs* d_out, uint32_t* d_status,
                             void* stream);

/* ---- allocation-free build (no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments, results and asynchrony as phobic_build_u64_device, but the
 * build's scratch (partitioning records, bin counters, scan totals, the
 * fulcrum table) lives in one fixed device array of
 * PHOBIC_NOALLOC_SCRATCH_BYTES per device, which each device the library
 * runs on holds until the process ends; the call makes no allocation.  A
 * build needs about 24 * (P + P/2) bytes per partition (partitions rounded
 * up to the routing's bin size) plus small tables: at P = 2048 up to about
 * 1.8 million keys fit.  Larger builds return PHOBIC_E_SCRATCH_LIMIT before
 * anything is enqueued.  Every call on a device shares that device's array
 * and the array keeps the last call's contents (what a build needs zeroed
 * it clears itself): calls that overlap in time on one device race and
 * give wrong results, so serialise them (one stream, or wait between
 * calls).  The context's stage timers record the build as for
 * phobic_build_u64_device. */
#define PHOBIC_NOALLOC_SCRATCH_BYTES (64u << 20)
int phobic_build_u64_device_noalloc(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n, const phobic_config* cfg,
                                    const phobic_outputs* d_out, uint32_t* d_status, void* stream);

/* ---- batched builds: many independent u64 key sets, one configuration ----
 * set s = d_keys[set_offsets[s] .. set_offsets[s+1]) (set_offsets: host,
 * sets+1 entries, set_offsets[0] = 0).  All sets' partitions are built as one
 * flattened device build; every set's pilots and offsets equal those of
 * building it alone.  Outputs of set s (device):
 *   pilots  at d_pilots  + B * part_base[s]   (np_s*B, pilots[p + b*np_s])
 *   offsets at d_offsets + part_base[s] + s   (np_s + 1)
 * with part_base[s] = sum over s' < s of ceil(n_s' / P) (phobic_batch_layout).
 * alpha must be 1.  The batch keeps device copies of the set layout; the
 * device scratch is the context's (one batch or build at a time per context).
 * phobic_batch_launch enqueues the kernels on `stream`;
 * phobic_batch_graph_launch captures them into a hipGraph on first use
 * (and again whenever the device pointers change) and replays it. */
typedef struct phobic_batch phobic_batch;
int phobic_batch_create(phobic_ctx* ctx, const phobic_config* cfg, const uint64_t* set_offsets, uint32_t sets,
                        phobic_batch** out);
int phobic_batch_layout(const phobic_batch* batch, uint32_t* part_base_out, uint32_t* buckets);
int phobic_batch_launch(phobic_batch* batch, const uint64_t* d_keys, uint32_t* d_pilots, uint32_t* d_offsets,
                        uint32_t* d_status, void* stream);
int phobic_batch_graph_launch(phobic_batch* batch, const uint64_t* d_keys, uint32_t* d_pilots,
                              uint32_t* d_offsets, uint32_t* d_status, void* stream);
int phobic_batch_destroy(phobic_batch* batch);

/* ---- serving: continuous-batching build server ----
 * Requests queue in arrival order.  Whenever one of `slots` workers is idle,
 * the scheduler takes the oldest request that fits the HBM budget (a later
 * one may overtake a request that does not fit yet) together with every
 * queued request of the same key kind, P, alpha and size class, whatever
 * its lambda (up to four lambdas per launch) and seed: small requests (at
 * most 4e6 keys) up to 64e6 keys in all, large ones up to 256e6, 1024
 * requests at most.  They are built as
 * one flattened launch whose per-request outputs equal building each alone.
 * Requests with a custom fulcrum table, and byte strings already in device
 * memory, run alone.  Each worker owns a context and stream, so admitted
 * launches run concurrently on the device; host inputs are uploaded through
 * the worker context's pinned staging chunks (two of 64 MiB per worker,
 * allocated by phobic_server_create).  Results stay in device memory until
 * phobic_server_release.  Host key buffers must stay valid until the
 * request completes.
 * Memory: hbm_in_use counts the outputs kept for unreleased requests, the
 * staging, outputs and scratch growth charged to running launches, and the
 * scratch each worker's context keeps between launches (reused by its next
 * launch), all in the library pool's 2 MiB units; admission keeps it within
 * hbm_budget.  A request whose own need exceeds the budget is refused at
 * submit.  No counterpart in the reference (one synchronous
 * MPHFbuilder::build per MPHF). */
#define PHOBIC_KEYS_U64 0
#define PHOBIC_KEYS_BYTES 1
typedef struct {
  int kind;                 /* PHOBIC_KEYS_U64 or PHOBIC_KEYS_BYTES */
  int keys_on_device;       /* 0: host pointers, 1: device pointers */
  const void* keys;         /* u64[n] or the concatenated bytes */
  const uint64_t* offsets;  /* bytes: n+1 non-decreasing offsets into keys */
  uint64_t bytes_len;       /* bytes: at least offsets[n] */
  uint64_t n;
  phobic_config cfg;
} phobic_request;
typedef struct {
  uint32_t slots;           /* concurrent launches (1..64) */
  int device;
  uint64_t hbm_budget;      /* device bytes the server may hold */
  uint32_t queue_capacity;  /* 0 = 65536 */
} phobic_server_config;
typedef struct {
  uint64_t submitted, completed, rejected, keys_built, hbm_in_use;
  double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion latency */
  double span_s;                          /* first submit -> last completion */
  uint64_t grouped_launches;              /* launches that built several requests */
  uint64_t grouped_requests;              /* requests built in such launches */
  uint64_t pool_used;                     /* bytes the library's device pool has handed out */
} phobic_server_stats;
typedef struct phobic_server phobic_server;
int phobic_server_create(const phobic_server_config* cfg, phobic_server** out);
/* PHOBIC_E_BUSY when the queue is full, PHOBIC_E_OOM when the request alone
 * exceeds the budget, PHOBIC_E_INVALID_CONFIG for a malformed request (host
 * string offsets must be non-decreasing and end within bytes_len). */
int phobic_server_submit(phobic_server* server, const phobic_request* req, uint64_t* id);
int phobic_server_poll(phobic_server* server, uint64_t id, int* done);
int phobic_server_wait(phobic_server* server, uint64_t id, int* status, double* latency_ms);
/* Sizes and device addresses of a completed request's outputs. */
int phobic_server_result(phobic_server* server, uint64_t id, uint32_t* np, uint32_t* buckets,
                         uint64_t* remap_entries, phobic_outputs* device_outputs);
/* Copies a completed request's outputs into host buffers. */
int phobic_server_fetch(phobic_server* server, uint64_t id, const phobic_outputs* host_out);
int phobic_server_release(phobic_server* server, uint64_t id);
int phobic_server_stats_get(phobic_server* server, phobic_server_stats* stats);
int phobic_server_destroy(phobic_server* server);

/* ---- launch shapes and the autotuner ----
 * search_waves: partitions (one wave64 each) per search workgroup (1,2,4,8);
 * route_per: records per thread of a level-1 (hash and route) tile of u64
 *   or pre-hashed keys; byte strings always use tiles of 1024 x 8;
 * max_fanout: bins per partitioning level (16..1024);
 * search_mode: pilot-search variants, a mask of PHOBIC_SEARCH_* (0 = all
 *   of the fast paths);
 * split_per: records per thread of a tile of levels 2..L (0 = route_per);
 * first_fanout: at most this many level-1 bins (0 = max_fanout, except
 *   that a two-level plan whose balanced fan-out would exceed 576 keeps
 *   level 1 at 576 bins); the later levels then take the larger share of
 *   the fan-out;
 * route_threads: threads of a routing workgroup, 512 (0) or 1024; with 512
 *   route_per and split_per are 8, 12, 16, 20 or 24, with 1024 they are 4,
 *   6, 8, 10 or 12.  A level whose tile would not fit the workgroup's LDS
 *   uses the largest smaller per that does.  Plain u64 builds at level 1,
 *   and splits of at most 1024 parent bins, run persistent workgroups (one
 *   per CU the device holds) over tiles of this shape.
 * Defaults {1, 12, 1024, 0, 12, 0, 1024}: tiles of 12,288 records, and two
 * levels of 576 x 848 bins for a billion keys (P 2048).
 * phobic_autotune times eight tile shapes x first_fanout {0, 48, 512} x
 * max_fanout {128, 1024} (on the partitioning stages' device time) and then
 * the search shapes {1,2,4,8} x {0,1,2,3} (on the search stage's time) on
 * `sample_keys` device-generated keys built with `cfg`, keeps the fastest
 * on the context and returns it.  Results do not depend on the tuning.  The
 * sample should span a few 1e4 partitions (tens of millions of keys): on
 * fewer the search stage underfills the GPU and favours wide workgroups. */
enum {
  PHOBIC_SEARCH_SINGLE_PILOT = 1, /* one pilot per hash round for every bucket */
  PHOBIC_SEARCH_NARROW = 2        /* scan passes of 2048 offsets for every table size */
};
typedef struct {
  uint32_t search_waves;
  uint32_t route_per;
  uint32_t max_fanout;
  uint32_t search_mode;
  uint32_t split_per;
  uint32_t first_fanout;
  uint32_t route_threads;
} phobic_tuning;
int phobic_get_tuning(phobic_ctx* ctx, phobic_tuning* tuning);
int phobic_set_tuning(phobic_ctx* ctx, const phobic_tuning* tuning);
int phobic_autotune(phobic_ctx* ctx, const phobic_config* cfg, uint64_t sample_keys, phobic_tuning* best,
                    float* best_ms);

/* Stage timings of the last build on this context (device events). */
int phobic_last_stage_times(phobic_ctx* ctx, phobic_stage_times* times);

/* ---- evaluation (MPHF::operator(), include/phobicGpu/mphf.hpp:77-86) ----
 * Device-side batched queries against an MPHF in HBM (pilots, offsets and,
 * for alpha < 1, the remap table, as a build wrote them): d_out[i] = the
 * position in [0, n) of key i, computed as the reference's operator() does
 * for the key type's initial hash (hasher.hpp:22-66).  A key that was not
 * built and falls in a partition without keys gets offsets[p].
 * Asynchronous on `stream`. */
/* uint64_t keys: XXH3-128 of the 8 native bytes (hasher.hpp:42-48). */
int phobic_eval_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t count,
                           const phobic_config* cfg, uint32_t np, const phobic_outputs* d_mphf,
                           uint32_t* d_out, void* stream);
/* std::string keys: key i = d_bytes[d_offsets[i] .. d_offsets[i+1])
 * (hasher.hpp:33-39). */
int phobic_eval_bytes_device(phobic_ctx* ctx, const uint8_t* d_bytes, const uint64_t* d_offsets, uint64_t count,
                             const phobic_config* cfg, uint32_t np, const phobic_outputs* d_mphf,
                             uint32_t* d_out, void* stream);
/* 16-byte keys hashed with XXH3-128: std::pair<uint64_t, uint64_t> (first,
 * second at d_pairs[2i], d_pairs[2i+1]) or a Key under the xxhash hasher
 * (hasher.hpp:51-66). */
int phobic_eval_pairs_device(phobic_ctx* ctx, const uint64_t* d_pairs, uint64_t count, const phobic_config* cfg,
                             uint32_t np, const phobic_outputs* d_mphf, uint32_t* d_out, void* stream);
/* Pre-hashed Keys (the nohash hasher, hasher.hpp:22-26). */
int phobic_eval_keys_device(phobic_ctx* ctx, const phobic_key* d_keys, uint64_t count, const phobic_config* cfg,
                            uint32_t np, const phobic_outputs* d_mphf, uint32_t* d_out, void* stream);

/* The validity contract of example.cpp:43-59 / benchmark.cpp:45-62 on the
 * device: d_bad (device u64) receives the number of positions in d_pos that
 * are >= n or repeat an earlier one; d_bitmap is caller scratch of
 * ceil(n/32) u32 words.  Asynchronous on `stream`. */
int phobic_check_bijection_device(phobic_ctx* ctx, const uint32_t* d_pos, uint64_t count, uint64_t n,
                                  uint32_t* d_bitmap, unsigned long long* d_bad, void* stream);

/* Device-to-device copy by a plain streaming kernel (16-byte words; both
 * addresses 16-byte aligned, bytes a multiple of 16), asynchronous on
 * `stream`.  No counterpart in the reference: it is the known-size kernel
 * against which HBM traffic counters (FETCH_SIZE, WRITE_SIZE) are scaled. */
int phobic_copy_device(void* d_dst, const void* d_src, uint64_t bytes, void* stream);

/* Host-to-device copy of `bytes` from host memory h_src (pageable) to d_dst
 * on `stream`, through the context's two pinned staging chunks for copies of
 * 16 MiB and more (the path the host-buffer builds and the server use);
 * returns when the copy has completed.  Replaces the staging-buffer upload
 * of the reference's host inputs (mphf_builder.h:231-238, a host-visible
 * Vulkan buffer filled by memcpy). */
int phobic_copy_to_device(phobic_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes, void* stream);

/* ---- synthetic inputs (device-side generators used by benchmarks) ----
 * d_keys[i] = splitmix64 output number i+1 from state `seed`
 * (state += 0x9E3779B97F4A7C15, then the 30/27/31 mix).  Distinct by
 * construction. */
int phobic_gen_splitmix64_device(uint64_t seed, uint64_t n, uint64_t* d_keys, void* stream);
/* String i fills d_bytes[d_offsets[i], d_offsets[i+1]) with printable ASCII
 * 0x20..0x7E drawn from a splitmix64 stream seeded by (seed, i).  Strings
 * are not made distinct: duplicates are unlikely at lengths of 8 and more,
 * and a build that meets one returns PHOBIC_E_DUPLICATE_KEYS. */
int phobic_gen_strings_device(uint64_t seed, uint64_t n, const uint64_t* d_offsets, uint8_t* d_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
