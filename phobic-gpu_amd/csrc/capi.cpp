// capi.cpp -- C ABI of the gfx950 PHOBIC construction path (include/phobic_hip.h).
//
// Host side of one build (the role of BuildInvocation::run,
// include/phobicGpu/mphf_builder.h:225-292): validate the configuration,
// derive np / B / fulcrums (mphf_config.h:26-55), size the context's device
// pool, enqueue the three stage kernels on one stream, and (for the
// host-buffer entry points) copy inputs in and pilots / offsets out.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "../../include/phobic_hip.h"
#include "kernels.hpp"

using namespace phobic;

namespace {

struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && ptr) return hipSuccess;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      cap = 0;
    }
    const size_t want = bytes ? bytes : 256;
    hipError_t e = hipMalloc(&ptr, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

constexpr int kStages = 3;
const char* const kStageNames[kStages] = {"hash_partition_scatter", "partition_offsets", "partition_search"};

}  // namespace

struct phobic_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // scratch owned by the context
  DevBuf pcount, rec_lower, rec_bkt, fulcs, status;
  // staging for host-buffer entry points
  DevBuf in_a, in_b, out_pilots, out_offsets;
  hipEvent_t ev[kStages + 1] = {};
  bool timed = false;
  // fulcrum cache
  double f_lambda = -1;
  uint32_t f_P = 0;
  uint32_t f_host[kFulcsTable];
};

namespace {

int hip_status(hipError_t e) { return e == hipSuccess ? PHOBIC_OK : (e == hipErrorOutOfMemory ? PHOBIC_E_OOM : PHOBIC_E_HIP); }

#define TRY_HIP(x)                         \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return hip_status(e_); \
  } while (0)

#define TRY(x)                \
  do {                        \
    int s_ = (x);             \
    if (s_ != PHOBIC_OK) return s_; \
  } while (0)

struct Plan {
  uint64_t n;
  uint32_t np, B, cap;
};

int validate(const phobic_config* cfg, uint64_t n, Plan* pl) {
  if (!cfg || !(cfg->lambda > 0.0) || !std::isfinite(cfg->lambda)) return PHOBIC_E_INVALID_CONFIG;
  if (cfg->partition_size < 2 || cfg->partition_size > (1u << 20) || cfg->flags != 0) return PHOBIC_E_INVALID_CONFIG;
  if (n > 0xFFFFFFFFull) return PHOBIC_E_INVALID_CONFIG;  // offsets are u32 (mphf_builder.h:254)
  const double Bd = std::ceil((double)cfg->partition_size / cfg->lambda);
  if (!(Bd >= 1.0) || Bd > 65535.0) return PHOBIC_E_INVALID_CONFIG;
  pl->n = n;
  pl->B = (uint32_t)Bd;
  pl->np = (uint32_t)((n + cfg->partition_size - 1) / cfg->partition_size);
  pl->cap = cfg->partition_size + cfg->partition_size / 2;  // mphf_config.h:61-64
  if (search_lds_bytes(pl->cap, pl->B) > 64 * 1024) return PHOBIC_E_INVALID_CONFIG;
  return PHOBIC_OK;
}

// MPHFconfig::setBucketer with OptBucketer (mphf_config.h:44-55,
// bucketers.h:28-39), evaluated in double precision on the host.
void compute_fulcrums(double lambda, uint32_t P, uint32_t* f) {
  const uint32_t B = (uint32_t)std::ceil((double)P / lambda);
  const double c = 0.2 * lambda / std::sqrt((double)P);
  const double scale = (double)(uint32_t)(B << 16);
  f[0] = 0;
  for (uint32_t xi = 1; xi < kFulcsTable - 1; xi++) {
    const double x = (double)xi / (double)(kFulcsTable - 1);
    const double y = (x + (1.0 - x) * std::log(1.0 - x)) * (1.0 - c) + c * x;
    f[xi] = (uint32_t)(y * scale);
  }
  f[kFulcsTable - 1] = B << 16;
}

int set_device(phobic_ctx* ctx) {
  TRY_HIP(hipSetDevice(ctx->device));
  return PHOBIC_OK;
}

// Enqueue one build.  Inputs/outputs are device pointers.
int enqueue_build(phobic_ctx* ctx, int kind, const KeySource& src, const phobic_config* cfg, const Plan& pl,
                  uint32_t* d_pilots, uint32_t* d_offsets, uint32_t* d_status, hipStream_t s) {
  if (ctx->f_lambda != cfg->lambda || ctx->f_P != cfg->partition_size) {
    compute_fulcrums(cfg->lambda, cfg->partition_size, ctx->f_host);
    ctx->f_lambda = cfg->lambda;
    ctx->f_P = cfg->partition_size;
    TRY_HIP(ctx->fulcs.ensure(kFulcsTable * 4));
    TRY_HIP(hipMemcpy(ctx->fulcs.ptr, ctx->f_host, kFulcsTable * 4, hipMemcpyHostToDevice));
  }
  const uint64_t slots = (uint64_t)pl.np * pl.cap;
  TRY_HIP(ctx->pcount.ensure((size_t)pl.np * 4));
  TRY_HIP(ctx->rec_lower.ensure(slots * 8));
  TRY_HIP(ctx->rec_bkt.ensure(slots * 2));
  TRY_HIP(launch_fill_u32(d_status, 1, 0, s));
  TRY_HIP(launch_fill_u32(ctx->pcount.as<uint32_t>(), pl.np, 0, s));
  TRY_HIP(hipEventRecord(ctx->ev[0], s));

  ScatterArgs sa{};
  sa.src = src;
  sa.n = pl.n;
  sa.seed = cfg->seed;
  sa.np = pl.np;
  sa.cap = pl.cap;
  sa.fulcs = ctx->fulcs.as<uint32_t>();
  sa.pcount = ctx->pcount.as<uint32_t>();
  sa.rec_lower = ctx->rec_lower.as<uint64_t>();
  sa.rec_bkt = ctx->rec_bkt.as<uint16_t>();
  sa.status = d_status;
  TRY_HIP(launch_hash_scatter(kind, sa, s));
  TRY_HIP(hipEventRecord(ctx->ev[1], s));

  TRY_HIP(launch_partition_offsets(ctx->pcount.as<uint32_t>(), pl.np, d_offsets, s));
  TRY_HIP(hipEventRecord(ctx->ev[2], s));

  SearchArgs ra{};
  ra.rec_lower = ctx->rec_lower.as<uint64_t>();
  ra.rec_bkt = ctx->rec_bkt.as<uint16_t>();
  ra.pcount = ctx->pcount.as<uint32_t>();
  ra.np = pl.np;
  ra.B = pl.B;
  ra.cap = pl.cap;
  ra.pilots = d_pilots;
  ra.status = d_status;
  TRY_HIP(launch_partition_search(ra, s));
  TRY_HIP(hipEventRecord(ctx->ev[3], s));
  ctx->timed = true;
  return PHOBIC_OK;
}

int status_from_bits(uint32_t bits) {
  if (!bits) return PHOBIC_OK;
  for (int c = 1; c < 32; c++)
    if (bits & (1u << c)) return c;
  return PHOBIC_E_HIP;
}

int fill_times(phobic_ctx* ctx, phobic_stage_times* t) {
  if (!t) return PHOBIC_OK;
  std::memset(t, 0, sizeof(*t));
  if (!ctx->timed) return PHOBIC_OK;
  t->count = kStages;
  for (int i = 0; i < kStages; i++) {
    float ms = 0;
    TRY_HIP(hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
    t->ms[i] = ms;
    std::snprintf(t->names[i], sizeof(t->names[i]), "%s", kStageNames[i]);
  }
  TRY_HIP(hipEventElapsedTime(&t->total_ms, ctx->ev[0], ctx->ev[kStages]));
  return PHOBIC_OK;
}

// Host-buffer build: stage inputs, enqueue on the context stream, copy out.
int host_build(phobic_ctx* ctx, int kind, const void* a, size_t a_bytes, const void* b, size_t b_bytes,
               uint64_t n, const phobic_config* cfg, uint32_t* pilots_out, uint32_t* offsets_out,
               phobic_stage_times* times) {
  if (!ctx || !pilots_out || !offsets_out) return PHOBIC_E_INVALID_CONFIG;
  Plan pl;
  TRY(validate(cfg, n, &pl));
  TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  TRY_HIP(ctx->in_a.ensure(a_bytes));
  if (a_bytes) TRY_HIP(hipMemcpyAsync(ctx->in_a.ptr, a, a_bytes, hipMemcpyHostToDevice, s));
  if (b) {
    TRY_HIP(ctx->in_b.ensure(b_bytes));
    TRY_HIP(hipMemcpyAsync(ctx->in_b.ptr, b, b_bytes, hipMemcpyHostToDevice, s));
  }
  const uint64_t NB = (uint64_t)pl.np * pl.B;
  TRY_HIP(ctx->out_pilots.ensure(NB * 4 + 4));  // + status word
  TRY_HIP(ctx->out_offsets.ensure(((uint64_t)pl.np + 1) * 4));
  uint32_t* d_status = ctx->out_pilots.as<uint32_t>() + NB;
  KeySource src{};
  if (kind == kKindU64) src.u64 = ctx->in_a.as<uint64_t>();
  if (kind == kKindBytes) {
    src.bytes = ctx->in_a.as<uint8_t>();
    src.offs = ctx->in_b.as<uint64_t>();
  }
  if (kind == kKindKey) src.keys = ctx->in_a.as<uint4>();
  TRY(enqueue_build(ctx, kind, src, cfg, pl, ctx->out_pilots.as<uint32_t>(), ctx->out_offsets.as<uint32_t>(),
                    d_status, s));
  uint32_t bits = 0;
  TRY_HIP(hipMemcpyAsync(&bits, d_status, 4, hipMemcpyDeviceToHost, s));
  if (NB) TRY_HIP(hipMemcpyAsync(pilots_out, ctx->out_pilots.ptr, NB * 4, hipMemcpyDeviceToHost, s));
  TRY_HIP(hipMemcpyAsync(offsets_out, ctx->out_offsets.ptr, ((uint64_t)pl.np + 1) * 4, hipMemcpyDeviceToHost, s));
  TRY_HIP(hipStreamSynchronize(s));
  TRY(fill_times(ctx, times));
  return status_from_bits(bits);
}

int device_build(phobic_ctx* ctx, int kind, const KeySource& src, uint64_t n, const phobic_config* cfg,
                 uint32_t* d_pilots, uint32_t* d_offsets, uint32_t* d_status, void* stream) {
  if (!ctx || !d_pilots || !d_offsets || !d_status) return PHOBIC_E_INVALID_CONFIG;
  Plan pl;
  TRY(validate(cfg, n, &pl));
  TRY(set_device(ctx));
  return enqueue_build(ctx, kind, src, cfg, pl, d_pilots, d_offsets, d_status, (hipStream_t)stream);
}

}  // namespace

extern "C" {

int phobic_ctx_create(int device, phobic_ctx** out) {
  if (!out) return PHOBIC_E_INVALID_CONFIG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PHOBIC_E_NO_DEVICE;
  if (device < 0 || device >= count) return PHOBIC_E_NO_DEVICE;
  TRY_HIP(hipSetDevice(device));
  phobic_ctx* ctx = new (std::nothrow) phobic_ctx();
  if (!ctx) return PHOBIC_E_OOM;
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  for (int i = 0; e == hipSuccess && i <= kStages; i++) e = hipEventCreate(&ctx->ev[i]);
  if (e != hipSuccess) {
    phobic_ctx_destroy(ctx);
    return hip_status(e);
  }
  *out = ctx;
  return PHOBIC_OK;
}

int phobic_ctx_destroy(phobic_ctx* ctx) {
  if (!ctx) return PHOBIC_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (DevBuf* b : {&ctx->pcount, &ctx->rec_lower, &ctx->rec_bkt, &ctx->fulcs, &ctx->status, &ctx->in_a, &ctx->in_b,
                    &ctx->out_pilots, &ctx->out_offsets})
    b->release();
  for (int i = 0; i <= kStages; i++)
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PHOBIC_OK;
}

const char* phobic_status_string(int s) {
  switch (s) {
    case PHOBIC_OK: return "ok";
    case PHOBIC_E_INVALID_CONFIG: return "invalid configuration or argument";
    case PHOBIC_E_HIP: return "HIP runtime error";
    case PHOBIC_E_OOM: return "out of device memory";
    case PHOBIC_E_NO_DEVICE: return "no usable device";
    case PHOBIC_E_PARTITION_OVERFLOW: return "a partition holds more than P + P/2 keys";
    case PHOBIC_E_BUCKET_OVERFLOW: return "a bucket holds more than 256 keys";
    case PHOBIC_E_DUPLICATE_KEYS: return "duplicate keys (identical 128-bit hashes)";
    case PHOBIC_E_PILOT_OVERFLOW: return "no pilot below 2^31 places a bucket";
    case PHOBIC_E_BUSY: return "request rejected: queue or memory budget full";
    case PHOBIC_E_NOT_FOUND: return "unknown request";
    default: return "unknown status";
  }
}

int phobic_query_sizes(const phobic_config* cfg, uint64_t n, uint32_t* np, uint32_t* buckets) {
  Plan pl;
  TRY(validate(cfg, n, &pl));
  if (np) *np = pl.np;
  if (buckets) *buckets = pl.B;
  return PHOBIC_OK;
}

int phobic_fulcrums(const phobic_config* cfg, uint32_t* out) {
  Plan pl;
  TRY(validate(cfg, 0, &pl));
  if (!out) return PHOBIC_E_INVALID_CONFIG;
  compute_fulcrums(cfg->lambda, cfg->partition_size, out);
  return PHOBIC_OK;
}

int phobic_build_u64(phobic_ctx* ctx, const uint64_t* keys, uint64_t n, const phobic_config* cfg,
                     uint32_t* pilots_out, uint32_t* offsets_out, phobic_stage_times* times) {
  if (n && !keys) return PHOBIC_E_INVALID_CONFIG;
  return host_build(ctx, kKindU64, keys, n * 8, nullptr, 0, n, cfg, pilots_out, offsets_out, times);
}

int phobic_build_bytes(phobic_ctx* ctx, const uint8_t* bytes, const uint64_t* offsets, uint64_t n,
                       const phobic_config* cfg, uint32_t* pilots_out, uint32_t* offsets_out,
                       phobic_stage_times* times) {
  if (!offsets || (n && !bytes && offsets[n] != offsets[0])) return PHOBIC_E_INVALID_CONFIG;
  for (uint64_t i = 0; i < n; i++)
    if (offsets[i + 1] < offsets[i]) return PHOBIC_E_INVALID_CONFIG;
  // keys index the byte buffer from offsets[0]; ship the whole prefix
  const uint64_t total = offsets[n];
  return host_build(ctx, kKindBytes, bytes, total, offsets, (n + 1) * 8, n, cfg, pilots_out, offsets_out, times);
}

int phobic_build_keys(phobic_ctx* ctx, const phobic_key* keys, uint64_t n, const phobic_config* cfg,
                      uint32_t* pilots_out, uint32_t* offsets_out, phobic_stage_times* times) {
  if (n && !keys) return PHOBIC_E_INVALID_CONFIG;
  return host_build(ctx, kKindKey, keys, n * 16, nullptr, 0, n, cfg, pilots_out, offsets_out, times);
}

int phobic_build_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n, const phobic_config* cfg,
                            uint32_t* d_pilots, uint32_t* d_offsets, uint32_t* d_status, void* stream) {
  KeySource src{};
  src.u64 = d_keys;
  if (n && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  return device_build(ctx, kKindU64, src, n, cfg, d_pilots, d_offsets, d_status, stream);
}

int phobic_build_bytes_device(phobic_ctx* ctx, const uint8_t* d_bytes, const uint64_t* d_offsets, uint64_t n,
                              const phobic_config* cfg, uint32_t* d_pilots, uint32_t* d_part_offsets,
                              uint32_t* d_status, void* stream) {
  KeySource src{};
  src.bytes = d_bytes;
  src.offs = d_offsets;
  if (!d_offsets) return PHOBIC_E_INVALID_CONFIG;
  return device_build(ctx, kKindBytes, src, n, cfg, d_pilots, d_part_offsets, d_status, stream);
}

int phobic_build_keys_device(phobic_ctx* ctx, const phobic_key* d_keys, uint64_t n, const phobic_config* cfg,
                             uint32_t* d_pilots, uint32_t* d_offsets, uint32_t* d_status, void* stream) {
  KeySource src{};
  src.keys = reinterpret_cast<const uint4*>(d_keys);
  if (n && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  return device_build(ctx, kKindKey, src, n, cfg, d_pilots, d_offsets, d_status, stream);
}

int phobic_last_stage_times(phobic_ctx* ctx, phobic_stage_times* times) {
  if (!ctx || !times) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  for (int i = 0; i <= kStages && ctx->timed; i++) TRY_HIP(hipEventSynchronize(ctx->ev[i]));
  return fill_times(ctx, times);
}

int phobic_eval_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t count, const phobic_config* cfg,
                           uint32_t np, const uint32_t* d_pilots, const uint32_t* d_offsets, uint32_t* d_out,
                           void* stream) {
  if (!ctx || (count && (!d_keys || !d_pilots || !d_offsets || !d_out)) || np == 0) return PHOBIC_E_INVALID_CONFIG;
  Plan pl;
  TRY(validate(cfg, 0, &pl));
  TRY(set_device(ctx));
  if (ctx->f_lambda != cfg->lambda || ctx->f_P != cfg->partition_size) {
    compute_fulcrums(cfg->lambda, cfg->partition_size, ctx->f_host);
    ctx->f_lambda = cfg->lambda;
    ctx->f_P = cfg->partition_size;
    TRY_HIP(ctx->fulcs.ensure(kFulcsTable * 4));
    TRY_HIP(hipMemcpy(ctx->fulcs.ptr, ctx->f_host, kFulcsTable * 4, hipMemcpyHostToDevice));
  }
  EvalArgs ea{};
  ea.keys = d_keys;
  ea.count = count;
  ea.seed = cfg->seed;
  ea.np = np;
  ea.fulcs = ctx->fulcs.as<uint32_t>();
  ea.pilots = d_pilots;
  ea.offsets = d_offsets;
  ea.out = d_out;
  TRY_HIP(launch_eval_u64(ea, (hipStream_t)stream));
  return PHOBIC_OK;
}

int phobic_check_bijection_device(phobic_ctx* ctx, const uint32_t* d_pos, uint64_t count, uint64_t n,
                                  uint32_t* d_bitmap, unsigned long long* d_bad, void* stream) {
  if (!ctx || (count && !d_pos) || !d_bitmap || !d_bad) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;
  TRY_HIP(launch_fill_u32(d_bitmap, (n + 31) / 32, 0, s));
  TRY_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(d_bad), 2, 0, s));
  TRY_HIP(launch_check_bijection(d_pos, count, n, d_bitmap, d_bad, s));
  return PHOBIC_OK;
}

int phobic_gen_splitmix64_device(uint64_t seed, uint64_t n, uint64_t* d_keys, void* stream) {
  if (n && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  TRY_HIP(launch_splitmix64(seed, n, d_keys, (hipStream_t)stream));
  return PHOBIC_OK;
}

}  // extern "C"
