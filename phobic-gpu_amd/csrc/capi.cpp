// capi.cpp -- C ABI of the gfx950 PHOBIC construction path (include/phobic_hip.h).
//
// Host side of one build (the role of BuildInvocation::run,
// include/phobicGpu/mphf_builder.h:225-292): validate the configuration,
// derive np / B / fulcrums (mphf_config.h:26-55), size the context's device
// pool, enqueue the three stage kernels on one stream, and (for the
// host-buffer entry points) copy inputs in and pilots / offsets out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/phobic_hip.h"
#include "internal.hpp"
#include "kernels.hpp"

using namespace phobic;

namespace {

// Device memory pools, one per device, created on first use.  The release
// threshold keeps freed memory in the pool.
std::mutex g_pool_mu;
hipMemPool_t g_pools[64] = {};

hipError_t pool_of_current(hipMemPool_t* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (!g_pools[dev]) {
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    e = hipMemPoolCreate(&pool, &props);
    if (e != hipSuccess) return e;
    uint64_t keep = ~0ull;
    e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    if (e != hipSuccess) {
      (void)hipMemPoolDestroy(pool);
      return e;
    }
    g_pools[dev] = pool;
  }
  *out = g_pools[dev];
  return hipSuccess;
}

// A growable device buffer; allocation and release are ordered on the
// stream that uses it.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes, hipStream_t s) {
    if (bytes <= cap && ptr) return hipSuccess;
    if (ptr) {
      hipError_t e = phobic::dev_free(ptr, s);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      cap = 0;
    }
    const size_t want = bytes ? bytes : 256;
    hipError_t e = phobic::dev_alloc(&ptr, want, s);
    if (e == hipSuccess) cap = want;
    else ptr = nullptr;
    return e;
  }
  void release(hipStream_t s) {
    if (ptr) (void)phobic::dev_free(ptr, s);
    ptr = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

constexpr int kStages = 4;
constexpr int kGraphScratch = 8;  // context buffers a captured build reads or writes
const char* const kStageNames[kStages] = {"hash_route", "partition_split", "partition_offsets", "partition_search"};
constexpr uint32_t kMaxLevels = 4;
// The two pinned host chunks of a context, 64 MiB each.  Staged uploads
// move through them in transfers of one or two units of 8 MiB (two for
// copies of 512 MiB and more), filled by up to kStageThreads host threads.
// Measured host-to-device rates: 8 GiB at 54 GB/s with 16 MiB transfers
// (47 GB/s with 4 MiB), 24-96 MiB at 41-48 GB/s with 8 MiB transfers.
constexpr size_t kStageChunk = 64ull << 20;
constexpr size_t kStageUnit = 8ull << 20;
constexpr unsigned kStageSlots = 2 * kStageChunk / kStageUnit;
constexpr unsigned kStageThreads = 4;
static_assert(kStageSlots % (2 * kStageThreads) == 0, "every transfer position belongs to one filling thread");

}  // namespace

struct phobic_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // scratch owned by the context
  DevBuf pcount, lvl_count, scan_tot, rec_lo[2], rec_meta[2], fulcs, status;
  // tables of a group build (several key sets in one launch)
  DevBuf grp;
  // staging for host-buffer entry points
  DevBuf in_a, in_b, out_pilots, out_offsets, out_remap, out_roff;
  // two pinned host chunks for large host <-> device copies (allocated on
  // first use) and the events that say when a chunk's copy has finished
  void* pinned[2] = {};
  hipEvent_t pin_ev[2] = {};
  // staged uploads: the event after each 8 MiB unit's last DMA
  hipEvent_t slot_ev[kStageSlots] = {};
  hipEvent_t ev[kStages + 1] = {};
  hipEvent_t ev_in = nullptr;  // host-buffer builds: before the first input copy
  bool timed = false;
  float in_ms = 0, out_ms = 0;  // copies of the last host-buffer build
  phobic_tuning tuning{1, 12, kMaxFanout, 0, 12, 0, 1024};
  // fulcrums currently in ctx->fulcs
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
  uint32_t a_ppm;  // load factor in ppm, 1e6 = alpha 1
  uint32_t tcap;   // largest table size: table_size(cap, a_ppm)
  // partitioning levels: level l (1..L) bins cover G[l] partitions, G[L] = 1
  uint32_t L, fanout;
  uint64_t G[kMaxLevels + 1];
};

// Level-1 bins of a default two-level plan whose balanced fan-out would be
// larger: level 1 writes a run per bin per tile, and longer runs pay more
// there than the wider second level costs (1e9 keys, P 2048: 576 x 848 bins
// route and split in 9.38 ms against 9.94 ms for 699 x 699, four
// alternating processes each; 512 x 954 took 9.36 ms).
constexpr uint32_t kLevel1Bins = 576;

// Levels of the partitioning: the fewest L such that first * f^(L-1) >= np
// with f <= max_fanout, where first (the level-1 bin count) is at most
// first_fanout (0: max_fanout, all levels alike, except that a two-level
// plan keeps level 1 at kLevel1Bins when its balanced fan-out is larger).
// Level l's bins cover G[l] = f^(L-l) partitions.  False (plan untouched)
// when that needs more than kMaxLevels levels or a level-1 bin of more
// than 65535 partitions.
bool plan_levels(Plan* pl, uint32_t max_fanout = kMaxFanout, uint32_t first_fanout = 0) {
  const uint32_t np = pl->np;
  if (first_fanout == 0 && max_fanout > kLevel1Bins && np > kLevel1Bins * kLevel1Bins &&
      (uint64_t)kLevel1Bins * max_fanout >= np)
    first_fanout = kLevel1Bins;
  const uint32_t F1 = first_fanout ? std::min(first_fanout, max_fanout) : max_fanout;
  auto power = [](uint64_t b, uint32_t e) {
    uint64_t r = 1;
    for (uint32_t i = 0; i < e; i++) r *= b;
    return r;
  };
  uint32_t L = 1, f = np ? np : 1;
  if (np > F1) {
    L = 2;
    while ((double)F1 * std::pow((double)max_fanout, (double)(L - 1)) < (double)np) L++;
    if (L > kMaxLevels) return false;
    if (first_fanout) {  // smallest f with F1 * f^(L-1) >= np
      f = (uint32_t)std::ceil(std::pow((double)np / F1, 1.0 / (L - 1)));
      while (f > 2 && (uint64_t)F1 * power(f - 1, L - 1) >= np) f--;
      while ((uint64_t)F1 * power(f, L - 1) < np) f++;
    } else {  // smallest f with f^L >= np
      f = (uint32_t)std::ceil(std::pow((double)np, 1.0 / L));
      while (f > 2 && power(f - 1, L) >= np) f--;
      while (power(f, L) < np) f++;
    }
    if (f > max_fanout) return false;
  }
  uint64_t G[kMaxLevels + 1] = {};
  for (uint32_t l = 1; l <= L; l++) G[l] = power(f, L - l);
  if (G[1] > 65535) return false;
  pl->L = L;
  pl->fanout = f;
  for (uint32_t l = 0; l <= kMaxLevels; l++) pl->G[l] = G[l];
  return true;
}

int validate(const phobic_config* cfg, uint64_t n, Plan* pl) {
  if (!cfg || !(cfg->lambda > 0.0) || !std::isfinite(cfg->lambda)) return PHOBIC_E_INVALID_CONFIG;
  if (cfg->partition_size < 2 || cfg->partition_size > (1u << 20)) return PHOBIC_E_INVALID_CONFIG;
  if (n > 0xFFFFFFFFull) return PHOBIC_E_INVALID_CONFIG;  // offsets are u32 (mphf_builder.h:254)
  if ((cfg->buckets != 0) != (cfg->fulcrums != nullptr)) return PHOBIC_E_INVALID_CONFIG;
  const double Bd = cfg->buckets ? (double)cfg->buckets : std::ceil((double)cfg->partition_size / cfg->lambda);
  if (!(Bd >= 1.0) || Bd > 65535.0) return PHOBIC_E_INVALID_CONFIG;
  // a custom table maps a key to (v1 + v2) >> 16 with v1 + v2 below the
  // table's largest entry (bucket_of): entries up to B << 16 keep every
  // bucket index below B
  if (cfg->fulcrums)
    for (uint32_t i = 0; i < kFulcsTable; i++)
      if ((uint64_t)cfg->fulcrums[i] > ((uint64_t)Bd << 16)) return PHOBIC_E_INVALID_CONFIG;
  pl->n = n;
  pl->B = (uint32_t)Bd;
  pl->np = (uint32_t)((n + cfg->partition_size - 1) / cfg->partition_size);
  pl->cap = cfg->partition_size + cfg->partition_size / 2;  // mphf_config.h:61-64
  const double alpha = cfg->alpha == 0.0 ? 1.0 : cfg->alpha;
  if (!(alpha >= 0.5 && alpha <= 1.0)) return PHOBIC_E_INVALID_CONFIG;
  pl->a_ppm = (uint32_t)std::llround(alpha * 1e6);
  pl->tcap = table_size(pl->cap, pl->a_ppm);
  if (search_lds_bytes(pl->tcap, pl->B, 1) > 64 * 1024) return PHOBIC_E_INVALID_CONFIG;
  if (pl->np > (1u << 21)) return PHOBIC_E_INVALID_CONFIG;  // partition index inside a bin must fit 16 bits
  if (!plan_levels(pl)) return PHOBIC_E_INVALID_CONFIG;
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

// The table a build uses: the caller's (custom Bucketer) or OptBucketer's.
void config_fulcrums(const phobic_config* cfg, uint32_t* out) {
  if (cfg->fulcrums) {
    std::memcpy(out, cfg->fulcrums, kFulcsTable * 4);
  } else {
    compute_fulcrums(cfg->lambda, cfg->partition_size, out);
  }
}

// Uploads the configuration's fulcrums unless the context already holds them.
int upload_fulcrums(phobic_ctx* ctx, const phobic_config* cfg, hipStream_t s) {
  uint32_t f[kFulcsTable];
  config_fulcrums(cfg, f);
  if (ctx->fulcs.ptr && std::memcmp(f, ctx->f_host, sizeof(f)) == 0) return PHOBIC_OK;
  std::memcpy(ctx->f_host, f, sizeof(f));
  TRY_HIP(ctx->fulcs.ensure(kFulcsTable * 4, s));
  TRY_HIP(hipMemcpyAsync(ctx->fulcs.ptr, ctx->f_host, kFulcsTable * 4, hipMemcpyHostToDevice, s));
  TRY_HIP(hipStreamSynchronize(s));
  return PHOBIC_OK;
}

int set_device(phobic_ctx* ctx) {
  TRY_HIP(hipSetDevice(ctx->device));
  return PHOBIC_OK;
}

// Batched builds: per-set key ranges and partition bases on the device.
struct BatchInfo {
  uint32_t sets = 0;
  const uint64_t* set_key_off = nullptr;
  const uint32_t* set_part_base = nullptr;
  uint32_t* global_offsets = nullptr;  // np + 1 scratch for the per-set offsets
  // alpha < 1: the global remap offsets and each partition's remap start in
  // the output (np + 1 scratch each), and where set s's remap entries start
  // in the remap output (sets entries)
  uint32_t* global_roff = nullptr;
  uint32_t* part_remap = nullptr;
  const uint32_t* set_remap_base = nullptr;
  // sets of several configurations (null / 0: one configuration): per set
  // its XXH3 seed, fulcrum table, buckets per partition and first pilot
  const uint64_t* set_seed = nullptr;
  const uint32_t* set_ftab = nullptr;
  const uint32_t* ftabs = nullptr;
  uint32_t ntabs = 0;
  const uint32_t* set_B = nullptr;
  const uint64_t* set_pilot_base = nullptr;
  // u64 sets read where they lie (null: concatenated at src.u64)
  const uint64_t* const* set_keys = nullptr;
};

// Layout of the context scratch for a plan.
struct Layout {
  uint32_t bins1;
  uint64_t slots;
  uint64_t cnt_off[kMaxLevels + 1];
  uint64_t cnt_total;
};

Layout layout_of(const Plan& pl) {
  Layout ly{};
  ly.bins1 = (uint32_t)((pl.np + pl.G[1] - 1) / pl.G[1]);
  ly.slots = (uint64_t)ly.bins1 * pl.G[1] * pl.cap;
  // per-level bin counters: level l has ceil(np / G[l-1]) * f bins (level 1: bins1)
  for (uint32_t l = 1; l < pl.L; l++) {
    ly.cnt_off[l] = ly.cnt_total;
    ly.cnt_total += l == 1 ? ly.bins1 : ((pl.np + pl.G[l - 1] - 1) / pl.G[l - 1]) * (uint64_t)pl.fanout;
  }
  return ly;
}

// The plan with the context's tuning applied (levels for its fan-out).
Plan tuned(const phobic_ctx* ctx, Plan pl) {
  if (!plan_levels(&pl, ctx->tuning.max_fanout, ctx->tuning.first_fanout)) plan_levels(&pl, kMaxFanout);
  return pl;
}

// Partitions per search workgroup: the tuned value, reduced until the
// workgroup's LDS fits 64 KiB.
uint32_t search_waves(const phobic_ctx* ctx, const Plan& pl) {
  uint32_t w = ctx->tuning.search_waves;
  while (w > 1 && search_lds_bytes(pl.tcap, pl.B, w) > 64 * 1024) w /= 2;
  return w;
}

// Device scratch of one build: partitioning records (two ping-pong
// buffers), bin counters, scan totals and the fulcrum table.
struct Scratch {
  uint32_t* pcount;
  uint32_t* lvl_count;
  uint32_t* scan_tot;
  uint64_t* rec_lo[2];
  uint32_t* rec_meta[2];
  uint32_t* fulcs;
};

// Bytes of each scratch region for a (tuned) plan, in Scratch's order.
struct ScratchSizes {
  uint64_t pcount, lvl_count, scan_tot, rec_lo, rec_meta, fulcs;
};
ScratchSizes scratch_sizes(const Plan& pl) {
  const Layout ly = layout_of(pl);
  ScratchSizes z;
  z.pcount = ((uint64_t)pl.np + pl.fanout) * 4;
  z.lvl_count = (ly.cnt_total + 1) * 4;
  z.scan_tot = (2ull * partition_scan_tiles(pl.np) + 2) * 4;
  z.rec_lo = ly.slots * 8;
  z.rec_meta = ly.slots * 4;
  z.fulcs = kFulcsTable * 4;
  return z;
}

Scratch ctx_scratch(const phobic_ctx* ctx) {
  Scratch sc;
  sc.pcount = ctx->pcount.as<uint32_t>();
  sc.lvl_count = ctx->lvl_count.as<uint32_t>();
  sc.scan_tot = ctx->scan_tot.as<uint32_t>();
  for (int i = 0; i < 2; i++) {
    sc.rec_lo[i] = ctx->rec_lo[i].as<uint64_t>();
    sc.rec_meta[i] = ctx->rec_meta[i].as<uint32_t>();
  }
  sc.fulcs = ctx->fulcs.as<uint32_t>();
  return sc;
}

// Allocations and the fulcrum upload: everything a build needs before its
// kernels are enqueued (not capturable into a graph).
int prepare_build(phobic_ctx* ctx, const phobic_config* cfg, const Plan& pl0, hipStream_t s) {
  const Plan pl = tuned(ctx, pl0);
  TRY(upload_fulcrums(ctx, cfg, s));
  const ScratchSizes z = scratch_sizes(pl);
  TRY_HIP(ctx->pcount.ensure(z.pcount, s));
  TRY_HIP(ctx->lvl_count.ensure(z.lvl_count, s));
  TRY_HIP(ctx->scan_tot.ensure(z.scan_tot, s));
  for (int i = 0; i < 2; i++) {
    TRY_HIP(ctx->rec_lo[i].ensure(z.rec_lo, s));
    TRY_HIP(ctx->rec_meta[i].ensure(z.rec_meta, s));
  }
  return PHOBIC_OK;
}

// Kernel launches of one build (fills included, no allocation, no host
// copies): capturable into a hipGraph when `timed` is false.
int launch_build(phobic_ctx* ctx, const Scratch& sc, int kind, const KeySource& src, const phobic_config* cfg,
                 const Plan& pl0, const phobic_outputs& d_out, uint32_t* d_status, hipStream_t s,
                 const BatchInfo& batch, bool timed) {
  const Plan pl = tuned(ctx, pl0);
  const uint32_t per = ctx->tuning.route_per;
  const uint32_t split_per = ctx->tuning.split_per ? ctx->tuning.split_per : per;
  // plain u64 builds carry the raw key through every level but the last;
  // the others carry hashed records; the last level writes what the search
  // reads (lower64 + bucket16)
  const bool raw = kind == kKindU64 && !batch.sets;
  auto fmt = [&](uint32_t l) { return l == pl.L ? kFmtFinal : raw ? kFmtRaw : kFmtRec; };
  uint32_t* d_pilots = d_out.pilots;
  const bool remap = pl.a_ppm < 1000000u;
  const uint32_t np = pl.np, L = pl.L, f = pl.fanout;
  const Layout ly = layout_of(pl);
  const uint32_t bins1 = ly.bins1;
  TRY_HIP(launch_fill_u32(d_status, 1, 0, s));
  TRY_HIP(launch_fill_u32(sc.pcount, (uint64_t)np + f, 0, s));
  TRY_HIP(launch_fill_u32(sc.lvl_count, ly.cnt_total, 0, s));
  if (timed) TRY_HIP(hipEventRecord(ctx->ev[0], s));

  // level l writes buffer 0 when L - l is even, so level L lands in buffer 0
  auto out_buf = [&](uint32_t l) { return (int)((L - l) & 1); };
  auto counts = [&](uint32_t l) {
    return l == L ? sc.pcount : sc.lvl_count + ly.cnt_off[l];
  };
  HashRouteArgs ha{};
  ha.src = src;
  ha.n = pl.n;
  ha.seed = cfg->seed;
  ha.np = np;
  ha.cap = pl.cap;
  ha.G = (uint32_t)pl.G[1];
  ha.fanout = bins1;
  ha.fulcs = sc.fulcs;
  ha.child_count = counts(1);
  ha.out_lo = sc.rec_lo[out_buf(1)];
  ha.out_meta = sc.rec_meta[out_buf(1)];
  ha.status = d_status;
  ha.sets = batch.sets;
  ha.set_key_off = batch.set_key_off;
  ha.set_part_base = batch.set_part_base;
  ha.set_keys = batch.set_keys;
  ha.set_seed = batch.set_seed;
  ha.set_ftab = batch.set_ftab;
  ha.ftabs = batch.ftabs;
  ha.ntabs = batch.ntabs;
  const uint32_t threads = ctx->tuning.route_threads ? ctx->tuning.route_threads : 512;
  TRY_HIP(launch_hash_route(kind, fmt(1), ha, threads, per, s));
  if (timed) TRY_HIP(hipEventRecord(ctx->ev[1], s));

  for (uint32_t l = 2; l <= L; l++) {
    SplitArgs sp{};
    const uint32_t parents = (uint32_t)((np + pl.G[l - 1] - 1) / pl.G[l - 1]);
    sp.in_lo = sc.rec_lo[out_buf(l - 1)];
    sp.in_meta = sc.rec_meta[out_buf(l - 1)];
    sp.parent_count = counts(l - 1);
    sp.parent_cap = pl.G[l - 1] * pl.cap;
    const uint32_t sper = split_tile_per(threads, split_per, fmt(l - 1), fmt(l), f);
    const uint64_t tile = (uint64_t)threads * sper;
    sp.slices = (uint32_t)((sp.parent_cap + tile - 1) / tile);
    sp.fanout = f;
    sp.G_parent = (uint32_t)pl.G[l - 1];
    sp.G_child = (uint32_t)pl.G[l];
    sp.cap = pl.cap;
    sp.np = np;
    sp.seed = cfg->seed;
    sp.fulcs = sc.fulcs;
    sp.child_count = counts(l);
    sp.out_lo = sc.rec_lo[out_buf(l)];
    sp.out_meta = sc.rec_meta[out_buf(l)];
    sp.status = d_status;
    TRY_HIP(launch_split_level(fmt(l - 1), fmt(l), sp, parents, threads, sper, s));
  }
  if (timed) TRY_HIP(hipEventRecord(ctx->ev[2], s));

  uint32_t* totals = sc.scan_tot;
  if (batch.sets) {  // global scans, then every set's own offsets
    TRY_HIP(launch_partition_offsets(sc.pcount, np, remap ? pl.a_ppm : 0, batch.global_offsets,
                                     remap ? batch.global_roff : nullptr, totals, s));
    TRY_HIP(launch_batch_offsets(batch.global_offsets, batch.set_part_base, batch.sets, np + batch.sets,
                                 d_out.offsets, s));
    if (remap) {
      TRY_HIP(launch_batch_offsets(batch.global_roff, batch.set_part_base, batch.sets, np + batch.sets,
                                   d_out.remap_offsets, s));
      TRY_HIP(launch_batch_remap_bases(batch.global_roff, batch.set_part_base, batch.set_remap_base, batch.sets, np,
                                       batch.part_remap, s));
    }
  } else {
    TRY_HIP(launch_partition_offsets(sc.pcount, np, remap ? pl.a_ppm : 0, d_out.offsets,
                                     remap ? d_out.remap_offsets : nullptr, totals, s));
  }
  if (timed) TRY_HIP(hipEventRecord(ctx->ev[3], s));

  SearchArgs ra{};
  ra.rec_lower = sc.rec_lo[0];
  ra.rec_bucket = reinterpret_cast<const uint16_t*>(sc.rec_meta[0]);
  ra.pcount = sc.pcount;
  ra.sorted = sc.rec_lo[1];  // free once partitioning is done
  ra.np = np;
  ra.B = pl.B;
  ra.cap = pl.cap;
  ra.a_ppm = pl.a_ppm;
  ra.tcap = pl.tcap;
  ra.pilots = d_pilots;
  ra.remap = remap ? d_out.remap : nullptr;
  ra.remap_offsets = remap ? (batch.sets ? batch.part_remap : d_out.remap_offsets) : nullptr;
  ra.status = d_status;
  ra.sets = batch.sets;
  ra.set_part_base = batch.set_part_base;
  ra.set_B = batch.set_B;
  ra.set_pilot_base = batch.set_pilot_base;
  ra.mode = ctx->tuning.search_mode;
  TRY_HIP(launch_partition_search(ra, search_waves(ctx, pl), s));
  if (timed) {
    TRY_HIP(hipEventRecord(ctx->ev[4], s));
    ctx->timed = true;
  }
  return PHOBIC_OK;
}

// Enqueue one build.  Inputs/outputs are device pointers.
int enqueue_build(phobic_ctx* ctx, int kind, const KeySource& src, const phobic_config* cfg, const Plan& pl,
                  const phobic_outputs& d_out, uint32_t* d_status, hipStream_t s) {
  TRY(prepare_build(ctx, cfg, pl, s));
  return launch_build(ctx, ctx_scratch(ctx), kind, src, cfg, pl, d_out, d_status, s, BatchInfo{}, true);
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
  t->input_transfer_ms = ctx->in_ms;
  t->result_transfer_ms = ctx->out_ms;
  return PHOBIC_OK;
}

uint64_t remap_capacity(const Plan& pl) {
  if (pl.a_ppm >= 1000000u) return 0;
  return (pl.n * (1000000u - pl.a_ppm) + pl.a_ppm - 1) / pl.a_ppm + pl.np;
}

bool outputs_ok(const phobic_outputs* out, const Plan& pl) {
  if (!out || !out->pilots || !out->offsets) return false;
  if (pl.a_ppm < 1000000u && (!out->remap || !out->remap_offsets)) return false;
  return true;
}

// Host-buffer build: stage inputs, enqueue on the context stream, copy out.
// Host <-> device copies of the host-buffer entry points.  Pageable memory
// moves at about half the link rate, so copies of at least kStageMin bytes
// go through the context's pinned chunks.  Uploads (stage_to_device) fill
// 4 MiB slots on several threads while earlier slots' DMAs run; downloads
// alternate between the two 64 MiB chunks, one chunk's DMA running while
// the other is copied out by kCopyThreads host threads.
constexpr size_t kStageMin = 4ull << 20;
constexpr unsigned kCopyThreads = 4;

// fill over [off, off + n) in kCopyThreads slices (multiples of 64 bytes),
// slice t writing dst + t * per; each slice runs on a thread of its own, and
// slices whose thread cannot be started (no exception may leave the C ABI)
// run here.
void par_fill(char* dst, uint64_t off, uint64_t n, const phobic::StageFill& fill) {
  const uint64_t per = ((n + kCopyThreads - 1) / kCopyThreads + 63) & ~uint64_t(63);
  auto slice = [&](unsigned t) {
    const uint64_t lo = t * per;
    if (lo < n) fill(dst + lo, off + lo, std::min(per, n - lo));
  };
  std::vector<std::thread> th;
  unsigned t = 1;
  try {
    th.reserve(kCopyThreads - 1);
    for (; t < kCopyThreads && t * per < n; t++) th.emplace_back(slice, t);
  } catch (...) {
  }
  for (unsigned u = t; u < kCopyThreads; u++) slice(u);
  slice(0);
  for (auto& x : th) x.join();
}

void par_memcpy(void* dst, const void* src, size_t n) {
  par_fill(static_cast<char*>(dst), 0, n,
           [src](char* d, uint64_t lo, uint64_t len) { std::memcpy(d, static_cast<const char*>(src) + lo, len); });
}

int ensure_pinned(phobic_ctx* ctx) {
  for (int i = 0; i < 2; i++) {
    if (!ctx->pinned[i]) {
      TRY_HIP(hipHostMalloc(&ctx->pinned[i], kStageChunk, hipHostMallocDefault));
      // touched once now: the first write of each page is slow, and would
      // otherwise fall on the first copies through it
      par_fill(static_cast<char*>(ctx->pinned[i]), 0, kStageChunk,
               [](char* d, uint64_t, uint64_t len) { std::memset(d, 0, len); });
    }
    if (!ctx->pin_ev[i]) TRY_HIP(hipEventCreateWithFlags(&ctx->pin_ev[i], hipEventDisableTiming));
  }
  for (unsigned i = 0; i < kStageSlots; i++)
    if (!ctx->slot_ev[i]) TRY_HIP(hipEventCreateWithFlags(&ctx->slot_ev[i], hipEventDisableTiming));
  return PHOBIC_OK;
}

// dst (device) <- src (host), ordered on s (src stays valid until the
// caller's final synchronisation)
int copy_to_device(phobic_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < kStageMin) {
    TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return PHOBIC_OK;
  }
  return phobic::stage_to_device(
      ctx, dst, bytes,
      [src](char* d, uint64_t lo, uint64_t len) { std::memcpy(d, static_cast<const char*>(src) + lo, len); }, s);
}

// dst (host) <- src (device) after the work enqueued on s; synchronous
int copy_to_host(phobic_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < kStageMin) {
    TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    TRY_HIP(hipStreamSynchronize(s));
    return PHOBIC_OK;
  }
  TRY(ensure_pinned(ctx));
  const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t i) -> int {
    const size_t off = i * kStageChunk;
    TRY_HIP(hipMemcpyAsync(ctx->pinned[i & 1], (const char*)src + off, std::min(kStageChunk, bytes - off),
                           hipMemcpyDeviceToHost, s));
    TRY_HIP(hipEventRecord(ctx->pin_ev[i & 1], s));
    return PHOBIC_OK;
  };
  // the chunks' earlier uploads and downloads are done
  for (unsigned i = 0; i < kStageSlots; i++) TRY_HIP(hipEventSynchronize(ctx->slot_ev[i]));
  TRY_HIP(hipEventSynchronize(ctx->pin_ev[0]));
  TRY_HIP(hipEventSynchronize(ctx->pin_ev[1]));
  TRY(issue(0));
  for (size_t i = 0; i < chunks; i++) {
    // chunk i + 1 goes to the other pinned chunk, whose last reader (the
    // copy-out of chunk i - 1) has finished
    if (i + 1 < chunks) TRY(issue(i + 1));
    TRY_HIP(hipEventSynchronize(ctx->pin_ev[i & 1]));
    const size_t off = i * kStageChunk;
    par_memcpy((char*)dst + off, ctx->pinned[i & 1], std::min(kStageChunk, bytes - off));
  }
  return PHOBIC_OK;
}

int host_build(phobic_ctx* ctx, int kind, const void* a, size_t a_bytes, const void* b, size_t b_bytes,
               uint64_t n, const phobic_config* cfg, const phobic_outputs* out, phobic_stage_times* times) {
  if (!ctx) return PHOBIC_E_INVALID_CONFIG;
  Plan pl;
  TRY(validate(cfg, n, &pl));
  if (!outputs_ok(out, pl)) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  TRY_HIP(ctx->in_a.ensure(a_bytes, s));
  TRY_HIP(hipEventRecord(ctx->ev_in, s));
  if (a_bytes) TRY(copy_to_device(ctx, ctx->in_a.ptr, a, a_bytes, s));
  if (b) {
    TRY_HIP(ctx->in_b.ensure(b_bytes, s));
    TRY(copy_to_device(ctx, ctx->in_b.ptr, b, b_bytes, s));
  }
  const uint64_t NB = (uint64_t)pl.np * pl.B;
  const bool remap = pl.a_ppm < 1000000u;
  TRY_HIP(ctx->out_pilots.ensure(NB * 4 + 4, s));  // + status word
  TRY_HIP(ctx->out_offsets.ensure(((uint64_t)pl.np + 1) * 4, s));
  if (remap) {
    TRY_HIP(ctx->out_remap.ensure(remap_capacity(pl) * 4, s));
    TRY_HIP(ctx->out_roff.ensure(((uint64_t)pl.np + 1) * 4, s));
  }
  uint32_t* d_status = ctx->out_pilots.as<uint32_t>() + NB;
  KeySource src{};
  if (kind == kKindU64) src.u64 = ctx->in_a.as<uint64_t>();
  if (kind == kKindBytes) {
    src.bytes = ctx->in_a.as<uint8_t>();
    src.offs = ctx->in_b.as<uint64_t>();
  }
  if (kind == kKindKey) src.keys = ctx->in_a.as<uint4>();
  phobic_outputs d_out{ctx->out_pilots.as<uint32_t>(), ctx->out_offsets.as<uint32_t>(),
                       remap ? ctx->out_remap.as<uint32_t>() : nullptr, remap ? ctx->out_roff.as<uint32_t>() : nullptr};
  TRY(enqueue_build(ctx, kind, src, cfg, pl, d_out, d_status, s));
  // the copies back cannot start before the last kernel ends: waiting for
  // it first separates the result transfer from the device time
  TRY_HIP(hipEventSynchronize(ctx->ev[kStages]));
  const auto t_out = std::chrono::steady_clock::now();
  uint32_t bits = 0;
  TRY_HIP(hipMemcpyAsync(&bits, d_status, 4, hipMemcpyDeviceToHost, s));
  if (NB) TRY(copy_to_host(ctx, out->pilots, ctx->out_pilots.ptr, NB * 4, s));
  TRY(copy_to_host(ctx, out->offsets, ctx->out_offsets.ptr, ((uint64_t)pl.np + 1) * 4, s));
  if (remap) TRY(copy_to_host(ctx, out->remap_offsets, ctx->out_roff.ptr, ((uint64_t)pl.np + 1) * 4, s));
  TRY_HIP(hipStreamSynchronize(s));
  if (remap && out->remap_offsets[pl.np])
    TRY(copy_to_host(ctx, out->remap, ctx->out_remap.ptr, (uint64_t)out->remap_offsets[pl.np] * 4, s));
  ctx->out_ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_out).count();
  TRY_HIP(hipEventElapsedTime(&ctx->in_ms, ctx->ev_in, ctx->ev[0]));
  TRY(fill_times(ctx, times));
  return status_from_bits(bits);
}

int device_build(phobic_ctx* ctx, int kind, const KeySource& src, uint64_t n, const phobic_config* cfg,
                 const phobic_outputs* d_out, uint32_t* d_status, void* stream) {
  if (!ctx || !d_status) return PHOBIC_E_INVALID_CONFIG;
  Plan pl;
  TRY(validate(cfg, n, &pl));
  if (!outputs_ok(d_out, pl)) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  ctx->in_ms = ctx->out_ms = 0;
  return enqueue_build(ctx, kind, src, cfg, pl, *d_out, d_status, (hipStream_t)stream);
}

}  // namespace

hipStream_t phobic::ctx_stream(phobic_ctx* ctx) { return ctx->stream; }

namespace {

constexpr uint64_t kPoolGrain = 2ull << 20;  // allocations are counted in whole 2 MiB units
uint64_t grain(uint64_t b) { return (b + kPoolGrain - 1) / kPoolGrain * kPoolGrain; }

// Byte offsets of a group build's tables inside ctx->grp.
struct GroupTables {
  uint64_t key_off, part_base, remap_base, goff, groff, part_remap;
  uint64_t seed, setB, ftab, pilot_base, ftabs;  // sets of several configurations
  uint64_t keys;                                 // u64 sets read in place
  uint64_t bytes;
};
GroupTables group_tables(uint32_t sets, uint32_t np, uint32_t ntabs, bool in_place) {
  GroupTables t;
  uint64_t at = 0;
  auto take = [&at](uint64_t b) {
    const uint64_t r = at;
    at += (b + 255) & ~uint64_t(255);
    return r;
  };
  t.key_off = take(8ull * (sets + 1));
  t.part_base = take(4ull * (sets + 1));
  t.remap_base = take(4ull * (sets + 1));
  t.goff = take(4ull * (np + 1));
  t.groff = take(4ull * (np + 1));
  t.part_remap = take(4ull * (np + 1));
  const bool mixed = ntabs != 0;
  t.seed = take(mixed ? 8ull * sets : 0);
  t.setB = take(mixed ? 4ull * sets : 0);
  t.ftab = take(mixed ? 4ull * sets : 0);
  t.pilot_base = take(mixed ? 8ull * (sets + 1) : 0);
  t.ftabs = take(4ull * kFulcsTable * ntabs);
  t.keys = take(in_place ? 8ull * sets : 0);
  t.bytes = at;
  return t;
}

// The flattened plan of a group: np = the sets' partitions together.
int group_plan(const phobic_config* cfg, const phobic::GroupLayout& gl, Plan* pl) {
  TRY(validate(cfg, 0, pl));
  if (gl.n > 0xFFFFFFFFull || gl.np > (1u << 21)) return PHOBIC_E_INVALID_CONFIG;
  pl->n = gl.n;
  pl->np = gl.np;
  if (gl.B) pl->B = gl.B;  // the largest of the sets' (the search's LDS layout)
  if (!plan_levels(pl)) return PHOBIC_E_INVALID_CONFIG;
  return PHOBIC_OK;
}

// Bytes each scratch buffer of the context must hold for a plan, in the
// order of scratch_bufs.
void scratch_needs(const phobic_ctx* ctx, const Plan& pl0, uint64_t grp_bytes, uint64_t (&need)[9]) {
  const ScratchSizes z = scratch_sizes(tuned(ctx, pl0));
  const uint64_t v[9] = {z.pcount, z.lvl_count, z.scan_tot, z.rec_lo, z.rec_lo, z.rec_meta, z.rec_meta, z.fulcs,
                         grp_bytes};
  for (int i = 0; i < 9; i++) need[i] = v[i];
}
void scratch_bufs(const phobic_ctx* ctx, const DevBuf* (&b)[9]) {
  const DevBuf* v[9] = {&ctx->pcount, &ctx->lvl_count, &ctx->scan_tot, &ctx->rec_lo[0], &ctx->rec_lo[1],
                        &ctx->rec_meta[0], &ctx->rec_meta[1], &ctx->fulcs, &ctx->grp};
  for (int i = 0; i < 9; i++) b[i] = v[i];
}

// from_empty: as if the context held no scratch yet
uint64_t growth(const phobic_ctx* ctx, const Plan& pl, uint64_t grp_bytes, bool from_empty) {
  uint64_t need[9];
  scratch_needs(ctx, pl, grp_bytes, need);
  const DevBuf* b[9];
  scratch_bufs(ctx, b);
  uint64_t g = 0;
  for (int i = 0; i < 9; i++) {
    const uint64_t cap = from_empty ? 0 : b[i]->cap;
    if (need[i] > cap) g += grain(need[i]) - grain(cap);
  }
  return g;
}

}  // namespace

int phobic::group_layout(const phobic_config* const* cfgs, const uint64_t* set_sizes, uint32_t sets,
                         GroupLayout* gl) {
  if (!cfgs || !gl || (sets && (!set_sizes || !cfgs[0]))) return PHOBIC_E_INVALID_CONFIG;
  const phobic_config* c0 = sets ? cfgs[0] : nullptr;
  if (!c0) return PHOBIC_E_INVALID_CONFIG;
  Plan one;
  TRY(validate(c0, 0, &one));
  auto alpha_of = [](const phobic_config* c) { return c->alpha == 0.0 ? 1.0 : c->alpha; };
  bool mixed = false;
  for (uint32_t s = 0; s < sets; s++) {
    const phobic_config* c = cfgs[s];
    if (!c || c->partition_size != c0->partition_size || alpha_of(c) != alpha_of(c0) ||
        c->buckets != c0->buckets || c->fulcrums != c0->fulcrums)
      return PHOBIC_E_INVALID_CONFIG;
    if (c->lambda != c0->lambda || c->seed != c0->seed) mixed = true;
  }
  if (mixed && c0->fulcrums) return PHOBIC_E_INVALID_CONFIG;  // one custom table for all
  try {
    gl->key_off.assign(sets + 1, 0);
    gl->part_base.assign(sets + 1, 0);
    gl->remap_base.assign(sets + 1, 0);
    gl->pilot_base.assign(sets + 1, 0);
    gl->seed.clear();
    gl->setB.clear();
    gl->ftab.clear();
    gl->ftabs.clear();
    if (mixed) {
      gl->seed.resize(sets);
      gl->setB.resize(sets);
      gl->ftab.resize(sets);
    }
  } catch (...) {
    return PHOBIC_E_OOM;
  }
  gl->sets = sets;
  gl->mixed = mixed;
  gl->ntabs = 0;
  gl->set_keys.clear();
  uint32_t Bmax = 0;
  std::vector<double> lambdas;  // of the fulcrum tables, in table order
  uint64_t parts = 0, rem = 0, pil = 0;
  for (uint32_t s = 0; s < sets; s++) {
    Plan ps;
    TRY(validate(cfgs[s], set_sizes[s], &ps));
    gl->key_off[s + 1] = gl->key_off[s] + set_sizes[s];
    parts += ps.np;
    rem += remap_capacity(ps);
    pil += (uint64_t)ps.np * ps.B;
    if (parts > (1u << 21) || rem > 0xFFFFFFFFull || gl->key_off[s + 1] > 0xFFFFFFFFull)
      return PHOBIC_E_INVALID_CONFIG;
    gl->part_base[s + 1] = (uint32_t)parts;
    gl->remap_base[s + 1] = (uint32_t)rem;
    gl->pilot_base[s + 1] = pil;
    Bmax = std::max(Bmax, ps.B);
    if (mixed) {
      uint32_t t = 0;
      while (t < lambdas.size() && lambdas[t] != cfgs[s]->lambda) t++;
      if (t == lambdas.size()) {
        if (t == kMaxFulcTabs) return PHOBIC_E_INVALID_CONFIG;
        uint32_t f[kFulcsTable];
        config_fulcrums(cfgs[s], f);
        try {
          lambdas.push_back(cfgs[s]->lambda);
          gl->ftabs.insert(gl->ftabs.end(), f, f + kFulcsTable);
        } catch (...) {
          return PHOBIC_E_OOM;
        }
      }
      gl->seed[s] = cfgs[s]->seed;
      gl->setB[s] = ps.B;
      gl->ftab[s] = t;
    }
  }
  gl->ntabs = (uint32_t)lambdas.size();
  gl->np = (uint32_t)parts;
  gl->n = gl->key_off[sets];
  gl->B = Bmax;
  gl->a_ppm = one.a_ppm;
  return PHOBIC_OK;
}

uint64_t phobic::group_scratch_growth(const phobic_ctx* ctx, const phobic_config* cfg, const GroupLayout& gl,
                                      bool from_empty) {
  Plan pl;
  if (group_plan(cfg, gl, &pl) != PHOBIC_OK) return 0;
  return growth(ctx, pl, group_tables(gl.sets, gl.np, gl.ntabs, !gl.set_keys.empty()).bytes, from_empty);
}

uint64_t phobic::build_scratch_growth(const phobic_ctx* ctx, const phobic_config* cfg, uint64_t n, bool from_empty) {
  Plan pl;
  if (validate(cfg, n, &pl) != PHOBIC_OK) return 0;
  return growth(ctx, pl, 0, from_empty);
}

// Transfer i of the staged upload (U units of 8 MiB) goes through ring
// position i mod (kStageSlots / U) and is filled and sent by worker i mod T.
// T divides the number of positions, so every position belongs to one
// worker, which waits for its units' previous DMAs before refilling them.
// The workers enqueue their copies on s concurrently; each copy has a
// destination of its own.
int phobic::stage_to_device(phobic_ctx* ctx, void* dst, uint64_t bytes, const StageFill& fill, hipStream_t s) {
  TRY(ensure_pinned(ctx));
  const unsigned U = bytes >= (512ull << 20) ? 2 : 1;
  const uint64_t tsize = U * kStageUnit, count = (bytes + tsize - 1) / tsize;
  const unsigned positions = kStageSlots / U;
  unsigned T = 1;
  while (T * 2 <= kStageThreads && T * 2 <= count) T *= 2;
  std::atomic<int> err{PHOBIC_OK};
  auto work = [&](unsigned t) {
    for (uint64_t i = t; i < count && err.load(std::memory_order_relaxed) == PHOBIC_OK; i += T) {
      const unsigned u0 = (unsigned)(i % positions) * U;  // first unit of the position
      char* buf = static_cast<char*>(ctx->pinned[u0 / (kStageSlots / 2)]) + (u0 % (kStageSlots / 2)) * kStageUnit;
      const uint64_t off = i * tsize, len = std::min<uint64_t>(tsize, bytes - off);
      hipError_t e = hipSuccess;
      for (unsigned u = 0; u < U && e == hipSuccess; u++) e = hipEventSynchronize(ctx->slot_ev[u0 + u]);
      if (e == hipSuccess) {
        fill(buf, off, len);
        e = hipMemcpyAsync(static_cast<char*>(dst) + off, buf, len, hipMemcpyHostToDevice, s);
      }
      for (unsigned u = 0; u < U && e == hipSuccess; u++) e = hipEventRecord(ctx->slot_ev[u0 + u], s);
      if (e != hipSuccess) {
        int want = PHOBIC_OK;
        err.compare_exchange_strong(want, hip_status(e));
      }
    }
  };
  // workers that cannot be started (no exception may leave the C ABI) run
  // here, one after another, after this thread's own share
  std::vector<std::thread> th;
  unsigned started = 1;
  try {
    th.reserve(T - 1);
    for (; started < T; started++) th.emplace_back(work, started);
  } catch (...) {
  }
  work(0);
  for (unsigned t = started; t < T; t++) work(t);
  for (auto& x : th) x.join();
  return err.load();
}

int phobic::ctx_prepare_staging(phobic_ctx* ctx) {
  TRY(set_device(ctx));
  return ensure_pinned(ctx);
}

void phobic::ctx_trim(phobic_ctx* ctx) {
  (void)hipSetDevice(ctx->device);
  for (DevBuf* b : {&ctx->pcount, &ctx->lvl_count, &ctx->scan_tot, &ctx->rec_lo[0], &ctx->rec_lo[1], &ctx->rec_meta[0],
                    &ctx->rec_meta[1], &ctx->fulcs, &ctx->grp, &ctx->status, &ctx->in_a, &ctx->in_b, &ctx->out_pilots,
                    &ctx->out_offsets, &ctx->out_remap, &ctx->out_roff})
    b->release(ctx->stream);
  std::memset(ctx->f_host, 0, sizeof(ctx->f_host));
  (void)hipStreamSynchronize(ctx->stream);
}

uint64_t phobic::ctx_scratch_bytes(const phobic_ctx* ctx) {
  const DevBuf* b[9];
  scratch_bufs(ctx, b);
  uint64_t t = 0;
  for (const DevBuf* x : b) t += grain(x->cap);
  for (const DevBuf* x : {&ctx->status, &ctx->in_a, &ctx->in_b, &ctx->out_pilots, &ctx->out_offsets, &ctx->out_remap,
                          &ctx->out_roff})
    t += grain(x->cap);
  return t;
}

uint64_t phobic::pool_bytes(uint64_t b) { return grain(b); }

uint64_t phobic::pool_used_bytes() {
  hipMemPool_t pool = nullptr;
  if (pool_of_current(&pool) != hipSuccess) return 0;
  uint64_t v = 0;
  if (hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &v) != hipSuccess) return 0;
  return v;
}

int phobic::group_build(phobic_ctx* ctx, const phobic_config* cfg, int kind, const void* d_keys,
                        const uint64_t* d_offs, const GroupLayout& gl, const phobic_outputs& block, uint32_t* d_status,
                        hipStream_t s) {
  if (!ctx || !d_status || gl.sets == 0 || (kind != kKindU64 && kind != kKindBytes)) return PHOBIC_E_INVALID_CONFIG;
  const bool in_place = !gl.set_keys.empty();
  if (in_place) {  // u64 sets read where they lie
    if (kind != kKindU64 || gl.set_keys.size() != gl.sets) return PHOBIC_E_INVALID_CONFIG;
    for (uint32_t x = 0; x < gl.sets; x++)
      if (gl.key_off[x + 1] > gl.key_off[x] && !gl.set_keys[x]) return PHOBIC_E_INVALID_CONFIG;
  } else if (gl.n && (!d_keys || (kind == kKindBytes && !d_offs))) {
    return PHOBIC_E_INVALID_CONFIG;
  }
  Plan pl;
  TRY(group_plan(cfg, gl, &pl));
  const bool remap = pl.a_ppm < 1000000u;
  if (!block.pilots || !block.offsets || (remap && (!block.remap || !block.remap_offsets))) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  const GroupTables t = group_tables(gl.sets, gl.np, gl.ntabs, !gl.set_keys.empty());
  TRY_HIP(ctx->grp.ensure(t.bytes, s));
  char* g = ctx->grp.as<char>();
  // from pageable memory: each copy has read its source when it returns
  TRY_HIP(hipMemcpyAsync(g + t.key_off, gl.key_off.data(), 8ull * (gl.sets + 1), hipMemcpyHostToDevice, s));
  TRY_HIP(hipMemcpyAsync(g + t.part_base, gl.part_base.data(), 4ull * (gl.sets + 1), hipMemcpyHostToDevice, s));
  TRY_HIP(hipMemcpyAsync(g + t.remap_base, gl.remap_base.data(), 4ull * (gl.sets + 1), hipMemcpyHostToDevice, s));
  if (gl.mixed) {
    if (gl.seed.size() != gl.sets || gl.setB.size() != gl.sets || gl.ftab.size() != gl.sets ||
        gl.pilot_base.size() != gl.sets + 1 || gl.ftabs.size() != (size_t)gl.ntabs * kFulcsTable || !gl.ntabs)
      return PHOBIC_E_INVALID_CONFIG;
    for (uint32_t x = 0; x < gl.sets; x++)
      if (gl.ftab[x] >= gl.ntabs || gl.setB[x] > gl.B) return PHOBIC_E_INVALID_CONFIG;
    TRY_HIP(hipMemcpyAsync(g + t.seed, gl.seed.data(), 8ull * gl.sets, hipMemcpyHostToDevice, s));
    TRY_HIP(hipMemcpyAsync(g + t.setB, gl.setB.data(), 4ull * gl.sets, hipMemcpyHostToDevice, s));
    TRY_HIP(hipMemcpyAsync(g + t.ftab, gl.ftab.data(), 4ull * gl.sets, hipMemcpyHostToDevice, s));
    TRY_HIP(hipMemcpyAsync(g + t.pilot_base, gl.pilot_base.data(), 8ull * (gl.sets + 1), hipMemcpyHostToDevice, s));
    TRY_HIP(hipMemcpyAsync(g + t.ftabs, gl.ftabs.data(), 4ull * gl.ftabs.size(), hipMemcpyHostToDevice, s));
  }
  if (in_place)
    TRY_HIP(hipMemcpyAsync(g + t.keys, gl.set_keys.data(), 8ull * gl.sets, hipMemcpyHostToDevice, s));
  TRY(prepare_build(ctx, cfg, pl, s));
  BatchInfo bi;
  bi.sets = gl.sets;
  bi.set_key_off = reinterpret_cast<const uint64_t*>(g + t.key_off);
  bi.set_part_base = reinterpret_cast<const uint32_t*>(g + t.part_base);
  bi.set_remap_base = reinterpret_cast<const uint32_t*>(g + t.remap_base);
  bi.global_offsets = reinterpret_cast<uint32_t*>(g + t.goff);
  bi.global_roff = reinterpret_cast<uint32_t*>(g + t.groff);
  bi.part_remap = reinterpret_cast<uint32_t*>(g + t.part_remap);
  if (gl.mixed) {
    bi.set_seed = reinterpret_cast<const uint64_t*>(g + t.seed);
    bi.set_B = reinterpret_cast<const uint32_t*>(g + t.setB);
    bi.set_ftab = reinterpret_cast<const uint32_t*>(g + t.ftab);
    bi.set_pilot_base = reinterpret_cast<const uint64_t*>(g + t.pilot_base);
    bi.ftabs = reinterpret_cast<const uint32_t*>(g + t.ftabs);
    bi.ntabs = gl.ntabs;
  }
  if (in_place) bi.set_keys = reinterpret_cast<const uint64_t* const*>(g + t.keys);
  KeySource src{};
  if (kind == kKindU64) {
    src.u64 = static_cast<const uint64_t*>(d_keys);
  } else {
    src.bytes = static_cast<const uint8_t*>(d_keys);
    src.offs = d_offs;
  }
  ctx->in_ms = ctx->out_ms = 0;
  return launch_build(ctx, ctx_scratch(ctx), kind, src, cfg, pl, block, d_status, s, bi, true);
}

int phobic::ctx_device(const phobic_ctx* ctx) { return ctx->device; }

int phobic::build_u64_in_scratch(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n, const phobic_config* cfg,
                                 const phobic_outputs* d_out, uint32_t* d_status, hipStream_t s, void* base,
                                 size_t bytes, uint32_t* fulcs_host) {
  if (!ctx || !d_status || !base || !fulcs_host || (n && !d_keys)) return PHOBIC_E_INVALID_CONFIG;
  Plan pl0;
  TRY(validate(cfg, n, &pl0));
  if (!outputs_ok(d_out, pl0)) return PHOBIC_E_INVALID_CONFIG;
  // regions at 256-byte boundaries; the fulcrum table always first, so the
  // copy recorded in fulcs_host stays valid from one call to the next
  const ScratchSizes z = scratch_sizes(tuned(ctx, pl0));
  uint64_t end = 0;
  auto take = [&](uint64_t sz) {
    const uint64_t at = end;
    end += (sz + 255) & ~uint64_t(255);
    return at;
  };
  const uint64_t o_f = take(z.fulcs), o_pc = take(z.pcount), o_lc = take(z.lvl_count), o_st = take(z.scan_tot);
  const uint64_t o_lo0 = take(z.rec_lo), o_lo1 = take(z.rec_lo), o_m0 = take(z.rec_meta), o_m1 = take(z.rec_meta);
  if (end > bytes) return PHOBIC_E_SCRATCH_LIMIT;
  TRY(set_device(ctx));
  char* b = static_cast<char*>(base);
  Scratch sc;
  sc.fulcs = reinterpret_cast<uint32_t*>(b + o_f);
  sc.pcount = reinterpret_cast<uint32_t*>(b + o_pc);
  sc.lvl_count = reinterpret_cast<uint32_t*>(b + o_lc);
  sc.scan_tot = reinterpret_cast<uint32_t*>(b + o_st);
  sc.rec_lo[0] = reinterpret_cast<uint64_t*>(b + o_lo0);
  sc.rec_lo[1] = reinterpret_cast<uint64_t*>(b + o_lo1);
  sc.rec_meta[0] = reinterpret_cast<uint32_t*>(b + o_m0);
  sc.rec_meta[1] = reinterpret_cast<uint32_t*>(b + o_m1);
  uint32_t f[kFulcsTable];
  config_fulcrums(cfg, f);
  if (std::memcmp(f, fulcs_host, sizeof(f)) != 0) {
    // from pageable memory: the runtime has read f when the call returns
    TRY_HIP(hipMemcpyAsync(sc.fulcs, f, sizeof(f), hipMemcpyHostToDevice, s));
    std::memcpy(fulcs_host, f, sizeof(f));
  }
  KeySource src{};
  src.u64 = d_keys;
  ctx->in_ms = ctx->out_ms = 0;
  return launch_build(ctx, sc, kKindU64, src, cfg, pl0, *d_out, d_status, s, BatchInfo{}, true);
}

hipError_t phobic::dev_alloc(void** p, size_t bytes, hipStream_t s) {
  hipMemPool_t pool = nullptr;
  const hipError_t e = pool_of_current(&pool);
  if (e != hipSuccess) return e;
  return hipMallocFromPoolAsync(p, bytes ? bytes : 256, pool, s);
}

hipError_t phobic::dev_free(void* p, hipStream_t s) { return p ? hipFreeAsync(p, s) : hipSuccess; }

void phobic::dev_trim() {
  hipMemPool_t pool = nullptr;
  if (pool_of_current(&pool) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
}

// A batch of independent key sets built together (one configuration).
struct phobic_batch {
  phobic_ctx* ctx = nullptr;
  phobic_config cfg{};
  std::vector<uint32_t> fulcs_copy;  // keeps cfg.fulcrums valid
  Plan pl{};
  uint32_t sets = 0;
  std::vector<uint32_t> part_base;  // sets + 1
  DevBuf d_key_off, d_part_base, d_goff;
  // graph captured for one set of device pointers
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* g_keys = nullptr;
  const void* g_pilots = nullptr;
  const void* g_offsets = nullptr;
  const void* g_status = nullptr;
  const void* g_scratch[kGraphScratch] = {};  // context scratch the graph's kernels use
};

namespace {

// Context scratch addresses baked into a captured graph.
void scratch_ptrs(const phobic_ctx* ctx, const void* (&p)[kGraphScratch]) {
  const void* v[kGraphScratch] = {ctx->pcount.ptr,    ctx->lvl_count.ptr,   ctx->scan_tot.ptr,
                                  ctx->rec_lo[0].ptr, ctx->rec_lo[1].ptr,   ctx->rec_meta[0].ptr,
                                  ctx->rec_meta[1].ptr, ctx->fulcs.ptr};
  for (int i = 0; i < kGraphScratch; i++) p[i] = v[i];
}

void batch_drop_graph(phobic_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int batch_enqueue(phobic_batch* b, const uint64_t* d_keys, uint32_t* d_pilots, uint32_t* d_offsets,
                  uint32_t* d_status, hipStream_t s) {
  KeySource src{};
  src.u64 = d_keys;
  BatchInfo bi;
  bi.sets = b->sets;
  bi.set_key_off = b->d_key_off.as<uint64_t>();
  bi.set_part_base = b->d_part_base.as<uint32_t>();
  bi.global_offsets = b->d_goff.as<uint32_t>();
  phobic_outputs out{d_pilots, d_offsets, nullptr, nullptr};
  return launch_build(b->ctx, ctx_scratch(b->ctx), kKindU64, src, &b->cfg, b->pl, out, d_status, s, bi, false);
}

}  // namespace

extern "C" {

int phobic_batch_create(phobic_ctx* ctx, const phobic_config* cfg, const uint64_t* set_offsets, uint32_t sets,
                        phobic_batch** out) {
  if (!ctx || !cfg || !set_offsets || !out || sets == 0) return PHOBIC_E_INVALID_CONFIG;
  *out = nullptr;
  const double alpha = cfg->alpha == 0.0 ? 1.0 : cfg->alpha;
  if (alpha != 1.0) return PHOBIC_E_INVALID_CONFIG;  // batches use the reference load factor
  Plan one;
  TRY(validate(cfg, 0, &one));
  // no exception may leave the C ABI: allocations that throw become E_OOM
  std::vector<uint32_t> pb;
  try {
    pb.assign(sets + 1, 0);
  } catch (...) {
    return PHOBIC_E_OOM;
  }
  uint64_t total_parts = 0;
  for (uint32_t s = 0; s < sets; s++) {
    if (set_offsets[s + 1] < set_offsets[s]) return PHOBIC_E_INVALID_CONFIG;
    const uint64_t ns = set_offsets[s + 1] - set_offsets[s];
    if (ns > 0xFFFFFFFFull) return PHOBIC_E_INVALID_CONFIG;
    total_parts += (ns + cfg->partition_size - 1) / cfg->partition_size;
    if (total_parts > 0xFFFFFFFFull) return PHOBIC_E_INVALID_CONFIG;
    pb[s + 1] = (uint32_t)total_parts;
  }
  const uint64_t n = set_offsets[sets] - set_offsets[0];
  if (set_offsets[0] != 0) return PHOBIC_E_INVALID_CONFIG;
  // the flattened build: np = total partitions (validate sizes it from n)
  Plan pl;
  {
    phobic_config c2 = *cfg;
    TRY(validate(&c2, 0, &pl));
    pl.n = n;
    pl.np = (uint32_t)total_parts;
    if (pl.np > (1u << 21)) return PHOBIC_E_INVALID_CONFIG;
    if (!plan_levels(&pl)) return PHOBIC_E_INVALID_CONFIG;
  }
  phobic_batch* b = new (std::nothrow) phobic_batch();
  if (!b) return PHOBIC_E_OOM;
  b->ctx = ctx;
  b->cfg = *cfg;
  try {
    if (cfg->fulcrums) {
      b->fulcs_copy.assign(cfg->fulcrums, cfg->fulcrums + kFulcsTable);
      b->cfg.fulcrums = b->fulcs_copy.data();
    }
    b->part_base = std::move(pb);
  } catch (...) {
    delete b;
    return PHOBIC_E_OOM;
  }
  b->pl = pl;
  b->sets = sets;
  int st = set_device(ctx);
  hipStream_t s = ctx->stream;
  auto fail = [&](int code) {
    b->d_key_off.release(s);
    b->d_part_base.release(s);
    b->d_goff.release(s);
    (void)hipStreamSynchronize(s);
    delete b;
    return code;
  };
  if (st) return fail(st);
  hipError_t e = b->d_key_off.ensure((uint64_t)(sets + 1) * 8, s);
  if (e == hipSuccess) e = b->d_part_base.ensure((uint64_t)(sets + 1) * 4, s);
  if (e == hipSuccess) e = b->d_goff.ensure(((uint64_t)pl.np + 1) * 4, s);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_key_off.ptr, set_offsets, (uint64_t)(sets + 1) * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_part_base.ptr, b->part_base.data(), (uint64_t)(sets + 1) * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(hip_status(e));
  st = prepare_build(ctx, &b->cfg, pl, s);
  if (st) return fail(st);
  *out = b;
  return PHOBIC_OK;
}

int phobic_batch_layout(const phobic_batch* b, uint32_t* part_base_out, uint32_t* buckets) {
  if (!b) return PHOBIC_E_INVALID_CONFIG;
  if (part_base_out) std::memcpy(part_base_out, b->part_base.data(), (b->sets + 1) * 4);
  if (buckets) *buckets = b->pl.B;
  return PHOBIC_OK;
}

int phobic_batch_launch(phobic_batch* b, const uint64_t* d_keys, uint32_t* d_pilots, uint32_t* d_offsets,
                        uint32_t* d_status, void* stream) {
  if (!b || !d_pilots || !d_offsets || !d_status || (b->pl.n && !d_keys)) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(b->ctx));
  // the context scratch may have been resized by another build in between
  TRY(prepare_build(b->ctx, &b->cfg, b->pl, (hipStream_t)stream));
  return batch_enqueue(b, d_keys, d_pilots, d_offsets, d_status, (hipStream_t)stream);
}

int phobic_batch_graph_launch(phobic_batch* b, const uint64_t* d_keys, uint32_t* d_pilots, uint32_t* d_offsets,
                              uint32_t* d_status, void* stream) {
  if (!b || !d_pilots || !d_offsets || !d_status || (b->pl.n && !d_keys)) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(b->ctx));
  hipStream_t s = (hipStream_t)stream;
  // fulcrums re-uploaded and scratch grown as needed; a replay is valid
  // only while every address the graph captured is unchanged
  TRY(prepare_build(b->ctx, &b->cfg, b->pl, s));
  const void* cur[kGraphScratch];
  scratch_ptrs(b->ctx, cur);
  bool same = b->exec && b->g_keys == d_keys && b->g_pilots == d_pilots && b->g_offsets == d_offsets &&
              b->g_status == d_status;
  for (int i = 0; i < kGraphScratch && same; i++) same = cur[i] == b->g_scratch[i];
  if (!same) {
    batch_drop_graph(b);
    hipStream_t cs = nullptr;
    TRY_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int st = e == hipSuccess ? batch_enqueue(b, d_keys, d_pilots, d_offsets, d_status, cs) : hip_status(e);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (st != PHOBIC_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    TRY_HIP(e2);
    b->graph = g;
    TRY_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->g_keys = d_keys;
    b->g_pilots = d_pilots;
    b->g_offsets = d_offsets;
    b->g_status = d_status;
    for (int i = 0; i < kGraphScratch; i++) b->g_scratch[i] = cur[i];
  }
  TRY_HIP(hipGraphLaunch(b->exec, s));
  return PHOBIC_OK;
}

int phobic_batch_destroy(phobic_batch* b) {
  if (!b) return PHOBIC_OK;
  (void)hipSetDevice(b->ctx->device);
  batch_drop_graph(b);
  // the batch's last launch may still read its tables on another stream
  (void)hipDeviceSynchronize();
  const hipStream_t s = b->ctx->stream;
  b->d_key_off.release(s);
  b->d_part_base.release(s);
  b->d_goff.release(s);
  (void)hipStreamSynchronize(s);
  delete b;
  return PHOBIC_OK;
}


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
  // the device address of the allocation-free entry point's scratch
  const int ns = phobic::noalloc_init(device);
  if (ns != PHOBIC_OK) {
    delete ctx;
    return ns;
  }
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  for (int i = 0; e == hipSuccess && i <= kStages; i++) e = hipEventCreate(&ctx->ev[i]);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev_in);
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
  // device builds may have run on caller streams
  (void)hipDeviceSynchronize();
  for (DevBuf* b : {&ctx->pcount, &ctx->lvl_count, &ctx->scan_tot, &ctx->rec_lo[0], &ctx->rec_lo[1], &ctx->rec_meta[0],
                    &ctx->rec_meta[1], &ctx->fulcs, &ctx->grp, &ctx->status, &ctx->in_a, &ctx->in_b, &ctx->out_pilots,
                    &ctx->out_offsets, &ctx->out_remap, &ctx->out_roff})
    b->release(ctx->stream);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  phobic::dev_trim();
  for (unsigned i = 0; i < kStageSlots; i++)
    if (ctx->slot_ev[i]) (void)hipEventSynchronize(ctx->slot_ev[i]);
  for (int i = 0; i < 2; i++) {
    if (ctx->pinned[i]) (void)hipHostFree(ctx->pinned[i]);
    if (ctx->pin_ev[i]) (void)hipEventDestroy(ctx->pin_ev[i]);
  }
  for (unsigned i = 0; i < kStageSlots; i++)
    if (ctx->slot_ev[i]) (void)hipEventDestroy(ctx->slot_ev[i]);
  for (int i = 0; i <= kStages; i++)
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->ev_in) (void)hipEventDestroy(ctx->ev_in);
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
    case PHOBIC_E_SCRATCH_LIMIT: return "build exceeds the allocation-free entry point's fixed scratch";
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
  config_fulcrums(cfg, out);
  return PHOBIC_OK;
}

int phobic_remap_capacity(const phobic_config* cfg, uint64_t n, uint64_t* entries) {
  Plan pl;
  TRY(validate(cfg, n, &pl));
  if (!entries) return PHOBIC_E_INVALID_CONFIG;
  *entries = remap_capacity(pl);
  return PHOBIC_OK;
}

int phobic_build_u64(phobic_ctx* ctx, const uint64_t* keys, uint64_t n, const phobic_config* cfg,
                     const phobic_outputs* out, phobic_stage_times* times) {
  if (n && !keys) return PHOBIC_E_INVALID_CONFIG;
  return host_build(ctx, kKindU64, keys, n * 8, nullptr, 0, n, cfg, out, times);
}

int phobic_build_bytes(phobic_ctx* ctx, const uint8_t* bytes, const uint64_t* offsets, uint64_t n,
                       const phobic_config* cfg, const phobic_outputs* out, phobic_stage_times* times) {
  if (!offsets || (n && !bytes && offsets[n] != offsets[0])) return PHOBIC_E_INVALID_CONFIG;
  for (uint64_t i = 0; i < n; i++)
    if (offsets[i + 1] < offsets[i]) return PHOBIC_E_INVALID_CONFIG;
  // keys index the byte buffer from offsets[0]; ship the whole prefix
  return host_build(ctx, kKindBytes, bytes, offsets[n], offsets, (n + 1) * 8, n, cfg, out, times);
}

int phobic_build_keys(phobic_ctx* ctx, const phobic_key* keys, uint64_t n, const phobic_config* cfg,
                      const phobic_outputs* out, phobic_stage_times* times) {
  if (n && !keys) return PHOBIC_E_INVALID_CONFIG;
  return host_build(ctx, kKindKey, keys, n * 16, nullptr, 0, n, cfg, out, times);
}

int phobic_build_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t n, const phobic_config* cfg,
                            const phobic_outputs* d_out, uint32_t* d_status, void* stream) {
  KeySource src{};
  src.u64 = d_keys;
  if (n && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  return device_build(ctx, kKindU64, src, n, cfg, d_out, d_status, stream);
}

int phobic_build_bytes_device(phobic_ctx* ctx, const uint8_t* d_bytes, const uint64_t* d_offsets, uint64_t n,
                              const phobic_config* cfg, const phobic_outputs* d_out, uint32_t* d_status,
                              void* stream) {
  KeySource src{};
  src.bytes = d_bytes;
  src.offs = d_offsets;
  if (!d_offsets) return PHOBIC_E_INVALID_CONFIG;
  return device_build(ctx, kKindBytes, src, n, cfg, d_out, d_status, stream);
}

int phobic_build_keys_device(phobic_ctx* ctx, const phobic_key* d_keys, uint64_t n, const phobic_config* cfg,
                             const phobic_outputs* d_out, uint32_t* d_status, void* stream) {
  KeySource src{};
  src.keys = reinterpret_cast<const uint4*>(d_keys);
  if (n && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  return device_build(ctx, kKindKey, src, n, cfg, d_out, d_status, stream);
}

int phobic_get_tuning(phobic_ctx* ctx, phobic_tuning* t) {
  if (!ctx || !t) return PHOBIC_E_INVALID_CONFIG;
  *t = ctx->tuning;
  return PHOBIC_OK;
}

int phobic_set_tuning(phobic_ctx* ctx, const phobic_tuning* t) {
  if (!ctx || !t) return PHOBIC_E_INVALID_CONFIG;
  const uint32_t w = t->search_waves, per = t->route_per, f = t->max_fanout;
  const uint32_t sp = t->split_per, f1 = t->first_fanout;
  const uint32_t th = t->route_threads;
  auto tile_per = [th](uint32_t x) {
    return th == 1024 ? (x == 4 || x == 6 || x == 8 || x == 10 || x == 12)
                      : (x == 8 || x == 12 || x == 16 || x == 20 || x == 24);
  };
  if (!(w == 1 || w == 2 || w == 4 || w == 8) || !(th == 0 || th == 512 || th == 1024) || !tile_per(per) || f < 16 ||
      f > kMaxFanout || t->search_mode > 3 || !(sp == 0 || tile_per(sp)) || (f1 != 0 && (f1 < 2 || f1 > f)))
    return PHOBIC_E_INVALID_CONFIG;
  ctx->tuning = *t;
  return PHOBIC_OK;
}

int phobic_autotune(phobic_ctx* ctx, const phobic_config* cfg, uint64_t sample_keys, phobic_tuning* best,
                    float* best_ms) {
  if (!ctx || sample_keys == 0) return PHOBIC_E_INVALID_CONFIG;
  Plan pl;
  TRY(validate(cfg, sample_keys, &pl));
  TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  uint64_t rcap = 0;
  TRY(phobic_remap_capacity(cfg, sample_keys, &rcap));
  // sample keys and outputs, freed on return
  DevBuf keys, pil, off, rm, ro, st;
  auto cleanup = [&]() {
    for (DevBuf* b : {&keys, &pil, &off, &rm, &ro, &st}) b->release(s);
    (void)hipStreamSynchronize(s);
  };
  hipError_t e = keys.ensure(sample_keys * 8, s);
  if (e == hipSuccess) e = pil.ensure((uint64_t)pl.np * pl.B * 4, s);
  if (e == hipSuccess) e = off.ensure(((uint64_t)pl.np + 1) * 4, s);
  if (e == hipSuccess) e = rm.ensure(rcap * 4 + 4, s);
  if (e == hipSuccess) e = ro.ensure(((uint64_t)pl.np + 1) * 4, s);
  if (e == hipSuccess) e = st.ensure(4, s);
  if (e == hipSuccess) e = launch_splitmix64(0x5eed, sample_keys, keys.as<uint64_t>(), s);
  if (e != hipSuccess) {
    cleanup();
    return hip_status(e);
  }
  KeySource src{};
  src.u64 = keys.as<uint64_t>();
  phobic_outputs outs{pil.as<uint32_t>(), off.as<uint32_t>(), rcap ? rm.as<uint32_t>() : nullptr,
                      rcap ? ro.as<uint32_t>() : nullptr};
  const phobic_tuning saved = ctx->tuning;
  ctx->in_ms = ctx->out_ms = 0;
  // The partitioning shape (route_per, max_fanout) only changes the stages
  // before the offsets and the search shape (search_waves, search_mode) only
  // the search, so each is tuned on its own stages' device time.
  auto time_stages = [&](const phobic_tuning& t, int e0, int e1, float* ms) -> int {
    ctx->tuning = t;
    *ms = 1e30f;
    for (int rep = 0; rep < 3; rep++) {  // first run warms up
      int r = enqueue_build(ctx, kKindU64, src, cfg, pl, outs, st.as<uint32_t>(), s);
      if (r == PHOBIC_OK && hipStreamSynchronize(s) != hipSuccess) r = PHOBIC_E_HIP;
      if (r != PHOBIC_OK) return r;
      float t_ms = 0;
      if (rep && hipEventElapsedTime(&t_ms, ctx->ev[e0], ctx->ev[e1]) == hipSuccess) *ms = std::min(*ms, t_ms);
    }
    return PHOBIC_OK;
  };
  phobic_tuning win = saved;
  float part_ms = 1e30f, search_ms = 1e30f;
  struct Shape {
    uint32_t threads, per, sper;
  };
  const Shape shapes[] = {{512, 16, 16}, {512, 20, 20}, {512, 24, 20}, {512, 24, 24}, {1024, 8, 8},
                          {1024, 10, 10}, {1024, 12, 10}, {1024, 12, 12}};
  for (const Shape& sh : shapes)
    for (uint32_t f1 : {0u, 48u, 512u})
      for (uint32_t f : {128u, 1024u}) {
        if (f1 > f) continue;
        phobic_tuning t = win;
        t.route_threads = sh.threads;
        t.route_per = sh.per;
        t.split_per = sh.sper;
        t.first_fanout = f1;
        t.max_fanout = f;
        float ms;
        const int r = time_stages(t, 0, kStages - 2, &ms);  // hash_route + partition_split
        if (r != PHOBIC_OK) {
          ctx->tuning = saved;
          cleanup();
          return r;
        }
        if (ms < part_ms) {
          part_ms = ms;
          win.route_threads = sh.threads;
          win.route_per = sh.per;
          win.split_per = sh.sper;
          win.first_fanout = f1;
          win.max_fanout = f;
        }
      }
  for (uint32_t w : {1u, 2u, 4u, 8u})
    for (uint32_t mode : {0u, 1u, 2u, 3u}) {
      phobic_tuning t = win;
      t.search_waves = w;
      t.search_mode = mode;
      float ms;
      const int r = time_stages(t, kStages - 1, kStages, &ms);  // partition_search
      if (r != PHOBIC_OK) {
        ctx->tuning = saved;
        cleanup();
        return r;
      }
      if (ms < search_ms) {
        search_ms = ms;
        win.search_waves = w;
        win.search_mode = mode;
      }
    }
  // the whole build with the winning shape
  float win_ms;
  {
    const int r = time_stages(win, 0, kStages, &win_ms);
    if (r != PHOBIC_OK) {
      ctx->tuning = saved;
      cleanup();
      return r;
    }
  }
  ctx->tuning = win;
  cleanup();
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return PHOBIC_OK;
}

int phobic_last_stage_times(phobic_ctx* ctx, phobic_stage_times* times) {
  if (!ctx || !times) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  for (int i = 0; i <= kStages && ctx->timed; i++) TRY_HIP(hipEventSynchronize(ctx->ev[i]));
  return fill_times(ctx, times);
}

}  // extern "C"

namespace {

int eval_device(phobic_ctx* ctx, int kind, const KeySource& src, uint64_t count, const phobic_config* cfg,
                uint32_t np, const phobic_outputs* d_mphf, uint32_t* d_out, void* stream) {
  Plan pl;
  TRY(validate(cfg, 0, &pl));
  if (!ctx || !d_mphf || np == 0 || (count && (!d_mphf->pilots || !d_mphf->offsets || !d_out)))
    return PHOBIC_E_INVALID_CONFIG;
  if (pl.a_ppm < 1000000u && (!d_mphf->remap || !d_mphf->remap_offsets)) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  TRY(upload_fulcrums(ctx, cfg, (hipStream_t)stream));
  EvalArgs ea{};
  ea.src = src;
  ea.count = count;
  ea.seed = cfg->seed;
  ea.np = np;
  ea.a_ppm = pl.a_ppm;
  ea.fulcs = ctx->fulcs.as<uint32_t>();
  ea.pilots = d_mphf->pilots;
  ea.offsets = d_mphf->offsets;
  ea.remap = d_mphf->remap;
  ea.remap_offsets = d_mphf->remap_offsets;
  ea.out = d_out;
  TRY_HIP(launch_eval(kind, ea, (hipStream_t)stream));
  return PHOBIC_OK;
}

}  // namespace

extern "C" {

int phobic_eval_u64_device(phobic_ctx* ctx, const uint64_t* d_keys, uint64_t count, const phobic_config* cfg,
                           uint32_t np, const phobic_outputs* d_mphf, uint32_t* d_out, void* stream) {
  if (count && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  KeySource src{};
  src.u64 = d_keys;
  return eval_device(ctx, kKindU64, src, count, cfg, np, d_mphf, d_out, stream);
}

int phobic_eval_bytes_device(phobic_ctx* ctx, const uint8_t* d_bytes, const uint64_t* d_offsets, uint64_t count,
                             const phobic_config* cfg, uint32_t np, const phobic_outputs* d_mphf, uint32_t* d_out,
                             void* stream) {
  if (count && !d_offsets) return PHOBIC_E_INVALID_CONFIG;
  KeySource src{};
  src.bytes = d_bytes;
  src.offs = d_offsets;
  return eval_device(ctx, kKindBytes, src, count, cfg, np, d_mphf, d_out, stream);
}

int phobic_eval_pairs_device(phobic_ctx* ctx, const uint64_t* d_pairs, uint64_t count, const phobic_config* cfg,
                             uint32_t np, const phobic_outputs* d_mphf, uint32_t* d_out, void* stream) {
  if (count && !d_pairs) return PHOBIC_E_INVALID_CONFIG;
  KeySource src{};
  src.keys = reinterpret_cast<const uint4*>(d_pairs);
  return eval_device(ctx, kKindPair, src, count, cfg, np, d_mphf, d_out, stream);
}

int phobic_eval_keys_device(phobic_ctx* ctx, const phobic_key* d_keys, uint64_t count, const phobic_config* cfg,
                            uint32_t np, const phobic_outputs* d_mphf, uint32_t* d_out, void* stream) {
  if (count && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  KeySource src{};
  src.keys = reinterpret_cast<const uint4*>(d_keys);
  return eval_device(ctx, kKindKey, src, count, cfg, np, d_mphf, d_out, stream);
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

int phobic_copy_device(void* d_dst, const void* d_src, uint64_t bytes, void* stream) {
  if (bytes % 16 || (bytes && (!d_dst || !d_src)) || (((uintptr_t)d_dst | (uintptr_t)d_src) & 15))
    return PHOBIC_E_INVALID_CONFIG;
  TRY_HIP(launch_copy16(d_dst, d_src, bytes / 16, (hipStream_t)stream));
  return PHOBIC_OK;
}

int phobic_copy_to_device(phobic_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes, void* stream) {
  if (!ctx || (bytes && (!d_dst || !h_src))) return PHOBIC_E_INVALID_CONFIG;
  TRY(set_device(ctx));
  if (!bytes) return PHOBIC_OK;
  hipStream_t s = (hipStream_t)stream;
  TRY(copy_to_device(ctx, d_dst, h_src, bytes, s));
  TRY_HIP(hipStreamSynchronize(s));
  return PHOBIC_OK;
}

int phobic_gen_splitmix64_device(uint64_t seed, uint64_t n, uint64_t* d_keys, void* stream) {
  if (n && !d_keys) return PHOBIC_E_INVALID_CONFIG;
  TRY_HIP(launch_splitmix64(seed, n, d_keys, (hipStream_t)stream));
  return PHOBIC_OK;
}

int phobic_gen_strings_device(uint64_t seed, uint64_t n, const uint64_t* d_offsets, uint8_t* d_bytes, void* stream) {
  if (n && (!d_offsets || !d_bytes)) return PHOBIC_E_INVALID_CONFIG;
  TRY_HIP(launch_gen_strings(seed, n, d_offsets, d_bytes, (hipStream_t)stream));
  return PHOBIC_OK;
}

}  // extern "C"
