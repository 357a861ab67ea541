// server.cpp -- continuous-batching build server (no counterpart in the
// reference, which builds one MPHF per synchronous MPHFbuilder::build call,
// mphf_builder.h:315-321).
//
// Requests queue in arrival order.  Whenever a worker slot is idle the
// scheduler forms a job from the oldest request that fits the HBM budget
// (backfill: a later request may overtake one that does not fit yet).  The
// request takes every queued request of its configuration and size class
// with it: the job is one flattened launch over all of their keys (a group,
// built by phobic::group_build), whose per-request outputs equal building
// each alone.  A request with a custom bucketer table, or device-resident
// strings, runs alone.  Every slot owns a phobic_ctx (its own stream and
// scratch), so jobs run concurrently on the device; results stay in device
// memory until released.  Host inputs are uploaded through the slot
// context's pinned staging chunks (allocated when the server starts).
//
// Memory accounting (all in the pool's 2 MiB units):
//   kept      outputs of completed, unreleased requests (one block per job,
//             freed with its last request)
//   running   staging + outputs + scratch growth charged to running jobs
//   retained  scratch each slot's context keeps between jobs
// hbm_in_use = kept + running + retained; a job is admitted when it fits
// the budget beside that (idle slots give their retained scratch back when
// nothing else would ever fit).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/phobic_hip.h"
#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

// Queued requests of one configuration share launches.  Small and large
// requests are grouped apart, so that a small request never waits for a
// large one's build: small groups hold up to kGroupKeys keys, large ones
// (requests above kSmallKeys each) up to kGroupKeysLarge.
constexpr uint64_t kSmallKeys = 4000000;
constexpr uint64_t kGroupKeys = 64000000;
constexpr uint64_t kGroupKeysLarge = 256000000;
constexpr uint32_t kGroupSets = 1024;  // requests of one shared launch at most
constexpr size_t kGroupScan = 4096;    // queued requests a job's leader looks through for its group

uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// Device memory holding the outputs of one job (one or several requests).
struct Block {
  void* ptr = nullptr;
  uint64_t bytes = 0;  // pool units
  uint32_t refs = 0;   // unreleased successful requests whose outputs live here
};

struct Request {
  uint64_t id = 0;
  phobic_request req{};
  phobic_config cfg{};
  std::vector<uint32_t> fulcs;  // copy of a custom table
  uint32_t np = 0, B = 0;
  uint64_t remap_cap = 0;
  Clock::time_point t_submit, t_start, t_done;
  int state = 0;  // 0 queued, 1 running, 2 done
  int status = PHOBIC_OK;
  std::shared_ptr<Block> block;
  phobic_outputs outs{};
};

// Output block of a single request: pilots | offsets | remap | remap offsets | status.
uint64_t single_out_bytes(const Request& r) {
  return align256((uint64_t)r.np * r.B * 4) + align256(((uint64_t)r.np + 1) * 4) + align256(r.remap_cap * 4) +
         align256(((uint64_t)r.np + 1) * 4) + 256;
}
uint64_t single_in_bytes(const Request& r) {
  if (r.req.keys_on_device) return 0;  // built in place
  if (r.req.kind == PHOBIC_KEYS_BYTES) return align256(r.req.bytes_len) + (r.req.n + 1) * 8 + 256;
  return r.req.n * 8;
}

// Host inputs below this size are copied directly (pageable memory); larger
// ones go through the slot context's pinned staging chunks.
constexpr uint64_t kDirectCopyMax = 1ull << 20;

int upload(phobic_ctx* ctx, void* dst, const void* src, uint64_t bytes, hipStream_t s) {
  if (!bytes) return PHOBIC_OK;
  if (bytes < kDirectCopyMax)
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) == hipSuccess ? PHOBIC_OK : PHOBIC_E_HIP;
  return phobic::stage_to_device(
      ctx, dst, bytes,
      [src](char* d, uint64_t lo, uint64_t len) { std::memcpy(d, static_cast<const char*>(src) + lo, len); }, s);
}

// Output block of a group: pilots | offsets | remap | remap offsets | status
// over the flattened build.
uint64_t group_out_bytes(const phobic::GroupLayout& gl) {
  const uint64_t entries = (uint64_t)gl.np + gl.sets;
  const bool remap = gl.a_ppm < 1000000u;
  return align256(gl.pilot_base[gl.sets] * 4) + align256(entries * 4) +
         (remap ? align256((uint64_t)gl.remap_base[gl.sets] * 4) + align256(entries * 4) : 0) + 256;
}

struct Job {
  std::vector<std::shared_ptr<Request>> reqs;
  bool group = false;
  phobic::GroupLayout gl;
  uint64_t in_bytes = 0;  // group staging: keys (+ string offsets)
  uint64_t charge = 0;    // counted in `running` while the job runs
};

struct Slot {
  phobic_ctx* ctx = nullptr;
  std::shared_ptr<Job> job;
  std::thread th;
  uint64_t retained = 0;
};

bool groupable(const Request& r) {
  return r.req.n > 0 && !r.cfg.fulcrums &&
         (r.req.kind == PHOBIC_KEYS_U64 || (r.req.kind == PHOBIC_KEYS_BYTES && !r.req.keys_on_device));
}

bool large(const Request& r) { return r.req.n > kSmallKeys; }

double alpha_of(const phobic_config& c) { return c.alpha == 0.0 ? 1.0 : c.alpha; }

// Requests one launch can build together: lambda and the seed may differ
// (at most kMaxFulcTabs lambdas in a group, counted in form_job).
bool same_class(const Request& a, const Request& b) {
  return a.req.kind == b.req.kind && a.cfg.partition_size == b.cfg.partition_size && a.cfg.buckets == b.cfg.buckets &&
         alpha_of(a.cfg) == alpha_of(b.cfg) && large(a) == large(b);
}

}  // namespace

struct phobic_server {
  phobic_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  std::vector<Slot> slots;
  hipStream_t free_stream = nullptr;  // frees of released outputs
  uint64_t next_id = 1;
  uint64_t kept = 0, running = 0;
  bool stop = false;
  // statistics
  std::vector<double> latencies_ms;
  uint64_t submitted = 0, completed = 0, rejected = 0, keys_built = 0;
  uint64_t grouped_launches = 0, grouped_requests = 0;
  Clock::time_point t_first{}, t_last{};
  bool any = false;
};

namespace {

uint64_t retained_total(const phobic_server* s) {
  uint64_t t = 0;
  for (const auto& sl : s->slots) t += sl.retained;
  return t;
}

uint64_t in_use(const phobic_server* s) { return s->kept + s->running + retained_total(s); }

// Charge of a job on a slot's context: staging, outputs, scratch growth.
uint64_t job_charge(const Job& j, const phobic_ctx* ctx) {
  const Request& r0 = *j.reqs[0];
  if (j.group) {
    return phobic::pool_bytes(j.in_bytes) + phobic::pool_bytes(group_out_bytes(j.gl)) +
           phobic::group_scratch_growth(ctx, &r0.cfg, j.gl);
  }
  return phobic::pool_bytes(single_in_bytes(r0)) + phobic::pool_bytes(single_out_bytes(r0)) +
         phobic::build_scratch_growth(ctx, &r0.cfg, r0.req.n);
}

// The job request `it` leads: itself, or a group with every queued small
// request of its configuration (in queue order, within the group limits).
// Caller holds s->mu.
std::shared_ptr<Job> form_job(phobic_server* s, std::deque<std::shared_ptr<Request>>::iterator it) {
  std::shared_ptr<Job> j;
  try {
    j = std::make_shared<Job>();
    j->reqs.push_back(*it);
    if (groupable(**it)) {
      uint64_t keys = (*it)->req.n;
      const uint64_t cap = large(**it) ? kGroupKeysLarge : kGroupKeys;
      size_t scanned = 0;  // admission looks at a bounded part of a long queue
      double lambdas[phobic::kMaxFulcTabs] = {(*it)->cfg.lambda};
      uint32_t nl = 1;
      for (auto jt = std::next(it); jt != s->queue.end() && j->reqs.size() < kGroupSets && scanned < kGroupScan;
           ++jt, ++scanned) {
        if (!groupable(**jt) || !same_class(**it, **jt) || keys + (*jt)->req.n > cap) continue;
        uint32_t l = 0;
        while (l < nl && lambdas[l] != (*jt)->cfg.lambda) l++;
        if (l == phobic::kMaxFulcTabs) continue;  // a lambda more than one launch holds
        if (l == nl) lambdas[nl++] = (*jt)->cfg.lambda;
        j->reqs.push_back(*jt);
        keys += (*jt)->req.n;
      }
    }
    if (j->reqs.size() > 1) {
      std::vector<uint64_t> sizes;
      std::vector<const phobic_config*> cfgs;
      sizes.reserve(j->reqs.size());
      cfgs.reserve(j->reqs.size());
      uint64_t in = 0;
      for (const auto& r : j->reqs) {
        sizes.push_back(r->req.n);
        cfgs.push_back(&r->cfg);
        // strings are staged; u64 keys in device memory are read in place
        in += r->req.kind == PHOBIC_KEYS_BYTES ? r->req.bytes_len : r->req.keys_on_device ? 0 : r->req.n * 8;
      }
      if (phobic::group_layout(cfgs.data(), sizes.data(), (uint32_t)sizes.size(), &j->gl) != PHOBIC_OK) {
        j->reqs.resize(1);  // build it alone
      } else {
        j->group = true;
        j->in_bytes = align256(in) + (j->reqs[0]->req.kind == PHOBIC_KEYS_BYTES ? (j->gl.n + 1) * 8 : 0) + 256;
      }
    }
  } catch (...) {
    return nullptr;
  }
  return j;
}

void start_job(phobic_server* s, Slot& sl, const std::shared_ptr<Job>& j) {
  for (const auto& r : j->reqs) {
    r->state = 1;
    auto qt = std::find(s->queue.begin(), s->queue.end(), r);
    if (qt != s->queue.end()) s->queue.erase(qt);
  }
  s->running += j->charge;
  sl.job = j;
  s->cv_work.notify_all();
}

void try_admit(phobic_server* s) {
  // caller holds s->mu
  for (;;) {
    bool idle = false, busy = false;
    for (const auto& sl : s->slots) (sl.job ? busy : idle) = true;
    if (!idle || s->queue.empty()) return;
    bool started = false;
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
      std::shared_ptr<Job> j = form_job(s, it);
      if (!j) return;
      // the idle slot whose retained scratch serves the job best
      Slot* best = nullptr;
      uint64_t best_charge = 0;
      for (auto& sl : s->slots) {
        if (sl.job) continue;
        const uint64_t c = job_charge(*j, sl.ctx);
        if (!best || c < best_charge) {
          best = &sl;
          best_charge = c;
        }
      }
      // a group that does not fit is built as its leading request alone
      if (j->group && in_use(s) + best_charge > s->cfg.hbm_budget) {
        j->reqs.resize(1);
        j->group = false;
        best_charge = job_charge(*j, best->ctx);
      }
      if (in_use(s) + best_charge > s->cfg.hbm_budget && !busy && retained_total(s)) {
        // nothing runs and retained scratch is in the way: the idle slots
        // give it back
        for (auto& sl : s->slots) {
          phobic::ctx_trim(sl.ctx);
          sl.retained = 0;
        }
        best_charge = job_charge(*j, best->ctx);
        if (j->group && in_use(s) + best_charge > s->cfg.hbm_budget) {
          j->reqs.resize(1);
          j->group = false;
          best_charge = job_charge(*j, best->ctx);
        }
      }
      if (in_use(s) + best_charge <= s->cfg.hbm_budget) {
        j->charge = best_charge;
        start_job(s, *best, j);  // erases the job's requests from the queue
        started = true;
        break;
      }
    }
    if (!started) return;
  }
}

void fail_outputs(Request& r) { r.outs = phobic_outputs{}; }

// One request built alone.  Returns its status; on success r.block holds its
// outputs (owned by the caller's accounting).
int run_single(phobic_ctx* ctx, Request& r, hipStream_t stream) {
  const uint64_t n = r.req.n, NB = (uint64_t)r.np * r.B;
  auto block = std::make_shared<Block>();
  block->bytes = phobic::pool_bytes(single_out_bytes(r));
  if (phobic::dev_alloc(&block->ptr, single_out_bytes(r), stream) != hipSuccess) return PHOBIC_E_OOM;
  r.block = block;
  uint8_t* p = static_cast<uint8_t*>(block->ptr);
  r.outs.pilots = reinterpret_cast<uint32_t*>(p);
  p += align256(NB * 4);
  r.outs.offsets = reinterpret_cast<uint32_t*>(p);
  p += align256(((uint64_t)r.np + 1) * 4);
  r.outs.remap = r.remap_cap ? reinterpret_cast<uint32_t*>(p) : nullptr;
  p += align256(r.remap_cap * 4);
  r.outs.remap_offsets = r.remap_cap ? reinterpret_cast<uint32_t*>(p) : nullptr;
  p += align256(((uint64_t)r.np + 1) * 4);
  uint32_t* d_status = reinterpret_cast<uint32_t*>(p);
  const void* keys = r.req.keys;
  const uint64_t* offs = r.req.offsets;
  void* staging = nullptr;
  auto done = [&](int st) {
    if (staging) (void)phobic::dev_free(staging, stream);
    return st;
  };
  if (!r.req.keys_on_device) {  // stage host keys
    const uint64_t kb = r.req.kind == PHOBIC_KEYS_BYTES ? r.req.bytes_len : n * 8;
    const uint64_t ob = r.req.kind == PHOBIC_KEYS_BYTES ? (n + 1) * 8 : 0;
    if (phobic::dev_alloc(&staging, single_in_bytes(r), stream) != hipSuccess) return done(PHOBIC_E_OOM);
    uint8_t* q = static_cast<uint8_t*>(staging);
    int ust = upload(ctx, q, r.req.keys, kb, stream);
    if (!ust) ust = upload(ctx, q + align256(kb), r.req.offsets, ob, stream);
    if (ust) return done(ust);
    keys = q;
    offs = reinterpret_cast<const uint64_t*>(q + align256(kb));
  }
  int st;
  if (r.req.kind == PHOBIC_KEYS_BYTES)
    st = phobic_build_bytes_device(ctx, static_cast<const uint8_t*>(keys), offs, n, &r.cfg, &r.outs, d_status,
                                   stream);
  else
    st = phobic_build_u64_device(ctx, static_cast<const uint64_t*>(keys), n, &r.cfg, &r.outs, d_status, stream);
  if (st) return done(st);
  uint32_t bits = 0;
  if (hipMemcpyAsync(&bits, d_status, 4, hipMemcpyDeviceToHost, stream) != hipSuccess) return done(PHOBIC_E_HIP);
  if (hipStreamSynchronize(stream) != hipSuccess) return done(PHOBIC_E_HIP);
  for (int c = 1; c < 32; c++)
    if (bits & (1u << c)) return done(c);
  return done(PHOBIC_OK);
}

// A group built as one launch.  On success every request's outputs point
// into the shared block.
int run_group(phobic_ctx* ctx, Job& j, hipStream_t stream) {
  const phobic::GroupLayout& gl = j.gl;
  const Request& r0 = *j.reqs[0];
  const bool bytes = r0.req.kind == PHOBIC_KEYS_BYTES;
  const bool remap = gl.a_ppm < 1000000u;
  auto block = std::make_shared<Block>();
  const uint64_t ob = group_out_bytes(gl);
  block->bytes = phobic::pool_bytes(ob);
  if (phobic::dev_alloc(&block->ptr, ob, stream) != hipSuccess) return PHOBIC_E_OOM;
  void* staging = nullptr;
  auto fail = [&](int st) {
    if (staging) (void)phobic::dev_free(staging, stream);
    (void)phobic::dev_free(block->ptr, stream);
    block->ptr = nullptr;
    return st;
  };
  if (phobic::dev_alloc(&staging, j.in_bytes, stream) != hipSuccess) return fail(PHOBIC_E_OOM);
  uint8_t* q = static_cast<uint8_t*>(staging);
  uint64_t key_bytes = 0;
  for (const auto& r : j.reqs) key_bytes += bytes ? r->req.bytes_len : r->req.n * 8;
  uint64_t* d_offs = reinterpret_cast<uint64_t*>(q + align256(key_bytes));
  // Each request's keys at `at[s]` of the staging area.  Runs of requests in
  // host memory are gathered into the pinned chunks and uploaded together;
  // u64 keys already in device memory are read where they lie (no staging;
  // groupable strings are all in host memory).
  std::vector<uint64_t> at(gl.sets + 1, 0);
  std::vector<const char*> src(gl.sets);
  for (uint32_t k = 0; k < gl.sets; k++) {
    const Request& r = *j.reqs[k];
    const uint64_t kb = bytes ? r.req.offsets[r.req.n] - r.req.offsets[0] : r.req.keys_on_device ? 0 : r.req.n * 8;
    src[k] = static_cast<const char*>(r.req.keys) + (bytes ? r.req.offsets[0] : 0);
    at[k + 1] = at[k] + kb;
  }
  for (uint32_t k = 0; k < gl.sets;) {
    if (j.reqs[k]->req.keys_on_device) {
      k++;
      continue;
    }
    uint32_t e = k + 1;
    while (e < gl.sets && !j.reqs[e]->req.keys_on_device) e++;
    // requests [k, e): bytes [at[k], at[e]) of the staging area
    const uint64_t base = at[k];
    auto gather = [&, k, e, base](char* d, uint64_t off, uint64_t len) {
      uint32_t t = (uint32_t)(std::upper_bound(at.begin() + k, at.begin() + e + 1, base + off) - at.begin()) - 1;
      for (uint64_t done = 0; done < len; t++) {
        const uint64_t from = base + off + done - at[t];
        const uint64_t take = std::min(at[t + 1] - at[t] - from, len - done);
        std::memcpy(d + done, src[t] + from, take);
        done += take;
      }
    };
    const int ust = at[e] > at[k] ? phobic::stage_to_device(ctx, q + base, at[e] - base, gather, stream) : PHOBIC_OK;
    if (ust) return fail(ust);
    k = e;
  }
  if (bytes) {
    // the group's n + 1 offsets: request s's offsets rebased onto at[s],
    // then the total, produced straight into the pinned chunks
    auto rebase = [&](char* d, uint64_t off, uint64_t len) {
      uint64_t* o = reinterpret_cast<uint64_t*>(d);
      const uint64_t i0 = off / 8, cnt = len / 8;
      uint32_t t = (uint32_t)(std::upper_bound(gl.key_off.begin(), gl.key_off.end(), i0) - gl.key_off.begin()) - 1;
      for (uint64_t i = i0; i < i0 + cnt; i++) {
        while (t < gl.sets && i >= gl.key_off[t + 1]) t++;
        if (t == gl.sets) {
          o[i - i0] = at[gl.sets];
          continue;
        }
        const uint64_t* ro = j.reqs[t]->req.offsets;
        o[i - i0] = at[t] + (ro[i - gl.key_off[t]] - ro[0]);
      }
    };
    const int ust = phobic::stage_to_device(ctx, d_offs, (gl.n + 1) * 8, rebase, stream);
    if (ust) return fail(ust);
  }
  uint8_t* p = static_cast<uint8_t*>(block->ptr);
  const uint64_t entries = (uint64_t)gl.np + gl.sets;
  phobic_outputs bo{};
  bo.pilots = reinterpret_cast<uint32_t*>(p);
  p += align256(gl.pilot_base[gl.sets] * 4);
  bo.offsets = reinterpret_cast<uint32_t*>(p);
  p += align256(entries * 4);
  if (remap) {
    bo.remap = reinterpret_cast<uint32_t*>(p);
    p += align256((uint64_t)gl.remap_base[gl.sets] * 4);
    bo.remap_offsets = reinterpret_cast<uint32_t*>(p);
    p += align256(entries * 4);
  }
  uint32_t* d_status = reinterpret_cast<uint32_t*>(p);
  if (!bytes) {
    j.gl.set_keys.resize(gl.sets);
    for (uint32_t k = 0; k < gl.sets; k++)
      j.gl.set_keys[k] = j.reqs[k]->req.keys_on_device ? static_cast<const uint64_t*>(j.reqs[k]->req.keys)
                                                       : reinterpret_cast<const uint64_t*>(q + at[k]);
  }
  int st = phobic::group_build(ctx, &r0.cfg, bytes ? 1 : 0, q, bytes ? d_offs : nullptr, gl, bo, d_status, stream);
  if (st) return fail(st);
  uint32_t bits = 0;
  if (hipMemcpyAsync(&bits, d_status, 4, hipMemcpyDeviceToHost, stream) != hipSuccess) return fail(PHOBIC_E_HIP);
  if (hipStreamSynchronize(stream) != hipSuccess) return fail(PHOBIC_E_HIP);
  for (int c = 1; c < 32; c++)
    if (bits & (1u << c)) return fail(c);
  (void)phobic::dev_free(staging, stream);
  for (uint32_t s = 0; s < gl.sets; s++) {
    Request& r = *j.reqs[s];
    r.block = block;
    r.outs.pilots = bo.pilots + gl.pilot_base[s];
    r.outs.offsets = bo.offsets + gl.part_base[s] + s;
    r.outs.remap = remap ? bo.remap + gl.remap_base[s] : nullptr;
    r.outs.remap_offsets = remap ? bo.remap_offsets + gl.part_base[s] + s : nullptr;
  }
  return PHOBIC_OK;
}

void worker(phobic_server* s, size_t idx, hipStream_t stream) {
  (void)hipSetDevice(s->cfg.device);
  phobic_ctx* ctx = s->slots[idx].ctx;
  for (;;) {
    std::shared_ptr<Job> job;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || s->slots[idx].job; });
      if (s->stop && !s->slots[idx].job) return;
      job = s->slots[idx].job;
      const auto now = Clock::now();
      for (const auto& r : job->reqs) r->t_start = now;
    }
    std::vector<int> st(job->reqs.size(), PHOBIC_OK);
    bool grouped = false;
    if (job->group && run_group(ctx, *job, stream) == PHOBIC_OK) {
      grouped = true;
    } else {
      // alone, or every request of a group that failed (each then reports
      // its own status)
      for (size_t i = 0; i < job->reqs.size(); i++) {
        st[i] = run_single(ctx, *job->reqs[i], stream);
        if (st[i] != PHOBIC_OK) {
          if (job->reqs[i]->block && job->reqs[i]->block->ptr) (void)phobic::dev_free(job->reqs[i]->block->ptr, stream);
          job->reqs[i]->block.reset();
          fail_outputs(*job->reqs[i]);
        }
      }
    }
    (void)hipStreamSynchronize(stream);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      const auto now = Clock::now();
      s->running -= job->charge;
      s->slots[idx].retained = phobic::ctx_scratch_bytes(ctx);
      s->slots[idx].job.reset();
      std::vector<Block*> counted;
      for (size_t i = 0; i < job->reqs.size(); i++) {
        Request& r = *job->reqs[i];
        r.status = st[i];
        r.t_done = now;
        r.state = 2;
        if (st[i] == PHOBIC_OK) {
          Block* b = r.block.get();
          if (std::find(counted.begin(), counted.end(), b) == counted.end()) {
            s->kept += b->bytes;
            counted.push_back(b);
          }
          b->refs++;
          s->keys_built += r.req.n;
        }
        s->completed++;
        try {  // a sample that cannot be stored is dropped from the statistics
          s->latencies_ms.push_back(std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count());
        } catch (...) {
        }
      }
      if (grouped) {
        s->grouped_launches++;
        s->grouped_requests += job->reqs.size();
      }
      s->t_last = now;
      try_admit(s);
    }
    s->cv_done.notify_all();
  }
}

std::shared_ptr<Request> find(phobic_server* s, uint64_t id) {
  auto it = s->all.find(id);
  return it == s->all.end() ? nullptr : it->second;
}

}  // namespace

extern "C" {

int phobic_server_create(const phobic_server_config* cfg, phobic_server** out) {
  if (!cfg || !out || cfg->slots == 0 || cfg->slots > 64 || cfg->hbm_budget == 0) return PHOBIC_E_INVALID_CONFIG;
  *out = nullptr;
  auto* s = new (std::nothrow) phobic_server();
  if (!s) return PHOBIC_E_OOM;
  s->cfg = *cfg;
  if (s->cfg.queue_capacity == 0) s->cfg.queue_capacity = 1u << 16;
  if (hipSetDevice(cfg->device) != hipSuccess ||
      hipStreamCreateWithFlags(&s->free_stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return PHOBIC_E_NO_DEVICE;
  }
  // no exception may leave the C ABI: allocations and threads that fail
  // become E_OOM, and the partly built server is torn down
  try {
    s->slots.resize(cfg->slots);
  } catch (...) {
    (void)hipStreamDestroy(s->free_stream);
    delete s;
    return PHOBIC_E_OOM;
  }
  for (auto& sl : s->slots) {
    // each slot's pinned staging chunks are allocated now, not by the first
    // request that uploads host keys
    int st = phobic_ctx_create(cfg->device, &sl.ctx);
    if (st == PHOBIC_OK) st = phobic::ctx_prepare_staging(sl.ctx);
    if (st != PHOBIC_OK) {
      for (auto& x : s->slots)
        if (x.ctx) phobic_ctx_destroy(x.ctx);
      (void)hipStreamDestroy(s->free_stream);
      delete s;
      return st;
    }
  }
  for (size_t i = 0; i < s->slots.size(); i++) {
    hipStream_t st = phobic::ctx_stream(s->slots[i].ctx);
    try {
      s->slots[i].th = std::thread(worker, s, i, st);
    } catch (...) {
      phobic_server_destroy(s);  // joins the workers already started
      return PHOBIC_E_OOM;
    }
  }
  *out = s;
  return PHOBIC_OK;
}

int phobic_server_submit(phobic_server* s, const phobic_request* req, uint64_t* id) {
  if (!s || !req || !id) return PHOBIC_E_INVALID_CONFIG;
  if (req->kind != PHOBIC_KEYS_U64 && req->kind != PHOBIC_KEYS_BYTES) return PHOBIC_E_INVALID_CONFIG;
  if (req->n && !req->keys) return PHOBIC_E_INVALID_CONFIG;
  if (req->kind == PHOBIC_KEYS_BYTES) {
    if (!req->offsets) return PHOBIC_E_INVALID_CONFIG;
    if (!req->keys_on_device) {  // the kernels read bytes [offsets[0], offsets[n])
      for (uint64_t i = 0; i < req->n; i++)
        if (req->offsets[i + 1] < req->offsets[i]) return PHOBIC_E_INVALID_CONFIG;
      if (req->offsets[req->n] > req->bytes_len) return PHOBIC_E_INVALID_CONFIG;
    }
  }
  std::shared_ptr<Request> r;
  try {
    r = std::make_shared<Request>();
    r->req = *req;
    r->cfg = req->cfg;
    if (req->cfg.fulcrums) {
      r->fulcs.assign(req->cfg.fulcrums, req->cfg.fulcrums + 2048);
      r->cfg.fulcrums = r->fulcs.data();
    }
  } catch (...) {
    return PHOBIC_E_OOM;
  }
  int st = phobic_query_sizes(&r->cfg, req->n, &r->np, &r->B);
  if (st) return st;
  st = phobic_remap_capacity(&r->cfg, req->n, &r->remap_cap);
  if (st) return st;
  // what it needs alone on a slot that holds no scratch yet
  const uint64_t need = phobic::pool_bytes(single_in_bytes(*r)) + phobic::pool_bytes(single_out_bytes(*r)) +
                        phobic::build_scratch_growth(s->slots[0].ctx, &r->cfg, req->n, true);
  std::lock_guard<std::mutex> lk(s->mu);
  if (need > s->cfg.hbm_budget) {
    s->rejected++;
    return PHOBIC_E_OOM;
  }
  if (s->queue.size() >= s->cfg.queue_capacity) {
    s->rejected++;
    return PHOBIC_E_BUSY;
  }
  r->id = s->next_id++;
  r->t_submit = Clock::now();
  try {  // both containers hold the request, or neither does
    s->queue.push_back(r);
  } catch (...) {
    return PHOBIC_E_OOM;
  }
  try {
    s->all[r->id] = r;
  } catch (...) {
    s->queue.pop_back();
    return PHOBIC_E_OOM;
  }
  if (!s->any) {
    s->t_first = r->t_submit;
    s->any = true;
  }
  s->submitted++;
  *id = r->id;
  try_admit(s);
  return PHOBIC_OK;
}

int phobic_server_wait(phobic_server* s, uint64_t id, int* status, double* latency_ms) {
  if (!s) return PHOBIC_E_INVALID_CONFIG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto r = find(s, id);
  if (!r) return PHOBIC_E_NOT_FOUND;
  s->cv_done.wait(lk, [&] { return r->state == 2; });
  if (status) *status = r->status;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  return PHOBIC_OK;
}

int phobic_server_poll(phobic_server* s, uint64_t id, int* done) {
  if (!s || !done) return PHOBIC_E_INVALID_CONFIG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto r = find(s, id);
  if (!r) return PHOBIC_E_NOT_FOUND;
  *done = r->state == 2;
  return PHOBIC_OK;
}

int phobic_server_result(phobic_server* s, uint64_t id, uint32_t* np, uint32_t* buckets, uint64_t* remap_entries,
                         phobic_outputs* device_outputs) {
  if (!s) return PHOBIC_E_INVALID_CONFIG;
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r = find(s, id);
    if (!r) return PHOBIC_E_NOT_FOUND;
    if (r->state != 2) return PHOBIC_E_BUSY;
  }
  if (r->status != PHOBIC_OK) return r->status;
  if (np) *np = r->np;
  if (buckets) *buckets = r->B;
  if (remap_entries) {
    uint32_t used = 0;
    (void)hipSetDevice(s->cfg.device);
    if (r->remap_cap && hipMemcpy(&used, r->outs.remap_offsets + r->np, 4, hipMemcpyDeviceToHost) != hipSuccess)
      return PHOBIC_E_HIP;
    *remap_entries = used;
  }
  if (device_outputs) *device_outputs = r->outs;
  return PHOBIC_OK;
}

int phobic_server_fetch(phobic_server* s, uint64_t id, const phobic_outputs* host_out) {
  if (!s || !host_out) return PHOBIC_E_INVALID_CONFIG;
  uint32_t np = 0, B = 0;
  uint64_t used = 0;
  phobic_outputs d{};
  int st = phobic_server_result(s, id, &np, &B, &used, &d);
  if (st) return st;
  (void)hipSetDevice(s->cfg.device);
  const uint64_t NB = (uint64_t)np * B;
  if (NB && host_out->pilots && hipMemcpy(host_out->pilots, d.pilots, NB * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return PHOBIC_E_HIP;
  if (host_out->offsets &&
      hipMemcpy(host_out->offsets, d.offsets, ((uint64_t)np + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return PHOBIC_E_HIP;
  if (d.remap && host_out->remap_offsets &&
      hipMemcpy(host_out->remap_offsets, d.remap_offsets, ((uint64_t)np + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return PHOBIC_E_HIP;
  if (d.remap && host_out->remap && used &&
      hipMemcpy(host_out->remap, d.remap, used * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return PHOBIC_E_HIP;
  return PHOBIC_OK;
}

int phobic_server_release(phobic_server* s, uint64_t id) {
  if (!s) return PHOBIC_E_INVALID_CONFIG;
  std::shared_ptr<Block> last;  // the block this release empties
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto r = find(s, id);
    if (!r) return PHOBIC_E_NOT_FOUND;
    if (r->state != 2) return PHOBIC_E_BUSY;
    s->all.erase(id);
    if (r->status == PHOBIC_OK && r->block && --r->block->refs == 0) last = r->block;
    r->block.reset();
  }
  if (last) {
    // freed before it leaves the count, so the count never falls below
    // what the pool holds
    (void)hipSetDevice(s->cfg.device);
    (void)phobic::dev_free(last->ptr, s->free_stream);
    (void)hipStreamSynchronize(s->free_stream);
    last->ptr = nullptr;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (last) s->kept -= last->bytes;
    try_admit(s);
  }
  return PHOBIC_OK;
}

int phobic_server_stats_get(phobic_server* s, phobic_server_stats* out) {
  if (!s || !out) return PHOBIC_E_INVALID_CONFIG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->keys_built = s->keys_built;
  out->hbm_in_use = in_use(s);
  out->grouped_launches = s->grouped_launches;
  out->grouped_requests = s->grouped_requests;
  (void)hipSetDevice(s->cfg.device);
  out->pool_used = phobic::pool_used_bytes();
  std::vector<double> v;
  try {
    v = s->latencies_ms;
  } catch (...) {
    return PHOBIC_E_OOM;
  }
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const size_t i = (size_t)std::min<double>((double)v.size() - 1, std::ceil(q * (double)v.size()) - 1);
      return v[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = v.back();
    double sum = 0;
    for (double x : v) sum += x;
    out->mean_ms = sum / (double)v.size();
    out->span_s = std::chrono::duration<double>(s->t_last - s->t_first).count();
  }
  return PHOBIC_OK;
}

int phobic_server_destroy(phobic_server* s) {
  if (!s) return PHOBIC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& sl : s->slots)
    if (sl.th.joinable()) sl.th.join();
  (void)hipSetDevice(s->cfg.device);
  std::vector<Block*> freed;
  for (auto& kv : s->all) {
    Block* b = kv.second->block.get();
    if (b && b->ptr && std::find(freed.begin(), freed.end(), b) == freed.end()) {
      (void)phobic::dev_free(b->ptr, s->free_stream);
      b->ptr = nullptr;
      freed.push_back(b);
    }
  }
  (void)hipStreamSynchronize(s->free_stream);
  (void)hipStreamDestroy(s->free_stream);
  for (auto& sl : s->slots)
    if (sl.ctx) phobic_ctx_destroy(sl.ctx);
  delete s;
  return PHOBIC_OK;
}

}  // extern "C"
