// server.cpp -- continuous-batching build server (no counterpart in the
// reference, which builds one MPHF per synchronous MPHFbuilder::build call,
// mphf_builder.h:315-321).
//
// Requests are queued; a scheduler admits them in arrival order, letting a
// later request overtake one that does not fit yet (backfill), whenever a
// worker slot is idle and the request's device-memory estimate fits the HBM
// budget.  Every slot owns a phobic_ctx (its own stream and scratch pool), so
// admitted builds run concurrently on the device; results stay in device
// memory until fetched or released.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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

struct Request {
  uint64_t id = 0;
  phobic_request req{};
  phobic_config cfg{};
  std::vector<uint32_t> fulcs;  // copy of a custom table
  uint32_t np = 0, B = 0;
  uint64_t remap_cap = 0;
  uint64_t est_bytes = 0;      // admission estimate (inputs + scratch + outputs)
  uint64_t keep_bytes = 0;     // outputs kept on the device after completion
  Clock::time_point t_submit, t_start, t_done;
  int state = 0;               // 0 queued, 1 running, 2 done
  int status = PHOBIC_OK;
  // device state
  void* d_in = nullptr;        // staged keys (host requests) or null
  void* d_out = nullptr;       // pilots | offsets | remap | remap_offsets | status
  phobic_outputs outs{};
  uint32_t* d_status = nullptr;
};

struct Slot {
  phobic_ctx* ctx = nullptr;
  std::shared_ptr<Request> job;
  std::thread th;
};

uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

}  // namespace

struct phobic_server {
  phobic_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  std::vector<Slot> slots;
  hipStream_t free_stream = nullptr;  // releases of kept outputs
  uint64_t next_id = 1;
  uint64_t mem_in_use = 0;
  bool stop = false;
  // statistics
  std::vector<double> latencies_ms;
  uint64_t submitted = 0, completed = 0, rejected = 0, keys_built = 0;
  Clock::time_point t_first{}, t_last{};
  bool any = false;
};

namespace {

// Device bytes a request may need: inputs, the build's scratch (about 56
// bytes per key for partitioning plus the search slice) and outputs.
void estimate(Request& r) {
  const uint64_t n = r.req.n;
  const uint64_t in = r.req.kind == PHOBIC_KEYS_BYTES ? (r.req.keys_on_device ? 0 : r.req.bytes_len + (n + 1) * 8)
                                                      : (r.req.keys_on_device ? 0 : n * 8);
  const uint64_t NB = (uint64_t)r.np * r.B;
  r.keep_bytes = align256(NB * 4) + align256(((uint64_t)r.np + 1) * 4) + align256(r.remap_cap * 4) +
                 align256(((uint64_t)r.np + 1) * 4) + 256;
  const uint64_t cap = r.cfg.partition_size + r.cfg.partition_size / 2;
  const uint64_t scratch = (uint64_t)(r.np + 128) * cap * 24 + (uint64_t)r.np * 16;
  r.est_bytes = in + scratch + r.keep_bytes;
}

void try_admit(phobic_server* s) {
  // caller holds s->mu
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    Slot* idle = nullptr;
    for (auto& sl : s->slots)
      if (!sl.job) {
        idle = &sl;
        break;
      }
    if (!idle) return;
    auto& r = *it;
    if (s->mem_in_use + r->est_bytes <= s->cfg.hbm_budget || s->mem_in_use == 0) {
      s->mem_in_use += r->est_bytes;
      r->state = 1;
      idle->job = r;
      it = s->queue.erase(it);
      s->cv_work.notify_all();
    } else {
      ++it;  // backfill: a smaller request behind it may fit
    }
  }
}

int run_request(phobic_ctx* ctx, Request& r, hipStream_t stream) {
  const uint64_t n = r.req.n, NB = (uint64_t)r.np * r.B;
  // outputs kept on the device until released
  if (phobic::dev_alloc(&r.d_out, r.keep_bytes, stream) != hipSuccess) return PHOBIC_E_OOM;
  uint8_t* p = static_cast<uint8_t*>(r.d_out);
  r.outs.pilots = reinterpret_cast<uint32_t*>(p);
  p += align256(NB * 4);
  r.outs.offsets = reinterpret_cast<uint32_t*>(p);
  p += align256(((uint64_t)r.np + 1) * 4);
  r.outs.remap = r.remap_cap ? reinterpret_cast<uint32_t*>(p) : nullptr;
  p += align256(r.remap_cap * 4);
  r.outs.remap_offsets = r.remap_cap ? reinterpret_cast<uint32_t*>(p) : nullptr;
  p += align256(((uint64_t)r.np + 1) * 4);
  r.d_status = reinterpret_cast<uint32_t*>(p);
  const void* keys = r.req.keys;
  const uint64_t* offs = r.req.offsets;
  if (!r.req.keys_on_device) {  // stage host keys
    const uint64_t kb = r.req.kind == PHOBIC_KEYS_BYTES ? r.req.bytes_len : n * 8;
    const uint64_t ob = r.req.kind == PHOBIC_KEYS_BYTES ? (n + 1) * 8 : 0;
    if (phobic::dev_alloc(&r.d_in, align256(kb) + ob + 256, stream) != hipSuccess) return PHOBIC_E_OOM;
    uint8_t* q = static_cast<uint8_t*>(r.d_in);
    if (kb && hipMemcpyAsync(q, r.req.keys, kb, hipMemcpyHostToDevice, stream) != hipSuccess) return PHOBIC_E_HIP;
    if (ob && hipMemcpyAsync(q + align256(kb), r.req.offsets, ob, hipMemcpyHostToDevice, stream) != hipSuccess)
      return PHOBIC_E_HIP;
    keys = q;
    offs = reinterpret_cast<const uint64_t*>(q + align256(kb));
  }
  int st;
  if (r.req.kind == PHOBIC_KEYS_BYTES)
    st = phobic_build_bytes_device(ctx, static_cast<const uint8_t*>(keys), offs, n, &r.cfg, &r.outs, r.d_status,
                                   stream);
  else
    st = phobic_build_u64_device(ctx, static_cast<const uint64_t*>(keys), n, &r.cfg, &r.outs, r.d_status, stream);
  if (st) return st;
  uint32_t bits = 0;
  if (hipMemcpyAsync(&bits, r.d_status, 4, hipMemcpyDeviceToHost, stream) != hipSuccess) return PHOBIC_E_HIP;
  if (hipStreamSynchronize(stream) != hipSuccess) return PHOBIC_E_HIP;
  if (r.d_in) {
    (void)phobic::dev_free(r.d_in, stream);
    r.d_in = nullptr;
  }
  for (int c = 1; c < 32; c++)
    if (bits & (1u << c)) return c;
  return PHOBIC_OK;
}

void worker(phobic_server* s, size_t idx, hipStream_t stream) {
  (void)hipSetDevice(s->cfg.device);
  for (;;) {
    std::shared_ptr<Request> job;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || s->slots[idx].job; });
      if (s->stop && !s->slots[idx].job) return;
      job = s->slots[idx].job;
      job->t_start = Clock::now();
    }
    const int st = run_request(s->slots[idx].ctx, *job, stream);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      job->status = st;
      job->t_done = Clock::now();
      job->state = 2;
      s->slots[idx].job.reset();
      // scratch and staging returned; outputs stay until released
      s->mem_in_use -= job->est_bytes - (st == PHOBIC_OK ? job->keep_bytes : 0);
      if (st != PHOBIC_OK && job->d_out) {
        (void)phobic::dev_free(job->d_out, stream);
        job->d_out = nullptr;
      }
      if (job->d_in) {  // staging left behind by a failed request
        (void)phobic::dev_free(job->d_in, stream);
        job->d_in = nullptr;
      }
      s->completed++;
      if (st == PHOBIC_OK) s->keys_built += job->req.n;
      try {  // a sample that cannot be stored is dropped from the statistics
        s->latencies_ms.push_back(std::chrono::duration<double, std::milli>(job->t_done - job->t_submit).count());
      } catch (...) {
      }
      s->t_last = job->t_done;
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
    const int st = phobic_ctx_create(cfg->device, &sl.ctx);
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
  if (req->kind == PHOBIC_KEYS_BYTES && !req->offsets) return PHOBIC_E_INVALID_CONFIG;
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
  estimate(*r);
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->est_bytes > s->cfg.hbm_budget) {
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
  }
  if (!r) return PHOBIC_E_NOT_FOUND;
  if (r->state != 2) return PHOBIC_E_BUSY;
  if (r->status != PHOBIC_OK) return r->status;
  if (np) *np = r->np;
  if (buckets) *buckets = r->B;
  if (remap_entries) {
    uint32_t used = 0;
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
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r = find(s, id);
    if (!r) return PHOBIC_E_NOT_FOUND;
    if (r->state != 2) return PHOBIC_E_BUSY;
    s->all.erase(id);
    if (r->d_out) s->mem_in_use -= r->keep_bytes;
  }
  (void)hipSetDevice(s->cfg.device);
  if (r->d_out) (void)phobic::dev_free(r->d_out, s->free_stream);
  r->d_out = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
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
  out->hbm_in_use = s->mem_in_use;
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
  for (auto& kv : s->all) {
    if (kv.second->d_out) (void)phobic::dev_free(kv.second->d_out, s->free_stream);
    if (kv.second->d_in) (void)phobic::dev_free(kv.second->d_in, s->free_stream);
  }
  (void)hipStreamSynchronize(s->free_stream);
  (void)hipStreamDestroy(s->free_stream);
  for (auto& sl : s->slots)
    if (sl.ctx) phobic_ctx_destroy(sl.ctx);
  delete s;
  return PHOBIC_OK;
}

}  // extern "C"
