// Served request loop (gss_server_* in gss.h): continuous batching of sort
// requests over one device-resident text.
//
// Requests queue in arrival order.  Each slot is a host thread owning a sort
// context (its own stream and workspace) and batch buffers sized from the
// HBM budget.  A free slot takes queued requests in order while they fit its
// capacity, gathers their offset arrays into one batch array with one
// kernel, sorts the batch (every request stays its own segment, so each is
// sorted exactly as alone), scatters the sorted offsets and LCPs back with
// one kernel and completes the requests.  Several slots run at once, so
// requests that arrive while batches are on the GPU are taken up as soon as a
// slot finishes, and the slots' kernels overlap on the device.
//
// Stream order: a request may name a HIP event whose work must finish before
// its offsets are read.  The event's state at submission is captured on the
// server's own submit stream (wait on it, then record a pooled event), so a
// later re-record of the caller's event does not move the dependency; the
// slot stream waits for the pooled event before its gather.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gss.h"
#include "gss_internal.h"
#include "host.h"

namespace gss {
namespace {

using Clock = std::chrono::steady_clock;

constexpr uint32_t kMaxSlots = 8;

struct Request {
  const uint32_t* in = nullptr;
  uint32_t* out = nullptr;
  uint64_t* lcp = nullptr;
  uint64_t n = 0;
  hipEvent_t ready = nullptr;  // pooled event captured at submission, or none
  Clock::time_point submitted;
  double latency_ms = 0;
  int status = GSS_OK;
  bool done = false;
  std::string error;
};

struct Slot {
  gss_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  char* dev = nullptr;            // one allocation for the arrays below
  uint32_t* d_in = nullptr;       // gathered request offsets
  uint32_t* d_out = nullptr;      // sorted batch
  uint64_t* d_lcp = nullptr;      // batch LCPs
  CopyChunk* d_chunks = nullptr;  // gather chunks, then scatter chunks
  CopyChunk* h_chunks = nullptr;  // pinned staging of the chunk table
  uint64_t chunk_cap = 0;
  std::vector<uint64_t> req;      // request table of the current batch
  std::vector<Request*> batch;
  std::thread thread;
};

uint64_t chunk_cap_for(uint64_t cap, uint32_t max_reqs) {
  // Offsets in, offsets out, LCP words out (two per entry): each request adds
  // at most one partial chunk to each.
  return 4 * (cap / kChunkWords) + 3ull * max_reqs + 8;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

size_t slot_buffer_bytes(uint64_t cap, uint32_t max_reqs) {
  return align256(4 * cap) + align256(4 * cap) + align256(8 * cap) +
         align256(sizeof(CopyChunk) * chunk_cap_for(cap, max_reqs));
}

size_t slot_bytes(uint64_t cap, uint32_t max_reqs, const gss_tuning& tun) {
  return workspace_bytes(cap, tun) + slot_buffer_bytes(cap, max_reqs);
}

double percentile(std::vector<double>& v, double p) {
  if (v.empty()) return 0;
  // Nearest rank: the smallest value with at least p% of the samples <= it.
  size_t rank = (size_t)std::ceil(p / 100.0 * (double)v.size());
  if (rank < 1) rank = 1;
  std::nth_element(v.begin(), v.begin() + (rank - 1), v.end());
  return v[rank - 1];
}

}  // namespace
}  // namespace gss

struct gss_server {
  int device = -1;
  const unsigned char* text = nullptr;
  uint64_t text_bytes = 0;
  gss_server_config cfg{};
  gss_tuning tuning{};  // the tuning current at creation: every batch uses it
  uint64_t cap = 0;     // strings one slot's batch may hold
  hipStream_t submit_stream = nullptr;  // captures the requests' ready events
  std::vector<hipEvent_t> free_events;  // pooled; guarded by mu
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<gss::Request*> queue;
  std::unordered_map<int64_t, gss::Request*> reqs;
  int64_t next_id = 0;
  bool stopping = false;
  std::vector<std::unique_ptr<gss::Slot>> slots;
  uint64_t submitted = 0, completed = 0, batches = 0, strings = 0;
  uint32_t max_batch_seen = 0;
  std::vector<double> latencies;
};

namespace gss {
namespace {

#define SRV_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(GSS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Queues the chunk-table upload, the gather, the sort and the scatter of a
// batch whose chunk table (g entries, the first ngather of them the gather)
// is in sl.h_chunks.
int launch_batch(gss_server* s, Slot& sl, uint64_t g, uint64_t ngather, bool want_lcp) {
  UploadBatch up;  // the chunk table, read by a kernel from coherent pinned memory
  if (add_upload(up, sl.h_chunks, sl.d_chunks, 0, sizeof(CopyChunk) * g))
    return fail(GSS_E_INTERNAL, "batch chunk table upload");
  SRV_HIP(launch_upload(up, sl.stream));
  SRV_HIP(launch_copy_chunks(sl.d_chunks, (uint32_t)ngather, sl.stream));
  const int rc = ctx_sort_batch_tuned(sl.ctx, s->text, s->text_bytes, sl.d_in, sl.d_out,
                                      want_lcp ? sl.d_lcp : nullptr, sl.req.data(),
                                      (uint32_t)sl.batch.size(), sl.stream, s->tuning);
  if (rc) return rc;
  SRV_HIP(launch_copy_chunks(sl.d_chunks + ngather, (uint32_t)(g - ngather), sl.stream));
  return GSS_OK;
}

// Gathers, sorts and scatters the slot's batch; returns its status.
int run_batch(gss_server* s, Slot& sl) {
  sl.req.assign(1, 0);
  uint64_t g = 0;
  bool want_lcp = false;
  for (Request* r : sl.batch) {
    const uint64_t base = sl.req.back();
    for (uint64_t k = 0; k < r->n; k += kChunkWords)
      sl.h_chunks[g++] =
          CopyChunk{r->in + k, sl.d_in + base + k, std::min<uint64_t>(kChunkWords, r->n - k)};
    sl.req.push_back(base + r->n);
    want_lcp |= r->lcp != nullptr && r->n > 1;
  }
  const uint64_t ngather = g;
  for (size_t i = 0; i < sl.batch.size(); ++i) {
    const Request* r = sl.batch[i];
    const uint64_t base = sl.req[i];
    for (uint64_t k = 0; k < r->n; k += kChunkWords)
      sl.h_chunks[g++] =
          CopyChunk{sl.d_out + base + k, r->out + k, std::min<uint64_t>(kChunkWords, r->n - k)};
    if (r->lcp && r->n > 1) {
      const uint64_t words = 2 * (r->n - 1);
      const uint32_t* src = reinterpret_cast<const uint32_t*>(sl.d_lcp + base);
      uint32_t* dst = reinterpret_cast<uint32_t*>(r->lcp);
      for (uint64_t k = 0; k < words; k += kChunkWords)
        sl.h_chunks[g++] = CopyChunk{src + k, dst + k, std::min<uint64_t>(kChunkWords, words - k)};
    }
  }
  if (g > sl.chunk_cap) return fail(GSS_E_INTERNAL, "batch chunk table overflow");
  for (const Request* r : sl.batch)
    if (r->ready) SRV_HIP(hipStreamWaitEvent(sl.stream, r->ready, 0));
  // From the first launch on, an error still drains the stream before it is
  // reported: queued copies read the chunk table and the requests' arrays.
  const int rc = launch_batch(s, sl, g, ngather, want_lcp);
  if (rc) {
    (void)hipStreamSynchronize(sl.stream);
    return rc;
  }
  return wait_stream(sl.stream);
}

void complete(gss_server* s, Slot& sl, int rc) {
  const Clock::time_point now = Clock::now();
  const std::string msg = rc ? std::string(gss_last_error()) : std::string();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request* r : sl.batch) {
      if (r->ready) {  // the batch's stream work is over: the event is free
        s->free_events.push_back(r->ready);
        r->ready = nullptr;
      }
      r->status = rc;
      r->error = msg;
      r->latency_ms = std::chrono::duration<double, std::milli>(now - r->submitted).count();
      r->done = true;
      s->completed += 1;
      s->strings += r->n;
      s->latencies.push_back(r->latency_ms);
    }
    s->batches += 1;
    s->max_batch_seen = std::max<uint32_t>(s->max_batch_seen, (uint32_t)sl.batch.size());
  }
  sl.batch.clear();
  s->cv_done.notify_all();
}

void worker(gss_server* s, Slot* sl) {
  const hipError_t e = hipSetDevice(s->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stopping || !s->queue.empty(); });
      if (s->queue.empty()) return;  // stopping, nothing left
      uint64_t total = 0;
      while (!s->queue.empty() && sl->batch.size() < s->cfg.max_batch_requests) {
        Request* r = s->queue.front();
        if (total + r->n > s->cap) break;
        total += r->n;
        sl->batch.push_back(r);
        s->queue.pop_front();
      }
      if (!s->queue.empty()) s->cv_work.notify_one();  // more work for another slot
    }
    int rc = GSS_OK;
    if (e != hipSuccess)
      rc = fail(GSS_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    else
      rc = run_batch(s, *sl);
    complete(s, *sl, rc);
  }
}

void free_slot(Slot& sl) {
  if (sl.ctx) gss_ctx_destroy(sl.ctx);
  if (sl.dev) (void)hipFree(sl.dev);
  if (sl.h_chunks) (void)hipHostFree(sl.h_chunks);
  sl.ctx = nullptr;
  sl.dev = nullptr;
  sl.h_chunks = nullptr;
}

int init_slot(Slot& sl, uint64_t cap, uint32_t max_reqs, const gss_tuning& tun) {
  sl.ctx = gss_ctx_create();
  if (!sl.ctx) return GSS_E_DEVICE;  // message set by gss_ctx_create
  sl.stream = reinterpret_cast<hipStream_t>(gss_ctx_stream(sl.ctx));
  int rc = ctx_reserve(sl.ctx, cap, max_reqs, tun);
  if (rc) return rc;
  sl.chunk_cap = chunk_cap_for(cap, max_reqs);
  const size_t bytes = slot_buffer_bytes(cap, max_reqs);
  if (hipMalloc((void**)&sl.dev, bytes) != hipSuccess) {
    (void)hipGetLastError();
    sl.dev = nullptr;
    return fail(GSS_E_NOMEM, "server batch buffers of " + std::to_string(bytes) +
                                 " bytes: allocation failed");
  }
  char* p = sl.dev;
  sl.d_in = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * cap);
  sl.d_out = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * cap);
  sl.d_lcp = reinterpret_cast<uint64_t*>(p);
  p += align256(8 * cap);
  sl.d_chunks = reinterpret_cast<CopyChunk*>(p);
  SRV_HIP(hipHostMalloc((void**)&sl.h_chunks, sizeof(CopyChunk) * sl.chunk_cap,
                        hipHostMallocCoherent));
  sl.req.reserve(max_reqs + 1);
  sl.batch.reserve(max_reqs);
  return GSS_OK;
}

void shutdown(gss_server* s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (auto& sl : s->slots)
    if (sl->thread.joinable()) sl->thread.join();
  for (auto& sl : s->slots) free_slot(*sl);
  for (auto& kv : s->reqs) {
    if (kv.second->ready) s->free_events.push_back(kv.second->ready);
    delete kv.second;
  }
  s->reqs.clear();
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  s->free_events.clear();
  if (s->submit_stream) (void)hipStreamDestroy(s->submit_stream);
  s->submit_stream = nullptr;
}

}  // namespace
}  // namespace gss

using namespace gss;

extern "C" {

gss_server* gss_server_create(const unsigned char* d_text, uint64_t text_bytes,
                              const gss_server_config* cfg) {
  if (!cfg) {
    fail(GSS_E_INVALID, "NULL server configuration");
    return nullptr;
  }
  if (!d_text || (reinterpret_cast<uintptr_t>(d_text) & 7)) {
    fail(GSS_E_INVALID, "the server text must be a non-NULL, 8-byte aligned device pointer");
    return nullptr;
  }
  if (cfg->slots < 1 || cfg->slots > kMaxSlots) {
    fail(GSS_E_INVALID, "slots must lie in [1, 8]");
    return nullptr;
  }
  if (cfg->max_batch_requests < 1 || cfg->max_batch_requests > (1u << 20)) {
    fail(GSS_E_INVALID, "max_batch_requests must lie in [1, 2^20]");
    return nullptr;
  }
  if (text_bytes == 0) {
    fail(GSS_E_INVALID, "empty server text");
    return nullptr;
  }
  // Largest per-slot capacity whose buffers fit the budget, under the tuning
  // current now (kept for the server's lifetime).
  const gss_tuning tun = current_tuning();
  const uint64_t per_slot = cfg->hbm_budget / cfg->slots;
  uint64_t lo = 0, hi = (1ull << 31) - 1;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo + 1) / 2;
    if (slot_bytes(mid, cfg->max_batch_requests, tun) <= per_slot) lo = mid; else hi = mid - 1;
  }
  if (lo == 0) {
    fail(GSS_E_NOMEM, "hbm_budget is too small for one slot's workspace (" +
                          std::to_string(slot_bytes(1, cfg->max_batch_requests, tun)) +
                          " bytes per slot needed)");
    return nullptr;
  }
  gss_server* s = new gss_server;
  s->text = d_text;
  s->text_bytes = text_bytes;
  s->cfg = *cfg;
  s->tuning = tun;
  s->cap = lo;
  if (hipGetDevice(&s->device) != hipSuccess ||
      hipStreamCreateWithFlags(&s->submit_stream, hipStreamNonBlocking) != hipSuccess) {
    fail(GSS_E_DEVICE, "hipGetDevice / hipStreamCreate failed");
    s->submit_stream = nullptr;
    delete s;
    return nullptr;
  }
  for (uint32_t i = 0; i < cfg->slots; ++i) {
    s->slots.emplace_back(new Slot);
    if (init_slot(*s->slots.back(), s->cap, cfg->max_batch_requests, s->tuning)) {
      const std::string msg = gss_last_error();
      shutdown(s);
      delete s;
      fail(GSS_E_NOMEM, "server slot setup: " + msg);
      return nullptr;
    }
  }
  for (auto& sl : s->slots) sl->thread = std::thread(worker, s, sl.get());
  return s;
}

int64_t gss_server_submit(gss_server* s, const uint32_t* d_off, uint32_t* d_out, uint64_t* d_lcp,
                          uint64_t n, void* ready_event) {
  if (!s) return fail(GSS_E_INVALID, "NULL server");
  if (n && (!d_off || !d_out)) return fail(GSS_E_INVALID, "NULL offsets or output");
  if (n > s->cap)
    return fail(GSS_E_TOO_LARGE, "request of " + std::to_string(n) +
                                     " strings exceeds the batch capacity of " +
                                     std::to_string(s->cap));
  Request* r = new Request;
  r->in = d_off;
  r->out = d_out;
  r->lcp = d_lcp;
  r->n = n;
  r->submitted = Clock::now();
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) {
      delete r;
      return fail(GSS_E_INVALID, "server is shutting down");
    }
    if (ready_event && n) {
      // The caller's event as recorded now: the submit stream waits for it
      // and a pooled event records that point.
      hipEvent_t ev = nullptr;
      if (!s->free_events.empty()) {
        ev = s->free_events.back();
        s->free_events.pop_back();
      } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        delete r;
        return fail(GSS_E_DEVICE, "hipEventCreate failed");
      }
      const hipError_t e1 =
          hipStreamWaitEvent(s->submit_stream, reinterpret_cast<hipEvent_t>(ready_event), 0);
      const hipError_t e2 = e1 == hipSuccess ? hipEventRecord(ev, s->submit_stream) : e1;
      if (e2 != hipSuccess) {
        s->free_events.push_back(ev);
        delete r;
        return fail(GSS_E_DEVICE, std::string("capturing the ready event: ") +
                                      hipGetErrorString(e2));
      }
      r->ready = ev;
    }
    id = s->next_id++;
    s->reqs.emplace(id, r);
    s->submitted += 1;
    if (n == 0) {  // nothing to sort
      r->done = true;
      s->completed += 1;
      s->latencies.push_back(0.0);
    } else {
      s->queue.push_back(r);
    }
  }
  if (n) s->cv_work.notify_one();
  return id;
}

int gss_server_wait(gss_server* s, int64_t id, double* latency_ms) {
  if (!s) return fail(GSS_E_INVALID, "NULL server");
  int status;
  std::string err;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(GSS_E_INVALID, "unknown request id (already waited for?)");
    Request* r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    status = r->status;
    err = r->error;
    if (latency_ms) *latency_ms = r->latency_ms;
    s->reqs.erase(it);
    delete r;
  }
  if (status) return fail(status, err);
  return GSS_OK;
}

int gss_server_done(gss_server* s, int64_t id) {
  if (!s) return fail(GSS_E_INVALID, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(GSS_E_INVALID, "unknown request id (already waited for?)");
  return it->second->done ? 1 : 0;
}

int gss_server_get_stats(gss_server* s, gss_server_stats* out) {
  if (!s || !out) return fail(GSS_E_INVALID, "NULL server or stats pointer");
  std::vector<double> lat;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->batches = s->batches;
    out->strings = s->strings;
    out->batch_capacity = s->cap;
    out->max_batch_requests_seen = s->max_batch_seen;
    lat = s->latencies;
  }
  if (!lat.empty()) {
    double sum = 0, mx = 0;
    for (double v : lat) {
      sum += v;
      mx = std::max(mx, v);
    }
    out->mean_ms = sum / (double)lat.size();
    out->max_ms = mx;
    out->p50_ms = percentile(lat, 50.0);
    out->p99_ms = percentile(lat, 99.0);
  }
  return GSS_OK;
}

void gss_server_destroy(gss_server* s) {
  if (!s) return;
  shutdown(s);
  delete s;
}

}  // extern "C"
