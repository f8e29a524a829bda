// Device-resident MSD radix string sort: sort contexts (stream, workspace,
// statistics), the pass loop, profiling, and the C entry points of gss.h
// that do not touch host strings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gss.h"
#include "gss_internal.h"
#include "host.h"

namespace gss {

namespace {

constexpr int kMaxDevices = 16;
constexpr size_t kPubHead = 256;  // the hand-over block's sequence word, padded

struct Workspace {
  char* base = nullptr;
  size_t bytes = 0;
  uint64_t n_cap = 0;
  uint32_t rb = 0;
  uint32_t threshold = 0;
  uint32_t tile = 0;
  Plan plan{};
  uint32_t* idx_copy = nullptr;
  // Pinned host mirrors.
  Counters* h_ctr = nullptr;
  LocalSeg* h_local = nullptr;  // initial local segments, 3 classes of cap_init
  uint32_t cap_init = 0;
  // Fine-grained (coherent) pinned blocks, one per direction.  Up (h_init,
  // h_local): the start of a sort, read by one upload launch.  Down (h_pub):
  // a sequence word (bytes [0, 8)) and from kPubHead on the statistics (local
  // statistics, per-pass active strings) and the counters, written by the
  // publish kernel after a local phase; h_lstats and h_pass point into it.
  char* h_pub = nullptr;
  uint64_t pub_seq = 0;
  // The first pass's segment aux table and tile table, built by the host
  // (adjacent on the device, uploaded with the start of the sort).
  char* h_up = nullptr;
  size_t up_tiles_off = 0;
  const LocalStats* h_lstats = nullptr;
  const unsigned long long* h_pass = nullptr;
  // The start of a sort: the counters and the initial segment list go up in
  // one upload that also clears the statistics (adjacent on the device).
  char* h_init = nullptr;
  size_t init_ctr_off = 0, init_big_off = 0;
};

struct ProfTotal {
  uint64_t launches = 0;
  double ms = 0;
  uint64_t items = 0;
};

struct Pending {
  hipEvent_t a, b;
  std::string name;
  uint64_t items;
};

std::mutex g_mu;  // tuning, profile totals, default contexts
gss_tuning g_tuning = {6144u, 11u, kTile, 64u};
std::atomic<bool> g_prof{false};    // HIP event brackets around every kernel and pass
std::atomic<bool> g_prof_pass{false};  // around passes only
std::atomic<bool> g_levels{false};  // counting variant of the local kernels
std::map<std::string, ProfTotal> g_prof_totals;
thread_local std::string g_last_error;
thread_local gss_stats g_last_stats;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int local_class(uint64_t size) {
  return size <= kLocalClassCap[0] ? 0 : (size <= kLocalClassCap[1] ? 1 : 2);
}

}  // namespace

// One independent sorter: its own stream and workspace.  Calls on one
// context are serialised; different contexts run concurrently.
struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  // The requests of a batch that start as local segments are sorted on
  // this stream while the large ones take their global passes on `stream`.
  hipStream_t side = nullptr;
  hipEvent_t side_done = nullptr;
  Workspace ws;
  // Global passes the last sort on this context took, and the strings its
  // first pass held: a sort of a similar size queues as many passes before
  // it first looks at the counters.
  uint32_t pass_hint = 1;
  uint64_t hint_strings = 0;
  // Segments of the largest local class the last sort on this context
  // sorted on its main stream (a sort of a similar size then launches that
  // class before it reads the counters back, for about as many segments).
  uint32_t hint_largest = 0;
  std::mutex mu;
  gss_stats stats{};
  std::vector<Pending> pending;
  std::vector<hipEvent_t> events;
};

}  // namespace gss

struct gss_ctx {
  gss::Context c;
};

namespace gss {

namespace {

gss_ctx* g_default[kMaxDevices] = {};

#define GSS_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(GSS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Waits until the hand-over block's sequence word reads seq.  The stream is
// queried now and then, so a failed launch ends the wait with its error.
int wait_published(const Workspace& w, hipStream_t st, uint64_t seq) {
  const volatile uint64_t* flag = reinterpret_cast<const volatile uint64_t*>(w.h_pub);
  for (uint32_t spin = 1;; ++spin) {
    if (*flag == seq) break;
    if (spin % 1024 == 0) {
      const hipError_t e = hipStreamQuery(st);
      if (e == hipSuccess) {
        if (*flag == seq) break;
        return fail(GSS_E_INTERNAL, "device hand-over did not arrive");
      }
      if (e != hipErrorNotReady)
        return fail(GSS_E_DEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return GSS_OK;
}

int init_context(Context& c) {
  GSS_HIP(hipGetDevice(&c.device));
  GSS_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  GSS_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
  GSS_HIP(hipEventCreateWithFlags(&c.side_done, hipEventDisableTiming));
  GSS_HIP(hipHostMalloc((void**)&c.ws.h_ctr, sizeof(Counters), hipHostMallocDefault));
  return GSS_OK;
}

int default_context(Context** out) {
  int dev = 0;
  GSS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(GSS_E_DEVICE, "device id out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_default[dev]) {
    gss_ctx* h = new gss_ctx;
    const int rc = init_context(h->c);
    if (rc) {
      delete h;
      return rc;
    }
    g_default[dev] = h;
  }
  *out = &g_default[dev]->c;
  return GSS_OK;
}

hipEvent_t take_event(Context& c) {
  if (!c.events.empty()) {
    hipEvent_t e = c.events.back();
    c.events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct ProfToken {
  hipEvent_t a = nullptr;
  const char* name = nullptr;
  uint64_t items = 0;
};

// A pass-level bracket (a global pass, the local phase) is recorded in both
// profiling modes, a kernel's only with every kernel bracketed.
ProfToken prof_begin(Context& c, const char* name, uint64_t items, hipStream_t stream,
                     bool pass_level = false) {
  ProfToken t;
  if (!g_prof.load(std::memory_order_relaxed) &&
      !(pass_level && g_prof_pass.load(std::memory_order_relaxed)))
    return t;
  t.a = take_event(c);
  t.name = name;
  t.items = items;
  if (t.a) (void)hipEventRecord(t.a, stream);
  return t;
}

void prof_end(Context& c, ProfToken& t, hipStream_t stream) {
  if (!t.a) return;
  hipEvent_t b = take_event(c);
  if (!b) return;
  (void)hipEventRecord(b, stream);
  c.pending.push_back(Pending{t.a, b, t.name, t.items});
}

// Folds the finished brackets into the global totals (after a sync).
void prof_collect(Context& c) {
  if (c.pending.empty()) return;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& p : c.pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      ProfTotal& tot = g_prof_totals[p.name];
      tot.launches += 1;
      tot.ms += ms;
      tot.items += p.items;
    }
    c.events.push_back(p.a);
    c.events.push_back(p.b);
  }
  c.pending.clear();
}

void free_workspace(Workspace& w) {
  if (w.base) (void)hipFree(w.base);
  if (w.h_pub) (void)hipHostFree(w.h_pub);
  if (w.h_init) (void)hipHostFree(w.h_init);
  if (w.h_up) (void)hipHostFree(w.h_up);
  w.h_up = nullptr;
  w.h_init = nullptr;
  w.base = nullptr;
  w.h_pub = nullptr;
  w.h_lstats = nullptr;
  w.h_pass = nullptr;
  w.bytes = 0;
  w.n_cap = 0;
}

// Capacities of the workspace tables for up to ncap strings and local
// threshold t.
struct Caps {
  uint64_t ncap, big, tiles, local[3], retire, bounds;
};

Caps caps_for(uint64_t ncap, uint32_t t, uint32_t tile) {
  Caps k;
  k.ncap = std::max<uint64_t>(ncap, 4096);
  k.big = k.ncap / (t + 1) + 2;
  k.tiles = k.ncap / tile + k.big + 2;
  k.local[0] = k.ncap / 2 + 2;
  k.local[1] = k.ncap / (kLocalClassCap[0] + 1) + 2;
  k.local[2] = k.ncap / (kLocalClassCap[1] + 1) + 2;
  k.retire = k.ncap + k.ncap / kRetireChunk + 8;  // single-string requests too
  k.bounds = k.ncap + 2;
  return k;
}

struct Item {
  void** dst;
  size_t bytes;
};

// The device arrays of a workspace, in allocation order.
std::vector<Item> workspace_items(Plan& p, uint32_t** idx_copy, const Caps& k, uint32_t rb) {
  const uint64_t nb = 1ull << rb;
  return {
      {(void**)&p.idx[1], 4 * k.ncap},
      {(void**)&p.idx[2], 4 * k.ncap},
      {(void**)idx_copy, 4 * k.ncap},
      {(void**)&p.keys, 8 * k.ncap},
      {(void**)&p.keys_next, 8 * k.ncap},
      {(void**)&p.ext, 8 * k.ncap},
      // lstats (+ the per-pass active strings), ctr, big[0]: uploaded
      // together at the start of a sort; lstats and ctr read back together.
      {(void**)&p.lstats, 2 * sizeof(LocalStats) + 8 * kMaxPasses},
      {(void**)&p.ctr, sizeof(Counters)},
      {(void**)&p.big[0], sizeof(BigSeg) * k.big},
      {(void**)&p.big[1], sizeof(BigSeg) * k.big},
      {(void**)&p.segmm, 16 * k.big},
      {(void**)&p.bstart, 4 * k.big * nb},
      {(void**)&p.alpha, 32},
      {(void**)&p.aux, sizeof(SegAux) * k.big},
      {(void**)&p.tiles, sizeof(Tile) * k.tiles},
      {(void**)&p.cnt, 2 * (k.tiles * nb + 8)},
      {(void**)&p.flat, 4 * (k.tiles * nb + 1)},
      {(void**)&p.flat_blk, 4 * ((k.tiles * nb) / kScanBlock + 2)},
      {(void**)&p.local[0], sizeof(LocalSeg) * k.local[0]},
      {(void**)&p.local[1], sizeof(LocalSeg) * k.local[1]},
      {(void**)&p.local[2], sizeof(LocalSeg) * k.local[2]},
      {(void**)&p.retire, sizeof(Retire) * k.retire},
      {(void**)&p.bounds, sizeof(Boundary) * k.bounds},
      {(void**)&p.tile_mm, 16 * k.tiles},
      {(void**)&p.tile_alpha, 32 * k.tiles},
      {(void**)&p.seg_alpha, 32 * k.big},
      {(void**)&p.guess, 4 * kGuessWords},
  };
}

size_t items_bytes(const std::vector<Item>& items) {
  size_t total = 0;
  for (auto& it : items) total += align_up(it.bytes, 256);
  return total;
}

// Sizes every device array for up to n strings, 2^rb digits, local
// threshold t and tiles of `tile` strings, and carves them out of one
// allocation.
int ensure_workspace(Context& c, uint64_t n, uint32_t rb, uint32_t t, uint32_t tile) {
  Workspace& w = c.ws;
  if (w.base && n <= w.n_cap && rb == w.rb && t == w.threshold && tile == w.tile) return GSS_OK;
  GSS_HIP(hipStreamSynchronize(c.stream));
  const Caps k = caps_for(std::max<uint64_t>(n, w.n_cap), t, tile);
  free_workspace(w);
  const uint64_t ncap = k.ncap, cap_big = k.big, cap_tiles = k.tiles, cap_retire = k.retire,
                 cap_bounds = k.bounds;
  const uint64_t* cap_local = k.local;
  Plan& p = w.plan;
  p = Plan{};
  std::vector<Item> items = workspace_items(p, &w.idx_copy, k, rb);
  const size_t total = items_bytes(items);
  char* base = nullptr;
  if (hipMalloc((void**)&base, total) != hipSuccess) {
    (void)hipGetLastError();
    return fail(GSS_E_NOMEM, "device workspace allocation of " + std::to_string(total) +
                                 " bytes failed");
  }
  size_t off = 0;
  for (auto& it : items) {
    *it.dst = base + off;
    off += align_up(it.bytes, 256);
  }
  p.pass_active = reinterpret_cast<unsigned long long*>(p.lstats + 2);
  if (hipMemset(p.guess, 0, 4 * kGuessWords) != hipSuccess) {  // no guess yet
    (void)hipGetLastError();
    (void)hipFree(base);
    return fail(GSS_E_INTERNAL, "device workspace initialisation failed");
  }
  w.base = base;
  w.bytes = total;
  w.n_cap = ncap;
  w.rb = rb;
  w.threshold = t;
  w.tile = tile;
  p.cap_big = (uint32_t)cap_big;
  p.cap_tiles = (uint32_t)cap_tiles;
  p.cap_retire = (uint32_t)cap_retire;
  p.cap_bounds = (uint32_t)cap_bounds;
  for (int i = 0; i < 3; ++i) p.cap_local[i] = (uint32_t)cap_local[i];
  w.up_tiles_off = (size_t)(reinterpret_cast<char*>(p.tiles) - reinterpret_cast<char*>(p.aux));
  GSS_HIP(hipHostMalloc((void**)&w.h_up, w.up_tiles_off + sizeof(Tile) * cap_tiles,
                        hipHostMallocCoherent));
  w.init_ctr_off = (size_t)(reinterpret_cast<char*>(p.ctr) - reinterpret_cast<char*>(p.lstats));
  const size_t pub_bytes = kPubHead + w.init_ctr_off + sizeof(Counters);
  GSS_HIP(hipHostMalloc((void**)&w.h_pub, pub_bytes, hipHostMallocCoherent));
  memset(w.h_pub, 0, pub_bytes);
  w.h_lstats = reinterpret_cast<const LocalStats*>(w.h_pub + kPubHead);
  w.h_pass = reinterpret_cast<const unsigned long long*>(w.h_lstats + 2);
  w.init_big_off = (size_t)(reinterpret_cast<char*>(p.big[0]) - reinterpret_cast<char*>(p.lstats));
  const size_t init_bytes = w.init_big_off + sizeof(BigSeg) * cap_big;
  GSS_HIP(hipHostMalloc((void**)&w.h_init, init_bytes, hipHostMallocCoherent));
  memset(w.h_init, 0, init_bytes);  // the statistics part is cleared on the device, never read
  return GSS_OK;
}

int ensure_init_lists(Workspace& w, uint32_t nreq) {
  if (nreq <= w.cap_init) return GSS_OK;
  if (w.h_local) (void)hipHostFree(w.h_local);
  w.h_local = nullptr;
  const uint32_t cap = std::max<uint32_t>(nreq, 64);
  // Three classes of initial local segments, then the initial retire runs.
  static_assert(sizeof(Retire) <= 2 * sizeof(LocalSeg), "layout of the initial lists");
  GSS_HIP(hipHostMalloc((void**)&w.h_local, 5 * sizeof(LocalSeg) * cap, hipHostMallocCoherent));
  w.cap_init = cap;
  return GSS_OK;
}

#define GSS_LAUNCH(name, items, call)                \
  do {                                               \
    ProfToken tok_ = prof_begin(c, name, items, st); \
    GSS_HIP(call);                                   \
    prof_end(c, tok_, st);                           \
  } while (0)

// Sorts the requests [req[r], req[r+1]) of the offset array independently
// (req == nullptr: one request of n strings).  req_depth (optional): the
// bytes every string of request r shares, where its sort starts, or
// kSpanKeep for a request left as it is (with out == in its offsets and
// LCPs are not touched).  Called with c.mu held.
int run_sort(Context& c, const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
             uint32_t* out, uint64_t* lcp, uint64_t n, const uint64_t* req, uint32_t nreq,
             hipStream_t stream, const gss_tuning* tuning, const uint32_t* req_depth = nullptr) {
  gss_stats& stats = c.stats;
  memset(&stats, 0, sizeof(stats));
  stats.n = n;
  stats.text_bytes = text_bytes;
  if (n == 0) return GSS_OK;
  if (!text || !in || !out) return fail(GSS_E_INVALID, "NULL text, input or output pointer");
  if (reinterpret_cast<uintptr_t>(text) & 7)
    return fail(GSS_E_INVALID, "the device text must be 8-byte aligned");
  if (text_bytes == 0) return fail(GSS_E_INVALID, "empty text: every string starts inside it");
  if (n >= (1ull << 31))
    return fail(GSS_E_TOO_LARGE, "at most 2^31-1 strings per sort on the 32-bit path");
  int dev = -1;
  GSS_HIP(hipGetDevice(&dev));
  if (dev != c.device) return fail(GSS_E_INVALID, "the context belongs to another device");
  const uint64_t one[2] = {0, n};
  if (!req) {
    req = one;
    nreq = 1;
  }
  if (req[0] != 0 || req[nreq] != n) return fail(GSS_E_INVALID, "requests must cover [0, n)");
  for (uint32_t r = 0; r < nreq; ++r)
    if (req[r + 1] < req[r]) return fail(GSS_E_INVALID, "request bounds must not decrease");
  hipStream_t st = stream ? stream : c.stream;

  gss_tuning tun;
  if (tuning) {
    tun = *tuning;
  } else {
    std::lock_guard<std::mutex> lk(g_mu);
    tun = g_tuning;
  }
  const uint32_t rb = tun.radix_bits;
  const uint32_t T = tun.local_threshold;
  const uint32_t tile = tun.tile;
  int rc = ensure_workspace(c, n, rb, T, tile);
  if (rc) return rc;
  rc = ensure_init_lists(c.ws, nreq);
  if (rc) return rc;
  Workspace& w = c.ws;
  Plan p = w.plan;
  p.text64 = reinterpret_cast<const uint64_t*>(text);
  p.text_bytes = text_bytes;
  p.out = out;
  p.lcp = lcp;
  p.local_threshold = T;
  p.tile = tile;
  p.key_bytes = tun.key_bytes;
  p.idx[0] = const_cast<uint32_t*>(in);
  if (out == in) {
    GSS_HIP(hipMemcpyAsync(w.idx_copy, in, 4 * n, hipMemcpyDeviceToDevice, st));
    p.idx[0] = w.idx_copy;
  }
  const uint32_t nb = 1u << rb;

  // Initial segments: one per request, finished locally when small enough;
  // a single string is final as it is (a retire run of one, nothing to fill).
  Counters& h0 = *w.h_ctr;
  memset(&h0, 0, sizeof(h0));
  LocalSeg* init[3] = {w.h_local, w.h_local + w.cap_init, w.h_local + 2 * w.cap_init};
  Retire* init_retire = reinterpret_cast<Retire*>(w.h_local + 3 * w.cap_init);
  BigSeg* init_big = reinterpret_cast<BigSeg*>(w.h_init + w.init_big_off);
  uint32_t nbig = 0;
  uint32_t tiles0 = 0;  // tiles of the first pass
  uint64_t pass0_strings = 0;
  SegAux* aux0 = reinterpret_cast<SegAux*>(w.h_up);
  Tile* tiles = reinterpret_cast<Tile*>(w.h_up + w.up_tiles_off);
  for (uint32_t r = 0; r < nreq; ++r) {
    const uint32_t b = (uint32_t)req[r], size = (uint32_t)(req[r + 1] - req[r]);
    const uint32_t dep = req_depth ? req_depth[r] : 0u;
    if (size == 0 || dep == kSpanKeep) continue;
    if (size == 1) {
      init_retire[h0.retire++] = Retire{b, 1, b, 0, 0, 1};
    } else if (size <= T) {
      const int cls = local_class(size);
      init[cls][h0.local[cls].segs++] = LocalSeg{b, size, dep, 0};
      h0.local[cls].strings += size;
    } else {
      if (nbig >= p.cap_big) return fail(GSS_E_INTERNAL, "segment table overflow");
      const uint32_t first = tiles0;
      for (uint32_t o = 0; o < size; o += tile) {
        if (tiles0 >= p.cap_tiles) return fail(GSS_E_INTERNAL, "tile table overflow");
        tiles[tiles0] = Tile{nbig, b + o, std::min<uint32_t>(tile, size - o), tiles0 - first};
        ++tiles0;
      }
      aux0[nbig] = SegAux{0, tiles0 - first, 0, first};
      init_big[nbig++] = BigSeg{b, size, dep, 0};
      pass0_strings += size;
    }
  }
  // The first pass's counts, as its set-up kernel would leave them (the
  // host built its tile table).
  h0.big_cur = nbig;
  h0.ntiles = tiles0;
  h0.active_cur = (uint32_t)pass0_strings;
  // With global passes to run, the initial local segments are sorted on the
  // side stream meanwhile; the main stream's local phase starts after them.
  // h0 lives in the pinned counter mirror, which the read-backs overwrite.
  const Counters start = h0;
  // The start of the sort goes up in one launch: the initial local lists,
  // the retire runs, and (one range) cleared statistics, the counters and
  // the first segment list.
  UploadBatch up;
  auto add = [&](const void* src, void* dst, size_t zero_bytes, size_t bytes) -> int {
    if (add_upload(up, src, dst, zero_bytes, bytes))
      return fail(GSS_E_INTERNAL, "upload batch: bad range");
    return GSS_OK;
  };
  for (int cls = 0; cls < 3; ++cls)
    if ((rc = add(init[cls], p.local[cls], 0, sizeof(LocalSeg) * h0.local[cls].segs))) return rc;
  if ((rc = add(init_retire, p.retire, 0, sizeof(Retire) * h0.retire))) return rc;
  memcpy(w.h_init + w.init_ctr_off, &h0, sizeof(Counters));
  if ((rc = add(w.h_init, p.lstats, w.init_ctr_off, w.init_big_off + sizeof(BigSeg) * nbig)))
    return rc;
  if (nbig && (rc = add(w.h_up, p.aux, 0, w.up_tiles_off + sizeof(Tile) * tiles0))) return rc;
  GSS_HIP(launch_upload(up, st));
  // Key levels are counted only while kernels are profiled.
  const bool count_levels = g_levels.load(std::memory_order_relaxed);
  // Local segments sorted on the side stream while global passes run: the
  // small requests of a batch from the start, and (when the host reads each
  // pass's counts) those a pass left when they hold at least half as many
  // strings as the next pass.  Against a heavier pass they would hold its
  // workgroups off the CUs (one local workgroup fills a CU's LDS).  The side
  // stream has its own statistics block.
  bool side_used = false;
  long side_tok[3] = {-1, -1, -1};
  uint32_t launched[3] = {0, 0, 0};  // local segments handed to a launch so far
  uint64_t early_strings = 0;        // their strings
  // Sorts local segments [launched, upto) of each class on the side stream,
  // after the main stream's work so far (which wrote them).
  auto launch_early = [&](const uint32_t* upto) -> int {
    Plan p_side = p;
    p_side.lstats = p.lstats + 1;
    GSS_HIP(hipEventRecord(c.side_done, st));
    GSS_HIP(hipStreamWaitEvent(c.side, c.side_done, 0));
    side_used = true;
    static const char* kEarlyNames[3] = {"local128_early", "local1024_early", "local6144_early"};
    for (int cls = 2; cls >= 0; --cls) {
      if (upto[cls] <= launched[cls]) continue;
      const size_t mark = c.pending.size();
      ProfToken tok = prof_begin(c, kEarlyNames[cls], 0, c.side);
      GSS_HIP(launch_local(p_side, cls, launched[cls], upto[cls] - launched[cls], c.side,
                           count_levels));
      prof_end(c, tok, c.side);
      if (c.pending.size() > mark) side_tok[cls] = (long)mark;
      launched[cls] = upto[cls];
    }
    return GSS_OK;
  };
  if (nbig && (start.local[0].segs || start.local[1].segs || start.local[2].segs)) {
    const uint32_t upto[3] = {start.local[0].segs, start.local[1].segs, start.local[2].segs};
    if ((rc = launch_early(upto))) return rc;
    for (int cls = 0; cls < 3; ++cls) early_strings += start.local[cls].strings;
  }

  // Launch bounds of the passes.  The first pass's segments and tiles are
  // known here; a later pass has at most one segment per T + 1 of the first
  // pass's strings and at most (those strings) / tile tiles beyond one per
  // segment.  A pass's kernels are launched with those bounds and their
  // blocks past the device counts exit at once, so passes follow each other
  // with no host hand-over.  An empty block costs about 0.4 ns to dispatch
  // (five such launches per pass), a hand-over about 20 us: past
  // kBoundedTilesMax tiles (first passes of some 50M strings, e.g. 100M
  // Zipf words) the host reads each later pass's counts after its set-up
  // kernel instead and launches exact grids.
  const uint64_t ub_segs = std::min<uint64_t>(p.cap_big, pass0_strings / ((uint64_t)T + 1) + 1);
  const uint64_t ub_tiles = std::min<uint64_t>(p.cap_tiles, pass0_strings / tile + ub_segs + 1);
  constexpr uint64_t kBoundedTilesMax = 20480;
  const bool read_counts = ub_tiles > kBoundedTilesMax;
  static const char* kPassNames[10] = {"pass0", "pass1", "pass2", "pass3", "pass4",
                                       "pass5", "pass6", "pass7", "pass8", "pass9+"};
  static const char* kLocalNames[3] = {"local128", "local1024", "local6144"};
  // Each pass either splits a segment or moves it >= 1 byte deeper, and no
  // string reaches past the text, so the pass count is bounded.
  const uint64_t max_passes = text_bytes + 64;
  uint32_t pass = 0, cur = 0;
  // The local phase's profile bracket and the largest class's launch token,
  // opened early when that class goes out before the read-back.
  ProfToken phase_tok{};
  bool phase_open = false;
  uint32_t largest_main = 0;  // largest-class segments the main stream took
  long local_tok[3] = {-1, -1, -1};
  // Publishes `bytes` from the start of the statistics block (all of it and
  // the counters behind it, or from the counters on) and waits for them.
  // (Queued and awaited in two steps, so that work queued in between runs
  // while the host waits.)
  uint64_t pub_wait = 0;
  auto publish = [&](bool counters_only) -> int {
    const uint64_t seq = ++w.pub_seq;
    const size_t skip = counters_only ? w.init_ctr_off : 0;
    GSS_HIP(launch_publish(reinterpret_cast<const char*>(p.lstats) + skip,
                           w.init_ctr_off + sizeof(Counters) - skip, nullptr, nullptr, 0, 8,
                           w.h_pub + kPubHead + skip, reinterpret_cast<uint64_t*>(w.h_pub), seq,
                           st));
    pub_wait = seq;
    return GSS_OK;
  };
  auto collect = [&]() -> int {
    const int r = wait_published(w, st, pub_wait);
    if (r) return r;
    memcpy(w.h_ctr, w.h_pub + kPubHead + w.init_ctr_off, sizeof(Counters));
    if (w.h_ctr->error)
      return fail(GSS_E_INTERNAL,
                  "device sort failed (code " + std::to_string(w.h_ctr->error) + ")");
    return GSS_OK;
  };
  auto read_back = [&](bool counters_only) -> int {
    const int r = publish(counters_only);
    return r ? r : collect();
  };
  // When the host has read the counters between two passes: the local
  // segments the passes so far left are sorted on the side stream while the
  // next pass runs if they hold at least half as many strings as it does.
  // On 100M Zipf words that gains 3% over waiting for the end, while on 10M
  // URL-like strings launching 2.9M local strings against a 7.1M-string pass
  // made the sort 19% slower (one local workgroup fills a CU's LDS).
  auto maybe_launch_early = [&](uint64_t next_strings) -> int {
    uint64_t waiting = 0;
    for (int cls = 0; cls < 3; ++cls) waiting += w.h_ctr->local[cls].strings;
    waiting -= early_strings;
    if (!waiting || 2 * waiting < next_strings) return GSS_OK;
    uint32_t upto[3];
    for (int cls = 0; cls < 3; ++cls) upto[cls] = std::min(w.h_ctr->local[cls].segs, p.cap_local[cls]);
    const int e = launch_early(upto);
    if (e) return e;
    early_strings += waiting;
    return GSS_OK;
  };
  // Queues one pass; false when its set-up kernel found no segment (only
  // known, and only checked, when the host reads the counts).
  auto run_pass = [&](bool* ran) -> int {
    *ran = true;
    if (pass > max_passes) return fail(GSS_E_INTERNAL, "global passes do not terminate");
    const uint64_t bound_tiles = pass == 0 ? tiles0 : ub_tiles;
    const uint64_t bound_segs = pass == 0 ? nbig : ub_segs;
    const BigSeg* segs = p.big[cur];
    BigSeg* next = p.big[cur ^ 1];
    ProfToken pass_tok = prof_begin(c, kPassNames[pass < 9 ? pass : 9], 0, st, true);
    if (pass > 0)  // the first pass's tables came with the start of the sort
      GSS_LAUNCH("plan", 0,
                 launch_plan(p, segs, pass, (uint32_t)bound_tiles, (uint32_t)bound_segs, st));
    uint32_t grid_tiles = (uint32_t)bound_tiles, grid_segs = (uint32_t)bound_segs;
    if (pass > 0 && read_counts) {
      const int r = read_back(true);
      if (r) return r;
      grid_tiles = w.h_ctr->ntiles;
      grid_segs = w.h_ctr->big_cur;
      if (!grid_segs) {
        prof_end(c, pass_tok, st);
        *ran = false;
        return GSS_OK;
      }
      const int e = maybe_launch_early(w.h_ctr->active_cur);
      if (e) return e;
    }
    // Passes alternate: one reads its keys (and the next 8 bytes) from the
    // text and hands the following pass its keys, which that pass takes as
    // they lie.
    p.keys_carried = pass % 2;
    if (p.keys_carried) std::swap(p.keys, p.keys_next);
    // A first pass over one segment counts its digits in the gather when
    // they are formed as in the last such sort (checked on the device).
    p.fuse_hist = pass == 0 && nbig == 1;
    GSS_LAUNCH("gather", 0, launch_gather_minmax(p, segs, grid_tiles, rb, st));
    GSS_LAUNCH("segreduce", 0, launch_segreduce(p, grid_segs, st));
    GSS_LAUNCH("layout", 0, launch_layout(p, rb, st));
    GSS_LAUNCH("histogram", 0, launch_histogram(p, grid_tiles, rb, st));
    GSS_LAUNCH("flat_scan", 0, launch_flat_scan(p, (uint64_t)grid_tiles * nb, st));
    GSS_LAUNCH("segscan", 0, launch_segscan(p, segs, next, grid_segs, rb, st));
    GSS_LAUNCH("scatter", 0, launch_scatter(p, segs, grid_tiles, rb, st));
    if (lcp) GSS_LAUNCH("boundary", 0, launch_boundary_lcp(p, st));
    p.fuse_hist = 0;
    prof_end(c, pass_tok, st);
    cur ^= 1;
    ++pass;
    return GSS_OK;
  };
  if (nbig) {
    // Passes are queued back to back, as many as the last sort on this
    // context took, before the host reads the counters; while they show
    // segments left for another pass, one more is queued.  With the host
    // reading each pass's counts anyway, it runs passes until one finds no
    // segment.
    bool ran = true;
    if (read_counts) {
      while (ran)
        if ((rc = run_pass(&ran))) return rc;
      if ((rc = read_back(true))) return rc;
    } else {
      // A hint from a single sort of another size would queue empty passes
      // (or too few): then one pass at a time.  The batches of a served
      // loop's slot mix many requests, and their pass counts vary little
      // with their sizes: they always take the last batch's count.
      const bool similar =
          c.hint_strings && (nreq > 1 || (pass0_strings <= 2 * c.hint_strings &&
                                          c.hint_strings <= 2 * pass0_strings));
      uint32_t queue = similar ? std::max<uint32_t>(1, c.pass_hint) : 1u;
      // After the hinted passes the largest local class usually has all its
      // segments: it is launched then, before the read-back, with a grid of
      // the most segments it can hold (blocks past the device count exit),
      // so the host round trip overlaps its run.  Segments a further pass
      // adds are launched with the rest of the local phase.
      bool speculate = similar && c.hint_largest;
      while (queue) {
        for (uint32_t k = 0; k < queue; ++k)
          if ((rc = run_pass(&ran))) return rc;
        if ((rc = publish(true))) return rc;
        uint32_t spec_grid = 0;
        if (speculate) {
          // An empty block still holds a CU's LDS while it reads the count,
          // so the grid is the last sort's count with some slack, not the
          // class's capacity; segments beyond it wait for the local phase.
          const uint64_t ub = std::min<uint64_t>(
              p.cap_local[2], n / ((uint64_t)kLocalClassCap[1] + 1) + nreq + 1);
          const uint64_t want = std::min<uint64_t>(
              ub, (uint64_t)launched[2] + c.hint_largest + c.hint_largest / 8 + 64);
          if (want > launched[2]) {
            phase_tok = prof_begin(c, "local_phase", 0, st, true);
            phase_open = true;
            spec_grid = (uint32_t)(want - launched[2]);
            const size_t mark = c.pending.size();
            GSS_LAUNCH(kLocalNames[2], 0,
                       launch_local(p, 2, launched[2] | 0x80000000u, spec_grid, st,
                                    count_levels));
            if (c.pending.size() > mark) local_tok[2] = (long)mark;
          }
        }
        if ((rc = collect())) return rc;
        if (spec_grid) {
          // It took the segments the counters show now, up to its grid.
          const uint32_t now = std::min(w.h_ctr->local[2].segs, p.cap_local[2]);
          const uint32_t took = std::min(now, launched[2] + spec_grid);
          largest_main += took - std::min(took, launched[2]);
          launched[2] = std::max(launched[2], took);
        }
        speculate = false;
        queue = w.h_ctr->big_next ? 1u : 0u;
        if (queue && (rc = maybe_launch_early(w.h_ctr->big_next_strings))) return rc;
      }
    }
  }
  // (With no global pass the counters are still the ones this call set.)
  // The local phase, sized by the counters: every retire run and every
  // local segment the passes produced beyond those the side stream took
  // (one bracket, "local_phase"; its items are filled in after the
  // read-back).
  const uint32_t nretire = std::min(w.h_ctr->retire, p.cap_retire);
  uint32_t nlocal[3];
  for (int cls = 0; cls < 3; ++cls)
    nlocal[cls] = std::min(w.h_ctr->local[cls].segs, p.cap_local[cls]) - launched[cls];
  if (!phase_open) phase_tok = prof_begin(c, "local_phase", 0, st, true);
  if (nretire) GSS_LAUNCH("retire", nretire, launch_retire(p, std::min<uint32_t>(nretire, 8192), st));
  for (int cls = 2; cls >= 0; --cls) {
    if (!nlocal[cls]) continue;
    const size_t mark = c.pending.size();
    GSS_LAUNCH(kLocalNames[cls], 0,
               launch_local(p, cls, launched[cls], nlocal[cls], st, count_levels));
    if (c.pending.size() > mark) local_tok[cls] = (long)mark;
  }
  const size_t before_end = c.pending.size();
  prof_end(c, phase_tok, st);
  const long phase_idx = c.pending.size() > before_end ? (long)before_end : -1;
  if (side_used) {  // the early local sorts must be done before the results are
    GSS_HIP(hipEventRecord(c.side_done, c.side));
    GSS_HIP(hipStreamWaitEvent(st, c.side_done, 0));
  }
  if ((rc = read_back(false))) return rc;
  // Event brackets are read once their stream has passed them.
  if (!c.pending.empty()) GSS_HIP(hipStreamSynchronize(st));
  uint32_t passes = 0;
  for (uint32_t i = 0; i < pass && i < (uint32_t)kMaxPasses; ++i) {
    const uint64_t act = i == 0 ? pass0_strings : w.h_pass[i];
    if (i < GSS_MAX_PASSES) stats.active[i] = act;
    if (act) passes = i + 1;
  }
  stats.global_passes = passes;
  stats.first_pass_fused = w.h_ctr->hist_fused;
  if (nbig) {
    c.pass_hint = std::max<uint32_t>(1, passes);
    c.hint_strings = pass0_strings;
    c.hint_largest = largest_main + nlocal[2];
  }
  for (int cls = 0; cls < 3; ++cls) {
    stats.local_segments += w.h_ctr->local[cls].segs;
    stats.local_strings += w.h_ctr->local[cls].strings;
  }
  // A local launch's unit of work is one active string in one key round;
  // block 0 of the statistics holds the main stream's launches, block 1 the
  // side stream's.  The profiled unit of a local launch is one string through
  // one 8-byte key level (the algorithmic unit of an MSD round with 8-byte
  // keys).
  uint64_t round_strings[2][3] = {{0, 0, 0}, {0, 0, 0}}, rounds_max = 0;
  uint64_t level_strings[2][3] = {{0, 0, 0}, {0, 0, 0}};
  for (int b = 0; b < 2; ++b)
    for (int slot = 0; slot < kStatSlots; ++slot) {
      for (int cls = 0; cls < 3; ++cls) {
        round_strings[b][cls] += w.h_lstats[b].v[slot][cls];
        level_strings[b][cls] += w.h_lstats[b].v[slot][4 + cls];
      }
      rounds_max = std::max<uint64_t>(rounds_max, w.h_lstats[b].v[slot][3]);
    }
  // The local kernels count the bytes their strings advanced through: eight
  // per 8-byte key level (a first round's key starts at any byte).
  for (int b = 0; b < 2; ++b)
    for (int cls = 0; cls < 3; ++cls) level_strings[b][cls] /= 8;
  uint64_t final_level_strings = 0;
  for (int cls = 0; cls < 3; ++cls) {
    stats.local_round_strings += round_strings[0][cls] + round_strings[1][cls];
    stats.local_level_strings += level_strings[0][cls] + level_strings[1][cls];
    final_level_strings += level_strings[0][cls];
    if (side_tok[cls] >= 0) c.pending[side_tok[cls]].items = level_strings[1][cls];
  }
  for (int cls = 0; cls < 3; ++cls)
    if (local_tok[cls] >= 0) c.pending[local_tok[cls]].items = level_strings[0][cls];
  stats.local_rounds_max = (uint32_t)rounds_max;
  if (phase_idx >= 0) c.pending[phase_idx].items = final_level_strings;
  // Pass brackets: the units are the pass's active strings.
  for (auto& pe : c.pending)
    if (pe.name.compare(0, 4, "pass") == 0) {
      const int r = pe.name == "pass9+" ? 9 : pe.name[4] - '0';
      if (r < 9 && r < (int)pass) pe.items = r == 0 ? pass0_strings : w.h_pass[r];
    }
  prof_collect(c);
  return GSS_OK;
}

int run_locked(Context& c, const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
               uint32_t* out, uint64_t* lcp, uint64_t n, const uint64_t* req, uint32_t nreq,
               hipStream_t stream, const gss_tuning* tuning = nullptr,
               const uint32_t* req_depth = nullptr) {
  std::lock_guard<std::mutex> lk(c.mu);
  const int rc =
      run_sort(c, text, text_bytes, in, out, lcp, n, req, nreq, stream, tuning, req_depth);
  if (rc) {
    // A failed sort may leave launches queued on both streams; they read the
    // caller's arrays and the workspace, so they finish before the error is
    // returned (and before the workspace can be reallocated).
    (void)hipStreamSynchronize(stream ? stream : c.stream);
    (void)hipStreamSynchronize(c.side);
  }
  g_last_stats = c.stats;
  return rc;
}

}  // namespace

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int wait_stream(hipStream_t st) {
  for (;;) {
    const hipError_t e = hipStreamQuery(st);
    if (e == hipSuccess) return GSS_OK;
    if (e != hipErrorNotReady)
      return fail(GSS_E_DEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
  }
}

int library_stream(hipStream_t* out) {
  Context* c = nullptr;
  const int rc = default_context(&c);
  if (rc) return rc;
  *out = c->stream;
  return GSS_OK;
}

int sort_device32(const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
                  uint32_t* out, uint64_t* lcp, uint64_t n, hipStream_t stream) {
  if (n && text_bytes == 0) return fail(GSS_E_INVALID, "empty text: every string starts inside it");
  Context* c = nullptr;
  int rc = default_context(&c);
  if (rc) return rc;
  gss_tuning tuned;
  bool use = false;
  rc = autotune_choose(text, text_bytes, in, n, stream, &tuned, &use);
  if (rc) return rc;
  return run_locked(*c, text, text_bytes, in, out, lcp, n, nullptr, 0, stream,
                    use ? &tuned : nullptr);
}

int sort_spans_device32(const unsigned char* text, uint64_t text_bytes, uint32_t* off,
                        uint64_t* lcp, const uint64_t* req, const uint32_t* req_depth,
                        uint32_t nreq, hipStream_t stream) {
  if (!req || !req_depth) return fail(GSS_E_INVALID, "NULL request table");
  Context* c = nullptr;
  const int rc = default_context(&c);
  if (rc) return rc;
  return run_locked(*c, text, text_bytes, off, off, lcp, req[nreq], req, nreq, stream, nullptr,
                    req_depth);
}

int sort_device32_tuned(const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
                        uint32_t* out, uint64_t n, hipStream_t stream, const gss_tuning& tuning) {
  Context* c = nullptr;
  const int rc = default_context(&c);
  if (rc) return rc;
  return run_locked(*c, text, text_bytes, in, out, nullptr, n, nullptr, 0, stream, &tuning);
}

int ctx_sort_batch_tuned(gss_ctx* h, const unsigned char* text, uint64_t text_bytes,
                         const uint32_t* in, uint32_t* out, uint64_t* lcp, const uint64_t* req,
                         uint32_t nreq, hipStream_t stream, const gss_tuning& tuning) {
  if (!h) return fail(GSS_E_INVALID, "NULL context");
  if (!req) return fail(GSS_E_INVALID, "NULL request table");
  return run_locked(h->c, text, text_bytes, in, out, lcp, req[nreq], req, nreq, stream, &tuning);
}

int check_tuning(const gss_tuning& t) {
  if (t.local_threshold < 64 || t.local_threshold > kLocalCapMax)
    return fail(GSS_E_INVALID, "local_threshold must lie in [64, " + std::to_string(kLocalCapMax) + "]");
  if (t.radix_bits != 8 && t.radix_bits != 11)
    return fail(GSS_E_INVALID, "radix_bits must be 8 or 11");
  if (t.tile != 2048 && t.tile != 4096) return fail(GSS_E_INVALID, "tile must be 2048 or 4096");
  if (t.key_bytes != 32 && t.key_bytes != 64)
    return fail(GSS_E_INVALID, "key_bytes must be 32 or 64");
  return GSS_OK;
}

gss_tuning current_tuning() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_tuning;
}

size_t workspace_bytes(uint64_t n, const gss_tuning& tun) {
  Plan p{};
  uint32_t* idx_copy = nullptr;
  return items_bytes(workspace_items(p, &idx_copy, caps_for(n, tun.local_threshold, tun.tile),
                                     tun.radix_bits));
}

int ctx_reserve(gss_ctx* h, uint64_t n, uint32_t nreq, const gss_tuning& tun) {
  if (!h) return fail(GSS_E_INVALID, "NULL context");
  Context& c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  int dev = -1;
  GSS_HIP(hipGetDevice(&dev));
  if (dev != c.device) return fail(GSS_E_INVALID, "the context belongs to another device");
  int rc = ensure_workspace(c, n, tun.radix_bits, tun.local_threshold, tun.tile);
  if (rc) return rc;
  return nreq ? ensure_init_lists(c.ws, nreq) : GSS_OK;
}

}  // namespace gss

using namespace gss;

extern "C" {

int gss_sort_device32(const unsigned char* d_text, uint64_t text_bytes, const uint32_t* d_off_in,
                      uint32_t* d_off_out, uint64_t* d_lcp, uint64_t n, void* stream) {
  return sort_device32(d_text, text_bytes, d_off_in, d_off_out, d_lcp, n,
                       reinterpret_cast<hipStream_t>(stream));
}

int gss_sort_batch_device32(const unsigned char* d_text, uint64_t text_bytes,
                            const uint32_t* d_off_in, uint32_t* d_off_out, uint64_t* d_lcp,
                            const uint64_t* req_begin, uint32_t nreq, void* stream) {
  if (!req_begin) return fail(GSS_E_INVALID, "NULL request table");
  Context* c = nullptr;
  const int rc = default_context(&c);
  if (rc) return rc;
  return run_locked(*c, d_text, text_bytes, d_off_in, d_off_out, d_lcp, req_begin[nreq],
                    req_begin, nreq, reinterpret_cast<hipStream_t>(stream));
}

gss_ctx* gss_ctx_create(void) {
  gss_ctx* h = new gss_ctx;
  if (init_context(h->c)) {
    delete h;
    return nullptr;
  }
  return h;
}

void gss_ctx_destroy(gss_ctx* h) {
  if (!h) return;
  {
    std::lock_guard<std::mutex> lk(h->c.mu);
    (void)hipStreamSynchronize(h->c.stream);
    free_workspace(h->c.ws);
    if (h->c.ws.h_ctr) (void)hipHostFree(h->c.ws.h_ctr);
    if (h->c.ws.h_local) (void)hipHostFree(h->c.ws.h_local);
    for (auto e : h->c.events) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(h->c.side);
    (void)hipEventDestroy(h->c.side_done);
    (void)hipStreamDestroy(h->c.side);
    (void)hipStreamDestroy(h->c.stream);
  }
  delete h;
}

void* gss_ctx_stream(gss_ctx* h) { return h ? reinterpret_cast<void*>(h->c.stream) : nullptr; }

int gss_ctx_sort_batch_device32(gss_ctx* h, const unsigned char* d_text, uint64_t text_bytes,
                                const uint32_t* d_off_in, uint32_t* d_off_out, uint64_t* d_lcp,
                                const uint64_t* req_begin, uint32_t nreq, void* stream) {
  if (!h) return fail(GSS_E_INVALID, "NULL context");
  if (!req_begin) return fail(GSS_E_INVALID, "NULL request table");
  return run_locked(h->c, d_text, text_bytes, d_off_in, d_off_out, d_lcp, req_begin[nreq],
                    req_begin, nreq, reinterpret_cast<hipStream_t>(stream));
}

int gss_ctx_reserve(gss_ctx* h, uint64_t n) { return ctx_reserve(h, n, 0, current_tuning()); }

int gss_ctx_get_stats(gss_ctx* h, gss_stats* out) {
  if (!h || !out) return fail(GSS_E_INVALID, "NULL context or stats pointer");
  std::lock_guard<std::mutex> lk(h->c.mu);
  *out = h->c.stats;
  return GSS_OK;
}

int gss_get_tuning(gss_tuning* out) {
  if (!out) return fail(GSS_E_INVALID, "NULL tuning pointer");
  std::lock_guard<std::mutex> lk(g_mu);
  *out = g_tuning;
  return GSS_OK;
}

int gss_set_tuning(const gss_tuning* in) {
  if (!in) return fail(GSS_E_INVALID, "NULL tuning pointer");
  const int rc = check_tuning(*in);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuning = *in;
  return GSS_OK;
}

int gss_get_stats(gss_stats* out) {
  if (!out) return fail(GSS_E_INVALID, "NULL stats pointer");
  *out = g_last_stats;
  return GSS_OK;
}

int gss_profile_enable(int mode) {
  if (mode & ~(GSS_PROFILE_EVENTS | GSS_PROFILE_LEVELS | GSS_PROFILE_PASSES))
    return fail(GSS_E_INVALID, "unknown profile mode bits");
  g_prof.store((mode & GSS_PROFILE_EVENTS) != 0);
  g_prof_pass.store((mode & GSS_PROFILE_PASSES) != 0);
  g_levels.store((mode & GSS_PROFILE_LEVELS) != 0);
  return GSS_OK;
}

int gss_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_prof_totals.clear();
  return GSS_OK;
}

int gss_profile_get(const char* name, uint64_t* launches, double* total_ms, uint64_t* items) {
  if (!name) return fail(GSS_E_INVALID, "NULL kernel name");
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_prof_totals.find(name);
  ProfTotal t;
  if (it != g_prof_totals.end()) t = it->second;
  if (launches) *launches = t.launches;
  if (total_ms) *total_ms = t.ms;
  if (items) *items = t.items;
  return GSS_OK;
}

const char* gss_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
