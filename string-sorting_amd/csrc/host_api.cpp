// Pointer front end: the routine.h drop-in (gss_msd, gss_msd_lcp).
//
// The caller's strings are NUL-terminated host strings.  The text moves into
// HBM once: when every string lies inside one readable address range of less
// than 4 GiB (sortstring's text buffer, the malloc heap of strdup'd strings,
// suffix pointers), that whole span is copied and each pointer becomes its
// 32-bit distance from the span start; otherwise the strings are packed one
// after another.  After the device sort the offsets turn back into the
// caller's own pointer values, so the output is a permutation of the input
// pointers, as the reference's routines produce (src/sortstring.c:434-436,
// unit-test/main.cpp:226-229).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cinttypes>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gss.h"
#include "gss_internal.h"
#include "host.h"

namespace gss {
namespace {

[[noreturn]] void die(const std::string& what) {
  fprintf(stderr, "gss_msd: %s\n", what.c_str());
  fflush(stderr);
  abort();
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

// True when [lo, hi) is covered by consecutive readable mappings of this
// process (the same /proc/self/maps source the reference's vma_info reads,
// src/util/vmainfo.c:83-129).
bool span_readable(uintptr_t lo, uintptr_t hi) {
  FILE* f = fopen("/proc/self/maps", "r");
  if (!f) return false;
  char line[512];
  uintptr_t covered = lo;
  bool ok = false;
  while (fgets(line, sizeof(line), f)) {
    unsigned long a = 0, b = 0;
    char perms[8] = {0};
    if (sscanf(line, "%lx-%lx %7s", &a, &b, perms) != 3) continue;
    if (b <= covered) continue;
    if (a > covered) break;  // a hole before `covered`
    if (perms[0] != 'r') break;
    covered = b;
    if (covered >= hi) {
      ok = true;
      break;
    }
  }
  fclose(f);
  return ok;
}

// A fork-join pool of host threads for the pointer front end's bulk host
// work (pointer scans and copies between the caller's pageable memory and
// pinned staging buffers).  Workers never call HIP.  run(fn) calls
// fn(worker, workers) on every worker, the calling thread being worker 0,
// and returns when all have finished.
class HostPool {
 public:
  explicit HostPool(int n) : n_(n < 1 ? 1 : n) {
    for (int w = 1; w < n_; ++w) th_.emplace_back([this, w] { loop(w); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_go_.notify_all();
    for (auto& t : th_) t.join();
  }
  int size() const { return n_; }
  void run(const std::function<void(int, int)>& fn) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_go_.notify_all();
    fn(0, n_);
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  void loop(int w) {
    uint64_t seen = 0;
    while (true) {
      const std::function<void(int, int)>* job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_go_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
      }
      (*job)(w, n_);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) cv_done_.notify_one();
    }
  }
  const int n_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_go_, cv_done_;
  const std::function<void(int, int)>* job_ = nullptr;
  uint64_t gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
};

HostPool& host_pool() {
  static HostPool pool(
      (int)std::min<unsigned>(8u, std::max<unsigned>(1u, std::thread::hardware_concurrency() / 2)));
  return pool;
}

// memcpy split over the pool (4 KiB-aligned pieces).
void par_memcpy(void* dst, const void* src, uint64_t bytes) {
  HostPool& pool = host_pool();
  pool.run([&](int w, int nw) {
    const uint64_t piece = ((bytes + nw - 1) / nw + 4095) & ~4095ull;
    const uint64_t a = std::min<uint64_t>(bytes, piece * w), b = std::min<uint64_t>(bytes, a + piece);
    if (b > a) memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
  });
}

// The 32-bit distances of n pointers from base, split over the pool.
void par_offsets(uint32_t* dst, unsigned char* const* ptrs, uint64_t n, uintptr_t base) {
  host_pool().run([&](int w, int nw) {
    const uint64_t a = n * (uint64_t)w / nw, b = n * (uint64_t)(w + 1) / nw;
    for (uint64_t i = a; i < b; ++i)
      dst[i] = (uint32_t)(reinterpret_cast<uintptr_t>(ptrs[i]) - base);
  });
}

// base + off[i] for n 32-bit offsets, split over the pool.
void par_pointers(unsigned char** dst, const uint32_t* off, uint64_t n, uintptr_t base) {
  host_pool().run([&](int w, int nw) {
    const uint64_t a = n * (uint64_t)w / nw, b = n * (uint64_t)(w + 1) / nw;
    for (uint64_t i = a; i < b; ++i) dst[i] = reinterpret_cast<unsigned char*>(base + off[i]);
  });
}

// Bytes [from, from + len) of the packed stream of m strings (string i at
// starts[i], lens[i] bytes with its NUL) into dst, split over the pool.
void par_pack(unsigned char* dst, unsigned char* const* ptrs, const uint32_t* starts,
              const uint64_t* lens, uint64_t m, uint64_t from, uint64_t len) {
  host_pool().run([&](int w, int nw) {
    const uint64_t a = from + len * (uint64_t)w / nw, b = from + len * (uint64_t)(w + 1) / nw;
    if (a >= b) return;
    // the string holding byte a: the last start <= a
    uint64_t i = (uint64_t)(std::upper_bound(starts, starts + m, (uint32_t)a) - starts) - 1;
    for (uint64_t pos = a; pos < b; ++i) {
      const uint64_t s0 = starts[i], e = std::min<uint64_t>(s0 + lens[i], b);
      memcpy(dst + (pos - from), ptrs[i] + (pos - s0), e - pos);
      pos = e;
    }
  });
}

// dst[i] = orig[idx[i]] for n indexes, split over the pool.
void par_gather(unsigned char** dst, const uint32_t* idx, uint64_t n, unsigned char* const* orig) {
  host_pool().run([&](int w, int nw) {
    const uint64_t a = n * (uint64_t)w / nw, b = n * (uint64_t)(w + 1) / nw;
    for (uint64_t i = a; i < b; ++i) dst[i] = orig[idx[i]];
  });
}

// Copies between the caller's pageable memory and the device through a ring
// of pinned chunks: host threads fill (or drain) one chunk while the DMA of
// the others runs.  Pageable copies go through the runtime's own staging at
// a fraction of the link's rate; this keeps the link busy.  A slot is reused
// only after the event of its last copy completed (also across calls).
class Staging {
 public:
  static constexpr uint64_t kChunk = 16ull << 20;
  static constexpr int kSlots = 4;

  void h2d(void* dst, const void* src, uint64_t bytes, hipStream_t st) {
    if (bytes < 2 * kChunk) {
      check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "hipMemcpy(H2D)");
      return;
    }
    ensure();
    int slot = 0;
    for (uint64_t off = 0; off < bytes; off += kChunk, slot = (slot + 1) % kSlots) {
      const uint64_t len = std::min(kChunk, bytes - off);
      settle(slot);
      par_memcpy(buf_ + slot * kChunk, static_cast<const char*>(src) + off, len);
      check(hipMemcpyAsync(static_cast<char*>(dst) + off, buf_ + slot * kChunk, len,
                           hipMemcpyHostToDevice, st),
            "hipMemcpy(H2D)");
      arm(slot, st, Drain::kNone, nullptr, 0);
    }
  }

  // Uploads m strings packed one after another (string i at starts[i],
  // lens[i] bytes with its NUL, `total` bytes in all): host threads copy
  // each chunk of the packed stream straight from the strings into a pinned
  // chunk.
  void h2d_packed(unsigned char* dst, unsigned char* const* ptrs, const uint32_t* starts,
                  const uint64_t* lens, uint64_t m, uint64_t total, hipStream_t st) {
    ensure();
    int slot = 0;
    for (uint64_t off = 0; off < total; off += kChunk, slot = (slot + 1) % kSlots) {
      const uint64_t len = std::min(kChunk, total - off);
      settle(slot);
      par_pack(buf_ + slot * kChunk, ptrs, starts, lens, m, off, len);
      check(hipMemcpyAsync(dst + off, buf_ + slot * kChunk, len, hipMemcpyHostToDevice, st),
            "hipMemcpy(H2D)");
      arm(slot, st, Drain::kNone, nullptr, 0);
    }
  }

  // Uploads the 32-bit distances from `base` of n string pointers: the host
  // threads compute them into the pinned chunks (half the bytes of the
  // pointers cross the link).
  void h2d_offsets(uint32_t* dst, unsigned char* const* ptrs, uint64_t n, uintptr_t base,
                   hipStream_t st) {
    ensure();
    constexpr uint64_t kPer = kChunk / 4;
    int slot = 0;
    for (uint64_t i = 0; i < n; i += kPer, slot = (slot + 1) % kSlots) {
      const uint64_t m = std::min(kPer, n - i);
      settle(slot);
      par_offsets(reinterpret_cast<uint32_t*>(buf_ + slot * kChunk), ptrs + i, m, base);
      check(hipMemcpyAsync(dst + i, buf_ + slot * kChunk, 4 * m, hipMemcpyHostToDevice, st),
            "hipMemcpy(H2D)");
      arm(slot, st, Drain::kNone, nullptr, 0);
    }
  }

  // Downloads n 32-bit values and turns each into a pointer in ptrs: base +
  // value (offsets into a copied span), or orig[value] (indexes of packed
  // strings) when orig is given.  Returns once all of them are there.
  void d2h_pointers(unsigned char** ptrs, const uint32_t* src, uint64_t n, uintptr_t base,
                    unsigned char* const* orig, hipStream_t st) {
    ensure();
    constexpr uint64_t kPer = kChunk / 4;
    int slot = 0;
    for (uint64_t i = 0; i < n; i += kPer, slot = (slot + 1) % kSlots) {
      const uint64_t m = std::min(kPer, n - i);
      settle(slot);
      check(hipMemcpyAsync(buf_ + slot * kChunk, src + i, 4 * m, hipMemcpyDeviceToHost, st),
            "hipMemcpy(D2H)");
      arm(slot, st, orig ? Drain::kGather : Drain::kWiden, ptrs + i, m, base, orig);
    }
    for (int q = 0; q < kSlots; ++q) settle((slot + q) % kSlots);  // drain in order
  }

  // Returns once the data is in `dst` (the stream's earlier work included).
  void d2h(void* dst, const void* src, uint64_t bytes, hipStream_t st) {
    if (bytes < 2 * kChunk) {
      check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "hipMemcpy(D2H)");
      check(hipStreamSynchronize(st), "hipStreamSynchronize");
      return;
    }
    ensure();
    int slot = 0;
    for (uint64_t off = 0; off < bytes; off += kChunk, slot = (slot + 1) % kSlots) {
      const uint64_t len = std::min(kChunk, bytes - off);
      settle(slot);
      check(hipMemcpyAsync(buf_ + slot * kChunk, static_cast<const char*>(src) + off, len,
                           hipMemcpyDeviceToHost, st),
            "hipMemcpy(D2H)");
      arm(slot, st, Drain::kBytes, static_cast<char*>(dst) + off, len);
    }
    for (int q = 0; q < kSlots; ++q) settle((slot + q) % kSlots);  // drain in order
  }

 private:
  // What happens to a chunk once its copy is done.
  enum class Drain { kNone, kBytes, kWiden, kGather };
  void ensure() {
    if (buf_) return;
    check(hipHostMalloc((void**)&buf_, kChunk * kSlots, hipHostMallocDefault), "hipHostMalloc");
    for (int q = 0; q < kSlots; ++q)
      check(hipEventCreateWithFlags(&ev_[q], hipEventDisableTiming), "hipEventCreate");
  }
  // Waits for the slot's last copy; a pending device-to-host chunk is then
  // moved on to its destination.
  void settle(int slot) {
    if (!armed_[slot]) return;
    check(hipEventSynchronize(ev_[slot]), "hipEventSynchronize");
    armed_[slot] = false;
    const uint32_t* vals = reinterpret_cast<const uint32_t*>(buf_ + slot * kChunk);
    switch (drain_[slot]) {
      case Drain::kNone:
        break;
      case Drain::kBytes:
        par_memcpy(out_[slot], buf_ + slot * kChunk, out_len_[slot]);
        break;
      case Drain::kWiden:
        par_pointers(static_cast<unsigned char**>(out_[slot]), vals, out_len_[slot], base_[slot]);
        break;
      case Drain::kGather:
        par_gather(static_cast<unsigned char**>(out_[slot]), vals, out_len_[slot], orig_[slot]);
        break;
    }
    drain_[slot] = Drain::kNone;
  }
  // out/len: where the chunk goes once its copy is done.
  void arm(int slot, hipStream_t st, Drain how, void* out, uint64_t len, uintptr_t base = 0,
           unsigned char* const* orig = nullptr) {
    check(hipEventRecord(ev_[slot], st), "hipEventRecord");
    armed_[slot] = true;
    drain_[slot] = how;
    out_[slot] = out;
    out_len_[slot] = len;
    base_[slot] = base;
    orig_[slot] = orig;
  }
  unsigned char* buf_ = nullptr;
  hipEvent_t ev_[kSlots] = {};
  bool armed_[kSlots] = {};
  Drain drain_[kSlots] = {};
  void* out_[kSlots] = {};
  uint64_t out_len_[kSlots] = {};
  uintptr_t base_[kSlots] = {};
  unsigned char* const* orig_[kSlots] = {};
};

Staging g_staging;

struct PtrBuffers {
  unsigned char* text = nullptr;
  uint64_t text_cap = 0;
  uint32_t* off = nullptr;
  uint32_t* out = nullptr;
  uint32_t* idx = nullptr;
  uint64_t* lcp = nullptr;
  uint64_t n_cap = 0;
};

std::mutex g_ptr_mu;  // one pointer sort at a time: they share g_buf and g_staging
PtrBuffers g_buf;

// Largest single device sort of the pointer path (gss_set_pointer_limits).
constexpr uint64_t kMaxPieceText = (1ull << 32) - 64;
constexpr uint64_t kMaxPieceStrings = (1ull << 31) - 1;
uint64_t g_max_text = kMaxPieceText;
uint64_t g_max_strings = kMaxPieceStrings;
gss_pointer_info g_info = {};

void ensure_text(uint64_t bytes) {
  if (bytes <= g_buf.text_cap) return;
  if (g_buf.text) check(hipFree(g_buf.text), "hipFree");
  const uint64_t cap = bytes + bytes / 8 + 64;
  check(hipMalloc((void**)&g_buf.text, cap), "hipMalloc(text)");
  g_buf.text_cap = cap;
}

void ensure_n(uint64_t n) {
  if (n <= g_buf.n_cap) return;
  if (g_buf.off) {
    check(hipFree(g_buf.off), "hipFree");
    check(hipFree(g_buf.out), "hipFree");
    check(hipFree(g_buf.idx), "hipFree");
    check(hipFree(g_buf.lcp), "hipFree");
  }
  const uint64_t cap = n + n / 8 + 16;
  check(hipMalloc((void**)&g_buf.off, 4 * cap), "hipMalloc(offsets)");
  check(hipMalloc((void**)&g_buf.out, 4 * cap), "hipMalloc(offsets)");
  check(hipMalloc((void**)&g_buf.idx, 4 * cap), "hipMalloc(indexes)");
  check(hipMalloc((void**)&g_buf.lcp, 8 * cap), "hipMalloc(lcp)");
  g_buf.n_cap = cap;
}

void device_sort(uint64_t text_len, uint64_t n, bool lcp, hipStream_t st) {
  check(hipMemsetAsync(g_buf.text + text_len, 0, 64, st), "hipMemset(pad)");
  const int rc = sort_device32(g_buf.text, text_len, g_buf.off, g_buf.out,
                               lcp ? g_buf.lcp : nullptr, n, st);
  if (rc) die(std::string("device sort failed: ") + gss_last_error());
  ++g_info.pieces;
  g_info.uploaded_text_bytes += text_len;
}

// Equal bytes of two NUL-terminated strings from byte `from` on, before the
// first mismatch or NUL (lcp() of src/mergesort_lcp.cpp:88-100 from a
// known common prefix).
uint64_t lcp_from(const unsigned char* a, const unsigned char* b, uint64_t from) {
  uint64_t i = from;
  while (a[i] && a[i] == b[i]) ++i;
  return i;
}

// Sorts the strings s[0..m) (lens[i] bytes each with the NUL, `total` bytes
// in all) as one packed device text: offsets are assigned in input order,
// so the sorted offsets map back to the caller's pointers through their
// input indexes.  lcp: m-1 entries, or NULL.
void sort_packed_piece(unsigned char** s, const uint64_t* lens, uint64_t m, uint64_t total,
                       size_t* lcp, hipStream_t st) {
  std::vector<uint32_t> starts(m);
  uint64_t pos = 0;
  for (uint64_t i = 0; i < m; ++i) {
    starts[i] = (uint32_t)pos;
    pos += lens[i];
  }
  ensure_n(m);
  ensure_text(total + 64);
  g_staging.h2d_packed(g_buf.text, s, starts.data(), lens, m, total, st);
  g_staging.h2d(g_buf.off, starts.data(), 4 * m, st);
  device_sort(total, m, lcp != nullptr, st);
  check(launch_off_to_index(g_buf.out, g_buf.off, g_buf.idx, m, st), "off_to_index");
  std::vector<unsigned char*> orig(s, s + m);
  g_staging.d2h_pointers(s, g_buf.idx, m, 0, orig.data(), st);
  if (lcp) g_staging.d2h(lcp, g_buf.lcp, 8 * (m - 1), st);
}

// The packed path.  A set whose packed bytes or count exceed one device
// sort's limits is first split on the host by the byte at the common
// depth, MSD fashion (src/msd_ce.cpp:151-165), until every piece fits:
// consecutive buckets share a piece while they fit, a bucket too large
// alone splits again one byte deeper (after skipping the prefix all its
// strings share), and the strings that end at the depth (bucket 0) are
// equal and final.  The distribution is stable, so the output is the one
// device sort's.  LCPs across piece borders are computed here.  `span`:
// bytes from the lowest string to the end of the highest; strings that do
// not overlap hold at most that many bytes together.
void sort_packed(unsigned char** s, size_t n, uint64_t span, size_t* lcp, hipStream_t st) {
  std::vector<uint64_t> lens(n);
  uint64_t total = 0;
  {
    HostPool& pool = host_pool();
    std::vector<uint64_t> wsum(pool.size(), 0);
    pool.run([&](int w, int nw) {
      const size_t a = n * (size_t)w / nw, b = n * (size_t)(w + 1) / nw;
      uint64_t t = 0;
      for (size_t i = a; i < b && t <= span; ++i)
        t += lens[i] = strlen(reinterpret_cast<char*>(s[i])) + 1;
      wsum[w] = t;
    });
    for (uint64_t t : wsum) total += t;
  }
  // Overlapping strings (suffixes of one text) pack to far more bytes than
  // their span: only the span path can take them, and their span is over
  // the limit of one device sort.
  if (total > span)
    die("the strings overlap (a suffix set?) over a span of " + std::to_string(span) +
        " bytes: overlapping strings must lie within " + std::to_string(g_max_text) +
        " bytes (the device text limit)");
  g_info.packed = 1;
  if (total <= g_max_text && n <= g_max_strings) {
    sort_packed_piece(s, lens.data(), n, total, lcp, st);
    return;
  }

  struct Range {
    uint64_t a, b, bytes, depth;
  };
  std::vector<Range> todo{{0, n, total, 0}};
  std::vector<uint64_t> borders;  // first index of every final piece
  std::vector<unsigned char*> tmp_s;
  std::vector<uint64_t> tmp_l;
  HostPool& pool = host_pool();
  const int nw = pool.size();
  std::vector<uint64_t> cnt((size_t)nw * 256), byt((size_t)nw * 256), put((size_t)nw * 256);
  while (!todo.empty()) {
    const Range r = todo.back();
    todo.pop_back();
    const uint64_t m = r.b - r.a;
    borders.push_back(r.a);
    if (m <= 1) continue;
    if (r.bytes <= g_max_text && m <= g_max_strings) {
      sort_packed_piece(s + r.a, lens.data() + r.a, m, r.bytes, lcp ? lcp + r.a : nullptr, st);
      continue;
    }
    // Stable distribution of s[a..b) by the byte at r.depth (every string
    // of the range has its first r.depth bytes equal and no NUL before).
    if (tmp_s.empty()) {
      tmp_s.resize(n);
      tmp_l.resize(n);
    }
    std::fill(cnt.begin(), cnt.end(), 0);
    std::fill(byt.begin(), byt.end(), 0);
    pool.run([&](int w, int nwk) {
      const uint64_t a = r.a + m * w / nwk, b = r.a + m * (w + 1) / nwk;
      for (uint64_t i = a; i < b; ++i) {
        const unsigned c = s[i][r.depth];
        ++cnt[(size_t)w * 256 + c];
        byt[(size_t)w * 256 + c] += lens[i];
      }
    });
    uint64_t bc[256], bb[256], run = r.a;
    for (int c = 0; c < 256; ++c) {
      bc[c] = bb[c] = 0;
      for (int w = 0; w < nw; ++w) {
        put[(size_t)w * 256 + c] = run;
        run += cnt[(size_t)w * 256 + c];
        bc[c] += cnt[(size_t)w * 256 + c];
        bb[c] += byt[(size_t)w * 256 + c];
      }
    }
    int only = -1;  // the one nonzero bucket holding every string, if any
    for (int c = 1; c < 256; ++c)
      if (bc[c] == m) only = c;
    if (only >= 0) {
      // No split at this depth: skip the bytes every string shares.
      std::vector<uint64_t> wmin(nw, UINT64_MAX);
      const unsigned char* first = s[r.a];
      pool.run([&](int w, int nwk) {
        const uint64_t a = r.a + m * w / nwk, b = r.a + m * (w + 1) / nwk;
        uint64_t lo = UINT64_MAX;
        for (uint64_t i = a; i < b; ++i) lo = std::min(lo, lcp_from(first, s[i], r.depth + 1));
        wmin[w] = lo;
      });
      uint64_t d = UINT64_MAX;
      for (uint64_t v : wmin) d = std::min(d, v);
      borders.pop_back();
      todo.push_back(Range{r.a, r.b, r.bytes, d});
      continue;
    }
    pool.run([&](int w, int nwk) {
      const uint64_t a = r.a + m * w / nwk, b = r.a + m * (w + 1) / nwk;
      uint64_t* p = &put[(size_t)w * 256];
      for (uint64_t i = a; i < b; ++i) {
        const uint64_t j = p[s[i][r.depth]]++;
        tmp_s[j] = s[i];
        tmp_l[j] = lens[i];
      }
    });
    pool.run([&](int w, int nwk) {
      const uint64_t a = r.a + m * w / nwk, b = r.a + m * (w + 1) / nwk;
      std::copy(tmp_s.begin() + a, tmp_s.begin() + b, s + a);
      std::copy(tmp_l.begin() + a, tmp_l.begin() + b, lens.begin() + a);
    });
    // Children in order: the strings ending here (equal, final), then runs
    // of buckets that fit one device sort, or a single bucket to split again.
    std::vector<Range> kids;
    uint64_t at = r.a;
    if (bc[0]) {
      borders.push_back(at);
      if (lcp)
        for (uint64_t i = at; i + 1 < at + bc[0]; ++i) lcp[i] = r.depth;
      at += bc[0];
    }
    Range group{at, at, 0, r.depth + 1};
    for (int c = 1; c < 256; ++c) {
      if (!bc[c]) continue;
      const bool alone = bb[c] > g_max_text || bc[c] > g_max_strings;
      if (alone || group.bytes + bb[c] > g_max_text || group.b - group.a + bc[c] > g_max_strings) {
        if (group.b > group.a) kids.push_back(group);
        group = Range{at, at, 0, r.depth + 1};
      }
      if (alone) {
        kids.push_back(Range{at, at + bc[c], bb[c], r.depth + 1});
        group = Range{at + bc[c], at + bc[c], 0, r.depth + 1};
      } else {
        group.b += bc[c];
        group.bytes += bb[c];
      }
      at += bc[c];
    }
    if (group.b > group.a) kids.push_back(group);
    for (size_t k = kids.size(); k-- > 0;) todo.push_back(kids[k]);
  }
  if (lcp)
    for (uint64_t x : borders)
      if (x > 0 && x < n) lcp[x - 1] = lcp_from(s[x - 1], s[x], 0);
}

// True when the strings' span is much larger than their bytes (a few strings
// spread over a large heap): estimated from the lengths of up to 1024 evenly
// spaced strings, each counted up to 256 bytes.
bool span_is_sparse(unsigned char* const* s, size_t n, uint64_t span) {
  constexpr uint64_t kSample = 1024, kLenCap = 256, kFloor = 16ull << 20;
  if (span <= kFloor) return false;
  const uint64_t k = std::min<uint64_t>(n, kSample);
  uint64_t sum = 0;
  for (uint64_t j = 0; j < k; ++j)
    sum += strnlen(reinterpret_cast<const char*>(s[j * n / k]), kLenCap) + 1;
  const double est = (double)sum / (double)k * (double)n;
  return (double)span > 2.0 * est + (double)kFloor;
}

void sort_pointers(unsigned char** strings, size_t n, size_t* lcp_out) {
  if (n <= 1) return;
  if (!strings) die("NULL string array");
  std::lock_guard<std::mutex> lk(g_ptr_mu);
  g_info = gss_pointer_info{};
  hipStream_t st = nullptr;
  if (library_stream(&st)) die(gss_last_error());

  // Lowest and highest pointer (a NULL shows up as lo == 0), split over the
  // host pool.
  uintptr_t lo = UINTPTR_MAX, hi = 0;
  {
    HostPool& pool = host_pool();
    std::vector<uintptr_t> wlo(pool.size(), UINTPTR_MAX), whi(pool.size(), 0);
    pool.run([&](int w, int nw) {
      const size_t a = n * (size_t)w / nw, b = n * (size_t)(w + 1) / nw;
      uintptr_t l = UINTPTR_MAX, h = 0;
      for (size_t i = a; i < b; ++i) {
        const uintptr_t p = reinterpret_cast<uintptr_t>(strings[i]);
        l = p < l ? p : l;
        h = p > h ? p : h;
      }
      wlo[w] = l;
      whi[w] = h;
    });
    for (int w = 0; w < pool.size(); ++w) {
      lo = std::min(lo, wlo[w]);
      hi = std::max(hi, whi[w]);
    }
  }
  if (lo == 0) die("NULL string pointer");
  const uintptr_t end = hi + strlen(reinterpret_cast<const char*>(hi)) + 1;
  const uint64_t span = end - lo;
  const bool use_span = span <= g_max_text && n <= g_max_strings &&
                        !span_is_sparse(strings, n, span) && span_readable(lo, end);
  if (!use_span) {
    sort_packed(strings, n, span, lcp_out, st);
    return;
  }
  // The span path: the whole address range moves, each pointer becomes its
  // distance from the range start, and the sorted distances come back as
  // pointers.
  ensure_n(n);
  ensure_text(span + 64);
  g_staging.h2d(g_buf.text, reinterpret_cast<const void*>(lo), span, st);
  g_staging.h2d_offsets(g_buf.off, strings, n, lo, st);
  device_sort(span, n, lcp_out != nullptr, st);
  g_staging.d2h_pointers(strings, g_buf.out, n, lo, nullptr, st);
  if (lcp_out) g_staging.d2h(lcp_out, g_buf.lcp, 8 * (n - 1), st);
}

}  // namespace
}  // namespace gss

extern "C" {

void gss_msd(unsigned char** strings, size_t n) { gss::sort_pointers(strings, n, nullptr); }

void gss_msd_lcp(unsigned char** strings, size_t n, size_t* lcp) {
  gss::sort_pointers(strings, n, lcp);
}

int gss_set_pointer_limits(uint64_t max_text_bytes, uint64_t max_strings) {
  const uint64_t t = max_text_bytes ? max_text_bytes : gss::kMaxPieceText;
  const uint64_t m = max_strings ? max_strings : gss::kMaxPieceStrings;
  if (t < 64 || t > gss::kMaxPieceText)
    return gss::fail(GSS_E_INVALID, "max_text_bytes must lie in [64, 2^32-64] (0: the default)");
  if (m < 2 || m > gss::kMaxPieceStrings)
    return gss::fail(GSS_E_INVALID, "max_strings must lie in [2, 2^31-1] (0: the default)");
  std::lock_guard<std::mutex> lk(gss::g_ptr_mu);
  gss::g_max_text = t;
  gss::g_max_strings = m;
  return GSS_OK;
}

int gss_pointer_last_info(gss_pointer_info* out) {
  if (!out) return gss::fail(GSS_E_INVALID, "NULL info");
  std::lock_guard<std::mutex> lk(gss::g_ptr_mu);
  *out = gss::g_info;
  return GSS_OK;
}

}  // extern "C"
