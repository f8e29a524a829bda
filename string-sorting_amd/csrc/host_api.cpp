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
      arm(slot, st, nullptr, 0);
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
      arm(slot, st, nullptr, 0);
    }
  }

  // Downloads n 32-bit offsets and stores base + offset into ptrs; returns
  // once all of them are there.
  void d2h_pointers(unsigned char** ptrs, const uint32_t* src, uint64_t n, uintptr_t base,
                    hipStream_t st) {
    ensure();
    constexpr uint64_t kPer = kChunk / 4;
    int slot = 0;
    for (uint64_t i = 0; i < n; i += kPer, slot = (slot + 1) % kSlots) {
      const uint64_t m = std::min(kPer, n - i);
      settle(slot);
      check(hipMemcpyAsync(buf_ + slot * kChunk, src + i, 4 * m, hipMemcpyDeviceToHost, st),
            "hipMemcpy(D2H)");
      arm(slot, st, ptrs + i, m, base, true);
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
      arm(slot, st, static_cast<char*>(dst) + off, len);
    }
    for (int q = 0; q < kSlots; ++q) settle((slot + q) % kSlots);  // drain in order
  }

 private:
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
    if (out_[slot] && widen_[slot])
      par_pointers(static_cast<unsigned char**>(out_[slot]),
                   reinterpret_cast<const uint32_t*>(buf_ + slot * kChunk), out_len_[slot],
                   base_[slot]);
    else if (out_[slot])
      par_memcpy(out_[slot], buf_ + slot * kChunk, out_len_[slot]);
    out_[slot] = nullptr;
  }
  // out/len: where the chunk goes once its copy is done (bytes, or offsets
  // to widen into pointers from base when widen is set).
  void arm(int slot, hipStream_t st, void* out, uint64_t len, uintptr_t base = 0,
           bool widen = false) {
    check(hipEventRecord(ev_[slot], st), "hipEventRecord");
    armed_[slot] = true;
    out_[slot] = out;
    out_len_[slot] = len;
    base_[slot] = base;
    widen_[slot] = widen;
  }
  unsigned char* buf_ = nullptr;
  hipEvent_t ev_[kSlots] = {};
  bool armed_[kSlots] = {};
  void* out_[kSlots] = {};
  uint64_t out_len_[kSlots] = {};
  uintptr_t base_[kSlots] = {};
  bool widen_[kSlots] = {};
};

Staging g_staging;

struct PtrBuffers {
  unsigned char* text = nullptr;
  uint64_t text_cap = 0;
  uint64_t* ptrs = nullptr;
  uint64_t* ptrs2 = nullptr;
  uint32_t* off = nullptr;
  uint32_t* out = nullptr;
  uint64_t* lcp = nullptr;
  uint64_t n_cap = 0;
  unsigned char* pinned = nullptr;  // packing staging
  uint64_t pinned_cap = 0;
};

std::mutex g_ptr_mu;
PtrBuffers g_buf;

void ensure_text(uint64_t bytes) {
  if (bytes <= g_buf.text_cap) return;
  if (g_buf.text) check(hipFree(g_buf.text), "hipFree");
  const uint64_t cap = bytes + bytes / 8 + 64;
  check(hipMalloc((void**)&g_buf.text, cap), "hipMalloc(text)");
  g_buf.text_cap = cap;
}

void ensure_n(uint64_t n) {
  if (n <= g_buf.n_cap) return;
  if (g_buf.ptrs) {
    check(hipFree(g_buf.ptrs), "hipFree");
    check(hipFree(g_buf.ptrs2), "hipFree");
    check(hipFree(g_buf.off), "hipFree");
    check(hipFree(g_buf.out), "hipFree");
    check(hipFree(g_buf.lcp), "hipFree");
  }
  const uint64_t cap = n + n / 8 + 16;
  check(hipMalloc((void**)&g_buf.ptrs, 8 * cap), "hipMalloc(ptrs)");
  check(hipMalloc((void**)&g_buf.ptrs2, 8 * cap), "hipMalloc(ptrs)");
  check(hipMalloc((void**)&g_buf.off, 4 * cap), "hipMalloc(offsets)");
  check(hipMalloc((void**)&g_buf.out, 4 * cap), "hipMalloc(offsets)");
  check(hipMalloc((void**)&g_buf.lcp, 8 * cap), "hipMalloc(lcp)");
  g_buf.n_cap = cap;
}

void sort_pointers(unsigned char** strings, size_t n, size_t* lcp_out) {
  if (n <= 1) return;
  if (!strings) die("NULL string array");
  if (n >= (1ull << 31)) die("more than 2^31-1 strings are not supported");
  std::lock_guard<std::mutex> lk(g_ptr_mu);
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
  const bool use_span = span < (1ull << 32) - 64 && span_readable(lo, end);

  ensure_n(n);
  uint64_t text_len = 0;
  if (use_span) {
    text_len = span;
    ensure_text(text_len + 64);
    g_staging.h2d(g_buf.text, reinterpret_cast<const void*>(lo), span, st);
    g_staging.h2d_offsets(g_buf.off, strings, n, lo, st);
  } else {
    std::vector<uint32_t> starts(n);
    uint64_t total = 0;
    for (size_t i = 0; i < n; ++i) {
      starts[i] = (uint32_t)total;
      total += strlen(reinterpret_cast<const char*>(strings[i])) + 1;
      if (total >= (1ull << 32) - 64) die("packed text exceeds 4 GiB");
    }
    text_len = total;
    if (g_buf.pinned_cap < total) {
      if (g_buf.pinned) check(hipHostFree(g_buf.pinned), "hipHostFree");
      check(hipHostMalloc((void**)&g_buf.pinned, total, hipHostMallocDefault), "hipHostMalloc");
      g_buf.pinned_cap = total;
    }
    for (size_t i = 0; i < n; ++i) {
      const size_t l = (i + 1 < n ? starts[i + 1] : total) - starts[i];
      memcpy(g_buf.pinned + starts[i], strings[i], l);
    }
    ensure_text(text_len + 64);
    check(hipMemcpyAsync(g_buf.text, g_buf.pinned, total, hipMemcpyHostToDevice, st),
          "hipMemcpy(text)");
    check(hipMemcpyAsync(g_buf.off, starts.data(), 4 * n, hipMemcpyHostToDevice, st),
          "hipMemcpy(offsets)");
    check(hipMemcpyAsync(g_buf.ptrs, strings, 8 * n, hipMemcpyHostToDevice, st),
          "hipMemcpy(pointers)");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
  }
  check(hipMemsetAsync(g_buf.text + text_len, 0, 64, st), "hipMemset(pad)");

  const int rc = sort_device32(g_buf.text, text_len, g_buf.off, g_buf.out,
                               lcp_out ? g_buf.lcp : nullptr, n, st);
  if (rc) die(std::string("device sort failed: ") + gss_last_error());

  if (use_span) {  // the sorted offsets come back and are widened on the host
    g_staging.d2h_pointers(strings, g_buf.out, n, lo, st);
  } else {
    check(launch_off_to_orig_ptr(g_buf.out, g_buf.off, g_buf.ptrs, g_buf.ptrs2, n, st),
          "off_to_orig_ptr");
    g_staging.d2h(strings, g_buf.ptrs2, 8 * n, st);
  }
  if (lcp_out) g_staging.d2h(lcp_out, g_buf.lcp, 8 * (n - 1), st);
}

}  // namespace
}  // namespace gss

extern "C" {

void gss_msd(unsigned char** strings, size_t n) { gss::sort_pointers(strings, n, nullptr); }

void gss_msd_lcp(unsigned char** strings, size_t n, size_t* lcp) {
  gss::sort_pointers(strings, n, lcp);
}

}  // extern "C"
