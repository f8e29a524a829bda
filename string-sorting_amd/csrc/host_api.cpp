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

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
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

  uintptr_t lo = UINTPTR_MAX, hi = 0;
  for (size_t i = 0; i < n; ++i) {
    const uintptr_t p = reinterpret_cast<uintptr_t>(strings[i]);
    if (!p) die("NULL string pointer");
    lo = p < lo ? p : lo;
    hi = p > hi ? p : hi;
  }
  const uintptr_t end = hi + strlen(reinterpret_cast<const char*>(hi)) + 1;
  const uint64_t span = end - lo;
  const bool use_span = span < (1ull << 32) - 64 && span_readable(lo, end);

  ensure_n(n);
  uint64_t text_len = 0;
  if (use_span) {
    text_len = span;
    ensure_text(text_len + 64);
    check(hipMemcpyAsync(g_buf.text, reinterpret_cast<const void*>(lo), span,
                         hipMemcpyHostToDevice, st),
          "hipMemcpy(text)");
    check(hipMemcpyAsync(g_buf.ptrs, strings, 8 * n, hipMemcpyHostToDevice, st),
          "hipMemcpy(pointers)");
    check(launch_ptr_to_off(g_buf.ptrs, lo, g_buf.off, n, st), "ptr_to_off");
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

  uint64_t* result = g_buf.ptrs;
  if (use_span) {
    check(launch_off_to_ptr(g_buf.out, lo, g_buf.ptrs, n, st), "off_to_ptr");
  } else {
    check(launch_off_to_orig_ptr(g_buf.out, g_buf.off, g_buf.ptrs, g_buf.ptrs2, n, st),
          "off_to_orig_ptr");
    result = g_buf.ptrs2;
  }
  check(hipMemcpyAsync(strings, result, 8 * n, hipMemcpyDeviceToHost, st), "hipMemcpy(result)");
  if (lcp_out)
    check(hipMemcpyAsync(lcp_out, g_buf.lcp, 8 * (n - 1), hipMemcpyDeviceToHost, st),
          "hipMemcpy(lcp)");
  check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

}  // namespace
}  // namespace gss

extern "C" {

void gss_msd(unsigned char** strings, size_t n) { gss::sort_pointers(strings, n, nullptr); }

void gss_msd_lcp(unsigned char** strings, size_t n, size_t* lcp) {
  gss::sort_pointers(strings, n, lcp);
}

}  // extern "C"
