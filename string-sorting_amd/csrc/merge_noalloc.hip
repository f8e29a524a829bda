// Allocation-free LCP merge (gss_merge_lcp_device32_noalloc in gss.h).
//
// The 2-way LCP merge needs one scratch array besides its inputs and
// outputs: the merge-path split of every 4096-output tile.  Here it lives in
// a fixed device array of the library instead of the per-device grow-only
// scratch, so a call allocates nothing (no hipMalloc, hipFree, pinned host
// memory or torch allocation) and never synchronizes: it only queues the two
// merge kernels on the caller's stream.  That suits callers that cannot
// allocate or block, e.g. code under HIP graph capture or a tight serving
// loop.
//
// Every call on one GPU shares that GPU's copy of the array, which keeps the
// last call's splits (each call writes every entry it reads, so nothing needs
// clearing): calls whose kernels overlap on one GPU (different streams, or
// host threads racing on one stream) race on it.  Calls queued in order on
// one stream are safe.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "gss.h"
#include "gss_internal.h"
#include "host.h"

namespace gss {

// Split entries of the largest merge this entry accepts: 262,144 entries of
// 8 bytes (2 MiB per GPU) cover merges of up to 262,143 tiles of 2048
// outputs, 536,868,864 strings.
constexpr uint64_t kNoAllocSplits = 262144;
constexpr uint64_t kNoAllocMaxStrings = (kNoAllocSplits - 1) * kMergeTile;
static_assert(kNoAllocMaxStrings == 536868864ull, "the limit gss.h states");

// The scratch array.  In host code `scratch` is only a stand-in for the
// symbol: its device address on each GPU comes from hipGetSymbolAddress.
static __device__ __align__(256) uint64_t scratch[kNoAllocSplits];

namespace {

constexpr int kMaxDevices = 16;
std::mutex g_mu;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device
// with that device current.
int device_scratch(uint64_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(GSS_E_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(GSS_E_DEVICE, "device id out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GSS_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                    hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return GSS_OK;
}

}  // namespace
}  // namespace gss

using namespace gss;

extern "C" {

int gss_merge_lcp_device32_noalloc(const unsigned char* d_text, uint64_t text_bytes,
                                   const uint32_t* d_a, const uint64_t* d_a_lcp, uint64_t na,
                                   const uint32_t* d_b, const uint64_t* d_b_lcp, uint64_t nb,
                                   uint32_t* d_out, uint64_t* d_out_lcp, void* stream) {
  const uint64_t total = na + nb;
  if (total == 0) return GSS_OK;
  int rc = merge_check(d_text, text_bytes, d_a, d_a_lcp, na, d_b, d_b_lcp, nb, d_out);
  if (rc) return rc;
  if (total > kNoAllocMaxStrings || merge_splits(total) > kNoAllocSplits)
    return fail(GSS_E_TOO_LARGE, "the allocation-free merge takes at most " +
                                     std::to_string(kNoAllocMaxStrings) + " strings (got " +
                                     std::to_string(total) + ")");
  uint64_t* split = nullptr;
  rc = device_scratch(&split);
  if (rc) return rc;
  const hipError_t e = merge_launch(reinterpret_cast<const uint64_t*>(d_text), text_bytes, d_a,
                                    d_a_lcp, na, d_b, d_b_lcp, nb, d_out, d_out_lcp, split,
                                    reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(GSS_E_DEVICE, std::string("merge launch: ") + hipGetErrorString(e));
  return GSS_OK;
}

}  // extern "C"
