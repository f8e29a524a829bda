// Host-side internals shared by the device-resident sorter and the pointer
// (routine.h) front end.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "gss.h"

namespace gss {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

// Sorts n 32-bit offsets into `text` (device memory) into `out` on the
// current device's default context; see gss_sort_device32 in gss.h.
int sort_device32(const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
                  uint32_t* out, uint64_t* lcp, uint64_t n, hipStream_t stream);

// The default context's stream on the current device (created on first use).
int library_stream(hipStream_t* out);

// Device bytes of a context's workspace sized for n strings under the
// current tuning.
size_t workspace_bytes(uint64_t n);

// Sizes a context's workspace for n strings and its initial request lists
// for nreq requests (0: leave them).
int ctx_reserve(gss_ctx* ctx, uint64_t n, uint32_t nreq);

}  // namespace gss
