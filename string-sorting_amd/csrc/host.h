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

// Request depth of sort_spans_device32 for a request left as it is.
constexpr uint32_t kSpanKeep = 0xFFFFFFFFu;

// Sorts, in place on the default context, the requests [req[r], req[r+1])
// of `off` (n = req[nreq] offsets) whose strings all share their first
// req_depth[r] bytes, starting there; requests with depth kSpanKeep keep
// their offsets and LCPs.  LCPs are written for the pairs inside each
// sorted request only.
int sort_spans_device32(const unsigned char* text, uint64_t text_bytes, uint32_t* off,
                        uint64_t* lcp, const uint64_t* req, const uint32_t* req_depth,
                        uint32_t nreq, hipStream_t stream);

// Waits for everything queued on `st` by polling it.  A blocking wait
// sleeps once the stream has been busy for a while and wakes tens of
// microseconds after the GPU finished, which a request's latency would pay.
int wait_stream(hipStream_t st);

// The default context's stream on the current device (created on first use).
int library_stream(hipStream_t* out);

// As sort_device32 on the default context, with an explicit tuning (the
// autotuner's trial sorts).
int sort_device32_tuned(const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
                        uint32_t* out, uint64_t n, hipStream_t stream, const gss_tuning& tuning);

// GSS_OK when the tuning is supported (else the GSS_E_INVALID of gss_set_tuning).
int check_tuning(const gss_tuning& t);
gss_tuning current_tuning();

// When the autotuner is on and n is large enough: *use = true and *out = the
// tuning for the input's signature (tuned on first sight).  Else *use = false.
int autotune_choose(const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
                    uint64_t n, hipStream_t stream, gss_tuning* out, bool* use);

// Device bytes of a context's workspace sized for n strings under `tuning`.
size_t workspace_bytes(uint64_t n, const gss_tuning& tuning);

// Sizes a context's workspace for n strings under `tuning` and its initial
// request lists for nreq requests (0: leave them).
int ctx_reserve(gss_ctx* ctx, uint64_t n, uint32_t nreq, const gss_tuning& tuning);

// gss_ctx_sort_batch_device32 with an explicit tuning instead of the current
// one (the served loop keeps the tuning its slots were sized for).
int ctx_sort_batch_tuned(gss_ctx* ctx, const unsigned char* text, uint64_t text_bytes,
                         const uint32_t* in, uint32_t* out, uint64_t* lcp, const uint64_t* req,
                         uint32_t nreq, hipStream_t stream, const gss_tuning& tuning);

}  // namespace gss
