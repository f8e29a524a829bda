// Autotuner: picks the local threshold (strings a workgroup finishes), the
// digit width and the tile size of the global passes, and the first-round
// key window of the largest local class per input signature (no reference
// counterpart: the reference fixes its thresholds at compile time, e.g.
// src/msd_ce.cpp:79 and :216, and its cached key width, src/msd_a.cpp:39,
// src/multikey_cache.cpp:39-55).
//
// The signature of an input is (mean string length, alphabet size, count),
// each as a power-of-two bucket, estimated from 512 strings at evenly spaced
// positions of the offset array.  Tuning an input sorts a strided sample of
// it (at most 2M strings, which keeps its length distribution and duplicate
// ratio) with the candidates and keeps the fastest: every threshold x digit
// width with the current tile and key window, then the best of those with
// each other tile, then with each other key window (the later knobs are
// tuned around the earlier ones rather than in all 32 combinations); the
// result is cached per signature.  When enabled, gss_sort_device32 uses the cached tuning of
// its input's signature and tunes signatures it has not seen yet.

#include <chrono>
#include <map>
#include <mutex>
#include <string>

#include "gss_internal.h"
#include "host.h"

namespace gss {

namespace {

constexpr uint32_t kSigStrings = 512;        // strings read for a signature
constexpr uint32_t kSigLenCap = 256;         // bytes scanned per string
constexpr uint64_t kTrialStrings = 1u << 21;  // strings sorted per candidate
constexpr uint64_t kAutoMin = 1u << 16;       // smaller inputs keep the current tuning
constexpr int kTimedTrials = 2;

constexpr uint32_t kThresholds[] = {1024, 2048, 4096, 6144};
constexpr uint32_t kRadixBits[] = {11, 8};
constexpr uint32_t kTiles[] = {4096, 2048};
constexpr uint32_t kKeyBytes[] = {64, 32};

struct SigOut {
  uint32_t alpha[8];            // byte values seen in the first 32 bytes
  unsigned long long len_sum;   // lengths, each capped at kSigLenCap
  uint32_t count;
  uint32_t pad;
};

// One block of kSigStrings threads: thread i reads string i * n / m.
__global__ __launch_bounds__(kSigStrings) void k_signature(const uint64_t* __restrict__ t64,
                                                          uint64_t tlen,
                                                          const uint32_t* __restrict__ off,
                                                          uint64_t n, SigOut* __restrict__ res) {
  __shared__ uint32_t s_alpha[8];
  __shared__ unsigned long long s_len;
  __shared__ uint32_t s_count;
  if (threadIdx.x < 8) s_alpha[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    s_len = 0;
    s_count = 0;
  }
  __syncthreads();
  const uint32_t m = n < kSigStrings ? (uint32_t)n : kSigStrings;
  if (threadIdx.x < m) {
    const uint64_t o = off[(uint64_t)threadIdx.x * n / m];
    uint32_t len = 0;
    bool ended = false;
    for (uint32_t k = 0; k < kSigLenCap && !ended; k += 8) {
      const uint64_t key = load_key(t64, tlen, o + k);  // big-endian, zero after NUL
      const uint32_t l = key_len(key);                    // bytes before the NUL, or 8
      len += l;
      ended = l < 8;
      if (k < 32)
        for (uint32_t b = 0; b < l; ++b) {
          const uint32_t v = (uint32_t)(key >> (56 - 8 * b)) & 0xFFu;
          atomicOr(&s_alpha[v >> 5], 1u << (v & 31));
        }
    }
    atomicAdd(&s_len, (unsigned long long)len);
    atomicAdd(&s_count, 1u);
  }
  __syncthreads();
  if (threadIdx.x < 8) res->alpha[threadIdx.x] = s_alpha[threadIdx.x];
  if (threadIdx.x == 0) {
    res->len_sum = s_len;
    res->count = s_count;
  }
}

// out[i] = in[i * n / m]: a strided sample that keeps the input's mix.
__global__ void k_sample(const uint32_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ out,
                         uint64_t m) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
    out[i] = in[i * n / m];
}

uint32_t log2_bucket(uint64_t v) {
  uint32_t b = 0;
  while ((1ull << (b + 1)) <= v && b < 63) ++b;
  return b;
}

struct State {
  std::mutex mu;
  bool enabled = false;
  std::map<uint32_t, gss_tuning> cache;
  SigOut* d_sig = nullptr;  // signature buffers, on `device`
  SigOut* h_sig = nullptr;
  int device = -1;
};

State g_state;

#define AT_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(GSS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Signature of the input (called with g_state.mu held).
int signature(const unsigned char* text, uint64_t text_bytes, const uint32_t* in, uint64_t n,
              hipStream_t st, uint32_t* sig) {
  int dev = -1;
  AT_HIP(hipGetDevice(&dev));
  if (g_state.device != dev) {
    if (g_state.d_sig) (void)hipFree(g_state.d_sig);
    if (g_state.h_sig) (void)hipHostFree(g_state.h_sig);
    g_state.d_sig = nullptr;
    g_state.h_sig = nullptr;
    AT_HIP(hipMalloc((void**)&g_state.d_sig, sizeof(SigOut)));
    AT_HIP(hipHostMalloc((void**)&g_state.h_sig, sizeof(SigOut), hipHostMallocDefault));
    g_state.device = dev;
  }
  k_signature<<<1, kSigStrings, 0, st>>>(reinterpret_cast<const uint64_t*>(text),
                                         text_bytes, in, n, g_state.d_sig);
  AT_HIP(hipGetLastError());
  AT_HIP(hipMemcpyAsync(g_state.h_sig, g_state.d_sig, sizeof(SigOut), hipMemcpyDeviceToHost, st));
  AT_HIP(hipStreamSynchronize(st));
  const SigOut& r = *g_state.h_sig;
  uint32_t alpha = 0;
  for (int w = 0; w < 8; ++w) alpha += __builtin_popcount(r.alpha[w]);
  const uint64_t mean_len = r.count ? r.len_sum / r.count : 0;
  *sig = log2_bucket(mean_len + 1) | (log2_bucket(alpha + 1) << 8) | (log2_bucket(n) << 16);
  return GSS_OK;
}

// Times every candidate on a strided sample of the input.
int run_trials(const unsigned char* text, uint64_t text_bytes, const uint32_t* in, uint64_t n,
               hipStream_t st, gss_autotune_result* res) {
  const uint64_t m = n < kTrialStrings ? n : kTrialStrings;
  uint32_t* buf = nullptr;
  if (hipMalloc((void**)&buf, 8 * (m ? m : 1)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(GSS_E_NOMEM, "autotuner sample buffers");
  }
  uint32_t* sample = buf;
  uint32_t* out = buf + m;
  int rc = GSS_OK;
  const gss_tuning def = current_tuning();
  res->best = def;
  res->best_ms = 1e30;
  res->default_ms = 0;
  res->sample = m;
  res->candidates = 0;
  k_sample<<<(uint32_t)((m + 255) / 256 < 4096 ? (m + 255) / 256 : 4096), 256, 0, st>>>(in, n,
                                                                                      sample, m);
  if (hipGetLastError() != hipSuccess) rc = fail(GSS_E_DEVICE, "autotuner sample launch");
  // The trial sorts may run on another stream: the sample must be complete.
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(GSS_E_DEVICE, "autotuner sample");
  // Times one candidate (a warm-up, then the best of kTimedTrials sorts).
  auto trial = [&](const gss_tuning& cand) {
    double best = 1e30;
    for (int t = 0; t <= kTimedTrials && !rc; ++t) {  // first run is a warm-up
      const auto t0 = std::chrono::steady_clock::now();
      rc = sort_device32_tuned(text, text_bytes, sample, out, m, st, cand);
      const double ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
              .count();
      if (t > 0 && ms < best) best = ms;
    }
    if (rc) return;
    ++res->candidates;
    if (cand.local_threshold == def.local_threshold && cand.radix_bits == def.radix_bits &&
        cand.tile == def.tile && cand.key_bytes == def.key_bytes)
      res->default_ms = best;
    if (best < res->best_ms) {
      res->best_ms = best;
      res->best = cand;
    }
  };
  for (uint32_t rb : kRadixBits)
    for (uint32_t th : kThresholds)
      if (!rc) trial(gss_tuning{th, rb, def.tile, def.key_bytes});
  for (uint32_t tile : kTiles)
    if (!rc && tile != res->best.tile) trial(gss_tuning{res->best.local_threshold,
                                                        res->best.radix_bits, tile,
                                                        res->best.key_bytes});
  for (uint32_t kb : kKeyBytes)
    if (!rc && kb != res->best.key_bytes) trial(gss_tuning{res->best.local_threshold,
                                                           res->best.radix_bits, res->best.tile,
                                                           kb});
  (void)hipStreamSynchronize(st);
  (void)hipFree(buf);
  return rc;
}

}  // namespace

int autotune_choose(const unsigned char* text, uint64_t text_bytes, const uint32_t* in,
                    uint64_t n, hipStream_t stream, gss_tuning* out, bool* use) {
  *use = false;
  if (n < kAutoMin || !text || !in) return GSS_OK;
  std::lock_guard<std::mutex> lk(g_state.mu);
  if (!g_state.enabled) return GSS_OK;
  uint32_t sig = 0;
  int rc = signature(text, text_bytes, in, n, stream, &sig);
  if (rc) return rc;
  auto it = g_state.cache.find(sig);
  if (it == g_state.cache.end()) {
    gss_autotune_result r{};
    rc = run_trials(text, text_bytes, in, n, stream, &r);
    if (rc) return rc;
    it = g_state.cache.emplace(sig, r.best).first;
  }
  *out = it->second;
  *use = true;
  return GSS_OK;
}

}  // namespace gss

using namespace gss;

extern "C" {

int gss_autotune_device32(const unsigned char* d_text, uint64_t text_bytes, const uint32_t* d_off,
                          uint64_t n, gss_autotune_result* out, void* stream) {
  if (!out) return fail(GSS_E_INVALID, "NULL result pointer");
  *out = gss_autotune_result{};
  if (!d_text || !d_off) return fail(GSS_E_INVALID, "NULL text or offsets");
  if (n < 2) return fail(GSS_E_INVALID, "nothing to tune on fewer than two strings");
  if (text_bytes == 0) return fail(GSS_E_INVALID, "empty text: every string starts inside it");
  if (reinterpret_cast<uintptr_t>(d_text) & 7)
    return fail(GSS_E_INVALID, "the device text must be 8-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_state.mu);
  uint32_t sig = 0;
  int rc = signature(d_text, text_bytes, d_off, n, st, &sig);
  if (rc) return rc;
  rc = run_trials(d_text, text_bytes, d_off, n, st, out);
  if (rc) return rc;
  out->signature = sig;
  g_state.cache[sig] = out->best;
  return GSS_OK;
}

int gss_autotune_enable(int on) {
  std::lock_guard<std::mutex> lk(g_state.mu);
  g_state.enabled = on != 0;
  return GSS_OK;
}

int gss_autotune_reset(void) {
  std::lock_guard<std::mutex> lk(g_state.mu);
  g_state.cache.clear();
  return GSS_OK;
}

int gss_autotune_cached(void) {
  std::lock_guard<std::mutex> lk(g_state.mu);
  return (int)g_state.cache.size();
}

}  // extern "C"
