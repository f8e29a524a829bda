// HIP kernels of the global MSD passes of the MI355X string sort (gfx950,
// wave64), plus small helpers.
//
// A sort runs global MSD passes over segments too large for one workgroup,
// then finishes every remaining segment inside one workgroup
// (local_sort.hip).  One global pass over the BigSeg list, tiles of p.tile
// strings that never span two segments:
//
//   gather_minmax  key = the 8 bytes at each string's depth (big-endian, NUL
//                  padded), cached in `keys` with the 8 bytes after it in
//                  `ext` -- or, every other pass, taken as the previous
//                  scatter left it in `keys`; per-tile min/max key and set
//                  of byte values seen.
//   segreduce      per segment: min/max key and byte set over its tiles.
//   layout         the pass alphabet; per segment its digit (first differing byte + k ranked
//                  characters, see digit_spec) and the place of its
//                  digit-major count matrix in the flat array.
//   histogram      per tile digit counts -> flat[base + d*ntiles + t].
//   scan (x3)      exclusive scan of the whole flat array: entry (d, t) now
//                  holds where tile t's strings of digit d go.
//   segscan        one workgroup per segment: bucket sizes, child segments
//                  (next pass, local, or final), retire runs, boundaries.
//   scatter        stable partition of the offsets by digit: wave
//                  match-ranking in LDS for the in-tile rank, staged in LDS
//                  so the global writes are contiguous runs; after a
//                  text-reading gather it also writes each string bound for
//                  the next pass its key at the child's depth.
//
//   boundary_lcp   LCP of each new adjacent pair of child buckets.
//
// Blocks of the tile kernels take tiles in XCD-aware order (xcd_tile).
// A later pass needs nothing from the host: a one-workgroup set-up kernel
// (plan) takes the segment list as the previous pass left it, numbers the
// tiles and writes the tile table (the host builds the first pass's); the
// pass's other kernels are launched with an upper bound of blocks and read
// the real tile or segment count from the device counters.  The host reads
// the counters when it needs them (publish writes them to pinned memory and
// then a sequence word that the host polls): after the queued passes, to
// size the local phase, and at the end.
//
// A segment whose keys are all equal costs the gather only: it keeps its
// place and buffer and moves 8 bytes deeper (or retires if the key holds the
// NUL, i.e. all its strings are identical).
//
// The reference's byte-at-a-time MSD (msd_CE2, src/msd_ce.cpp:143-178) and
// the 8-byte cache rules of multikey_cache (src/multikey_cache.cpp:107-125,
// 207-211) define the order and the retirement rule used here.

#include "gss_internal.h"

namespace gss {

// ---------------------------------------------------------------------------
// Small helpers

// Packed texts: offsets were assigned in input order, so the input index of
// a sorted offset is found by binary search in the ascending start array.
__global__ void k_off_to_index(const uint32_t* __restrict__ off,
                               const uint32_t* __restrict__ starts, uint32_t* __restrict__ idx,
                               uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t o = off[i];
    uint64_t lo = 0, hi = n;  // first index with starts[idx] >= o
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (starts[mid] < o) lo = mid + 1; else hi = mid;
    }
    idx[i] = (uint32_t)lo;
  }
}

// Host-to-device upload of a sort's tables from fine-grained pinned host
// memory, up to kUploadRanges ranges in one launch (blockIdx.y = range).
// Words [0, nzero) of a range are cleared, words [nzero, nwords) read from
// the host with system-scope loads, so a launch always sees what the host
// wrote before it whatever the GPU caches hold of an earlier upload.  Being
// a kernel on the sort's own stream, it needs no copy-engine hand-off
// between the passes.
__global__ __launch_bounds__(256) void k_upload(UploadBatch b) {
  const UploadRange r = b.r[blockIdx.y];
  const uint32_t stride = gridDim.x * 256;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < r.nwords; i += stride)
    r.dst[i] = i < r.nzero
                   ? 0ull
                   : __hip_atomic_load(r.src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Device-to-host hand-over: words [0, na) of a, then the first
// min(*count, cap) records (rec_words words each) of b, are written to the
// fine-grained pinned block dst with system-scope stores; once every store
// has completed, the sequence number goes to *flag.  The host polls the
// flag instead of waiting for a copy and a stream: both would reach it tens
// of microseconds later.  Each lane loads kPublishWords words before it
// stores any (one memory round trip per 4096 words), each wave's release
// fence waits for its own stores, and the barrier puts the flag after all.
constexpr int kPublishThreads = 256;
constexpr int kPublishWords = 16;

__global__ __launch_bounds__(kPublishThreads) void k_publish(
    const uint64_t* __restrict__ a, uint32_t na, const uint64_t* __restrict__ b,
    const uint32_t* __restrict__ count, uint32_t cap, uint32_t rec_words, uint64_t* dst,
    uint64_t* flag, uint64_t seq) {
  const uint32_t total = na + (b ? min(*count, cap) * rec_words : 0u);
  for (uint32_t i0 = 0; i0 < total; i0 += kPublishThreads * kPublishWords) {
    uint64_t v[kPublishWords];
#pragma unroll
    for (int q = 0; q < kPublishWords; ++q) {
      // Clamped to the last word, so every load is issued unconditionally.
      const uint32_t i = min(i0 + q * kPublishThreads + threadIdx.x, total - 1);
      v[q] = i < na ? a[i] : b[i - na];
    }
#pragma unroll
    for (int q = 0; q < kPublishWords; ++q) {
      const uint32_t i = i0 + q * kPublishThreads + threadIdx.x;
      if (i < total) __hip_atomic_store(dst + i, v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(256) void k_copy_chunks(const CopyChunk* __restrict__ chunks) {
  const CopyChunk c = chunks[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < c.words; i += 256) c.dst[i] = c.src[i];
}

// ---------------------------------------------------------------------------
// Global pass, step 0 (one workgroup): the pass's segments are the list the
// previous pass (or the start of the sort) left, big_next of them.  Segment
// i gets ceil(size / tile) tiles, numbered in segment order (tile0 = the
// exclusive scan of the counts); tile t finds its segment by a binary search
// over tile0.  The counts go to big_cur / ntiles for the pass's kernels, the
// boundaries emitted so far are marked done (this pass's boundary kernel
// takes the ones it adds), and the pass's active strings go to the
// statistics.
constexpr int kPlanThreads = 1024;
constexpr uint32_t kPlanLdsSegs = 8192;  // segments whose first tiles the search finds in LDS

__global__ __launch_bounds__(kPlanThreads) void k_plan(Plan p, const BigSeg* segs, uint32_t pass,
                                                      uint32_t grid_tiles, uint32_t grid_segs) {
  __shared__ uint32_t s_scan[kPlanThreads / kWave + 1];
  __shared__ unsigned long long s_act[kPlanThreads / kWave];
  __shared__ uint32_t s_tile0[kPlanLdsSegs];
  const uint32_t nsegs = min(p.ctr->big_next, p.cap_big);
  const bool in_lds = nsegs <= kPlanLdsSegs;
  uint32_t carry = 0;
  unsigned long long act = 0;
  for (uint32_t s0 = 0; s0 < nsegs; s0 += kPlanThreads) {
    const uint32_t i = s0 + threadIdx.x;
    uint32_t nt = 0;
    if (i < nsegs) {
      const uint32_t sz = segs[i].size;
      nt = (sz + p.tile - 1) / p.tile;
      act += sz;
    }
    uint32_t total;
    const uint32_t pre = block_excl_scan<kPlanThreads>(nt, s_scan, &total);
    if (i < nsegs) {
      p.aux[i] = SegAux{0, nt, 0, carry + pre};
      if (in_lds) s_tile0[i] = carry + pre;
    }
    carry += total;
  }
  const uint32_t ntiles = min(carry, p.cap_tiles);
  __syncthreads();  // the first tiles are visible to the whole workgroup
  for (uint32_t t = threadIdx.x; t < ntiles; t += kPlanThreads) {
    uint32_t lo = 0, hi = nsegs - 1;  // the last segment whose first tile is <= t
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if ((in_lds ? s_tile0[mid] : p.aux[mid].tile0) <= t) lo = mid; else hi = mid - 1;
    }
    const BigSeg sg = segs[lo];
    const uint32_t k = t - (in_lds ? s_tile0[lo] : p.aux[lo].tile0);
    p.tiles[t] = Tile{lo, sg.begin + k * p.tile, min(p.tile, sg.size - k * p.tile), k};
  }
  // Active strings of the pass: a block sum of the 64-bit partial sums.
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) act += __shfl_xor(act, o, kWave);
  if (lane_id() == 0) s_act[threadIdx.x / kWave] = act;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kPlanThreads / kWave; ++w) tot += s_act[w];
    // The pass's kernels are launched with grid_tiles / grid_segs blocks:
    // counts beyond them would leave work undone.
    if (carry > p.cap_tiles || carry > grid_tiles || nsegs > grid_segs)
      atomicOr(&p.ctr->error, 128u);
    p.ctr->big_cur = nsegs;
    p.ctr->ntiles = ntiles;
    p.ctr->active_cur = (uint32_t)tot;
    p.ctr->bound_done = p.ctr->boundary;
    if (pass < (uint32_t)kMaxPasses) p.pass_active[pass] = tot;
  }
}



// ---------------------------------------------------------------------------
// Digits of a global pass.
//
// A pass records which byte values occur in its keys (the alphabet, A
// symbols) and ranks them; rank order is byte order, so a digit formed as
// the mixed-radix number of k consecutive ranks orders strings exactly as
// their bytes do.  k is the largest count with A^k <= 2^radix_bits, e.g. 2
// characters of [a-z0-9./:-] or 4 of {A,C,G,T} per 11-bit digit.  A segment's
// digit starts at the first key byte where its min and max key differ.

struct DigitSpec {
  uint32_t fb0;  // first key byte that differs inside the segment
  uint32_t k;    // characters in the digit
  uint32_t A;    // alphabet size of the pass
  uint32_t nbu;  // digit values in use: A^k
};

__device__ __forceinline__ uint32_t alpha_size(const uint32_t* __restrict__ alpha) {
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < 8; ++w) total += __popc(alpha[w]);
  return total;
}

// Fills rank[0..256) from the pass alphabet and returns its size.  All
// threads of the block call it; rank is ready after the next barrier.
__device__ __forceinline__ uint32_t build_rank(const uint32_t* __restrict__ alpha,
                                               uint8_t* rank) {
  uint32_t words[8];
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < 8; ++w) {
    words[w] = alpha[w];
    total += __popc(words[w]);
  }
  for (uint32_t t = threadIdx.x; t < 256; t += blockDim.x) {
    uint32_t r = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w)
      if ((uint32_t)w < (t >> 5)) r += __popc(words[w]);
    r += __popc(words[t >> 5] & ((1u << (t & 31)) - 1u));
    rank[t] = (uint8_t)r;
  }
  return total;
}

// The digit of a segment whose keys first differ at byte fb0.
template <int RB>
__device__ __forceinline__ DigitSpec digit_spec_at(uint32_t fb0, uint32_t A) {
  DigitSpec ds;
  ds.fb0 = fb0;
  ds.A = A < 2 ? 2 : A;
  uint32_t k = 1, v = ds.A;
  while (k < 8 && (uint64_t)v * ds.A <= (1ull << RB)) {
    v *= ds.A;
    ++k;
  }
  ds.k = min(k, 8 - ds.fb0);
  uint32_t nbu = 1;
  for (uint32_t c = 0; c < ds.k; ++c) nbu *= ds.A;
  ds.nbu = nbu;
  return ds;
}

template <int RB>
__device__ __forceinline__ DigitSpec digit_spec(uint64_t mn, uint64_t mx, uint32_t A) {
  return digit_spec_at<RB>((uint32_t)__builtin_clzll(mn ^ mx) >> 3, A);
}

__device__ __forceinline__ uint32_t digit_of(uint64_t key, const DigitSpec& ds,
                                             const uint8_t* rank) {
  uint32_t d = 0;
  for (uint32_t c = 0; c < ds.k; ++c)
    d = d * ds.A + rank[(uint32_t)(key >> (56 - 8 * (ds.fb0 + c))) & 0xFFu];
  return d;
}

// Bits needed to tell digits below nbu apart.
__device__ __forceinline__ uint32_t digit_bits(uint32_t nbu) {
  return nbu <= 1 ? 1u : 32u - (uint32_t)__builtin_clz(nbu - 1);
}

// Lanes of the wave whose digit equals this lane's, among `live`; the digit
// has `bits` significant bits (wave-uniform).
__device__ __forceinline__ uint64_t match_digit_n(uint32_t d, uint64_t live, uint32_t bits) {
  uint64_t m = live;
  for (uint32_t b = 0; b < bits; ++b) {
    const uint64_t bal = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

// ---------------------------------------------------------------------------
// Global pass, step 1: gather keys; per tile the min/max key and the set of
// byte values seen (no atomics: thousands of tiles would serialise on one
// segment's words).  A first pass over one segment (fuse_hist) also counts
// the tile's digits as the guess forms them, into the segment's count
// column (its counts start the flat array); k_layout tells k_histogram
// whether they stand.

template <int RB, int TI>
__device__ __forceinline__ void gather_tile(Plan& p, const BigSeg* segs, uint32_t b,
                                            uint32_t nt) {
  constexpr uint32_t NB = 1u << RB;
  __shared__ uint64_t s_mn[kTileThreads / kWave], s_mx[kTileThreads / kWave];
  __shared__ uint8_t s_present[256];
  __shared__ uint32_t s_h[NB];
  __shared__ uint8_t s_rank[256];
  if (threadIdx.x < 256) s_present[threadIdx.x] = 0;
  const uint32_t ti = xcd_tile(b, nt);
  const Tile t = p.tiles[ti];
  const BigSeg sg = segs[t.seg];
  const uint32_t* __restrict__ src = p.idx[sg.buf];
  uint32_t o[TI];
  uint64_t k[TI];
  uint64_t mn = ~0ull, mx = 0ull;
  if (p.keys_carried) {
    // The previous pass left these keys at the strings' positions.
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const uint32_t i = j * kTileThreads + threadIdx.x;
      k[j] = i < t.len ? p.keys[t.begin + i] : 0ull;
    }
  } else {
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const uint32_t i = j * kTileThreads + threadIdx.x;
      o[j] = i < t.len ? src[t.begin + i] : 0u;
    }
    // Unconditional (out-of-tile slots read offset 0) so all loads issue
    // before the first wait.  16 bytes: the key and the 8 bytes after it,
    // from which the scatter forms the next pass's keys.
    uint64_t x[TI];
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      uint64_t kk[2];
      load_keys<2>(p.text64, p.text_bytes, (uint64_t)o[j] + sg.depth, kk);
      k[j] = kk[0];
      x[j] = kk[1];
    }
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const uint32_t i = j * kTileThreads + threadIdx.x;
      if (i < t.len) {
        p.keys[t.begin + i] = k[j];
        p.ext[t.begin + i] = x[j];
      }
    }
  }
  // The guessed digit (fuse_hist is a kernel argument: uniform).
  bool fuse = false;
  DigitSpec ds{};
  if (p.fuse_hist && !p.keys_carried) {
    uint32_t g[kGuessWords];
#pragma unroll
    for (int w = 0; w < kGuessWords; ++w) g[w] = p.guess[w];
    fuse = g[9] != 0;
    if (fuse) {
      ds = digit_spec_at<RB>(g[8], build_rank(g, s_rank));
      for (uint32_t d = threadIdx.x; d < ds.nbu; d += kTileThreads) s_h[d] = 0;
    }
  }
  __syncthreads();  // s_present, s_h cleared; s_rank ready
#pragma unroll
  for (int j = 0; j < TI; ++j) {
    const uint32_t i = j * kTileThreads + threadIdx.x;
    if (i < t.len) {
      mn = k[j] < mn ? k[j] : mn;
      mx = k[j] > mx ? k[j] : mx;
#pragma unroll
      for (int b = 0; b < 8; ++b) s_present[(uint32_t)(k[j] >> (8 * b)) & 0xFFu] = 1;
      if (fuse) {
        // A byte outside the guessed alphabet ranks A and can form a digit
        // >= nbu (the guess then fails its check): not counted.
        const uint32_t d = digit_of(k[j], ds, s_rank);
        if (d < ds.nbu) atomicAdd(&s_h[d], 1u);
      }
    }
  }
  mn = wave_min64(mn);
  mx = wave_max64(mx);
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_mn[wid] = mn;
    s_mx[wid] = mx;
  }
  __syncthreads();
  // 256 presence flags folded into 8 words by ballots.
  if (threadIdx.x < 256) {
    const uint64_t bal = __ballot(s_present[threadIdx.x] != 0);
    if (lane_id() == 0) {
      uint32_t* ta = p.tile_alpha + 8ull * ti + threadIdx.x / 32;
      ta[0] = (uint32_t)bal;
      ta[1] = (uint32_t)(bal >> 32);
    }
  }
  if (threadIdx.x == 0) {
    for (int w = 1; w < kTileThreads / kWave; ++w) {
      mn = s_mn[w] < mn ? s_mn[w] : mn;
      mx = s_mx[w] > mx ? s_mx[w] : mx;
    }
    p.tile_mm[2ull * ti] = mn;
    p.tile_mm[2ull * ti + 1] = mx;
  }
  if (fuse) {  // the only segment's counts start at flat entry 0
    const uint32_t ntiles = p.aux[t.seg].ntiles;
    uint16_t* col = p.cnt + t.tidx;
    for (uint32_t d = threadIdx.x; d < ds.nbu; d += kTileThreads)
      col[(uint64_t)d * ntiles] = (uint16_t)s_h[d];
  }
}

// One tile per block; a launch may hold more blocks than the pass has tiles
// (an upper bound), the rest exit at once.  (A loop over tiles would let one
// launch take any count, but the compiler then keeps the per-thread indices
// of every item live across it: far more registers, lower occupancy.)
template <int RB, int TI>
__global__ __launch_bounds__(kTileThreads) void k_gather_minmax(Plan p, const BigSeg* segs) {
  const uint32_t nt = p.ctr->ntiles;
  if (blockIdx.x < nt) gather_tile<RB, TI>(p, segs, blockIdx.x, nt);
}

// One workgroup per segment: min/max key and byte set over its tiles.  A
// first pass's single segment has thousands of tiles: 1024 threads with
// unrolled loads keep its latency down.
constexpr int kSegReduceThreads = 1024;

__device__ __forceinline__ void segreduce_seg(Plan& p, uint32_t si) {
  constexpr int T = kSegReduceThreads;
  __shared__ uint64_t s_mn[T / kWave], s_mx[T / kWave];
  __shared__ uint32_t s_al[8];
  const SegAux ax = p.aux[si];
  if (threadIdx.x < 8) s_al[threadIdx.x] = 0;
  uint64_t mn = ~0ull, mx = 0ull;
  uint32_t al = 0;
  const uint64_t* __restrict__ mm = p.tile_mm + 2ull * ax.tile0;
#pragma unroll 4
  for (uint32_t t = threadIdx.x; t < ax.ntiles; t += T) {
    const uint64_t a = mm[2ull * t], b = mm[2ull * t + 1];
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  // Byte sets: thread i ORs word i & 7 of tiles i / 8, i / 8 + T / 8, ...
  const uint32_t* __restrict__ ta = p.tile_alpha + 8ull * ax.tile0;
#pragma unroll 8
  for (uint32_t i = threadIdx.x; i < 8 * ax.ntiles; i += T) al |= ta[i];
  mn = wave_min64(mn);
  mx = wave_max64(mx);
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_mn[wid] = mn;
    s_mx[wid] = mx;
  }
  __syncthreads();
  static_assert(T % 8 == 0, "word index of i is i & 7");
  if (al) atomicOr(&s_al[threadIdx.x & 7], al);
  __syncthreads();
  if (threadIdx.x < 8) p.seg_alpha[8ull * si + threadIdx.x] = s_al[threadIdx.x];
  if (threadIdx.x == 0) {
    for (int w = 1; w < T / kWave; ++w) {
      mn = s_mn[w] < mn ? s_mn[w] : mn;
      mx = s_mx[w] > mx ? s_mx[w] : mx;
    }
    p.segmm[2ull * si] = mn;
    p.segmm[2ull * si + 1] = mx;
  }
}

__global__ __launch_bounds__(kSegReduceThreads) void k_segreduce(Plan p) {
  if (blockIdx.x < p.ctr->big_cur) segreduce_seg(p, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Global pass, step 2: digit ranges and the flat layout of the counts.

template <int RB>
__global__ __launch_bounds__(1024) void k_layout(Plan p) {
  __shared__ uint32_t s_scan[1024 / kWave + 1];
  __shared__ uint32_t s_al[8];
  const uint32_t nsegs = p.ctr->big_cur;
  // The pass alphabet: union of the segments' byte sets.
  if (threadIdx.x < 8) s_al[threadIdx.x] = 0;
  __syncthreads();
  uint32_t al = 0;
  for (uint32_t i = threadIdx.x; i < 8 * nsegs; i += 1024) al |= p.seg_alpha[i];
  if (al) atomicOr(&s_al[threadIdx.x & 7], al);  // 1024 % 8 == 0: word index = i & 7
  __syncthreads();
  if (threadIdx.x < 8) p.alpha[threadIdx.x] = s_al[threadIdx.x];
  const uint32_t A = alpha_size(s_al);
  uint32_t carry = 0;
  for (uint32_t s0 = 0; s0 < nsegs; s0 += 1024) {
    const uint32_t s = s0 + threadIdx.x;
    uint32_t entries = 0, nbu = 0;
    if (s < nsegs) {
      const uint64_t mn = p.segmm[2 * s], mx = p.segmm[2 * s + 1];
      nbu = mn == mx ? 0u : digit_spec<RB>(mn, mx, A).nbu;
      entries = nbu * p.aux[s].ntiles;
    }
    uint32_t total;
    const uint32_t pre = block_excl_scan<1024>(entries, s_scan, &total);
    if (s < nsegs) {
      p.aux[s].flat_base = carry + pre;
      p.aux[s].nbu = nbu;
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    // The gather's counts stand when this pass's one segment forms its
    // digit as guessed: the same alphabet and first differing byte.  The
    // pass's own values become the next guess.
    uint32_t hit = 0;
    if (p.fuse_hist && !p.keys_carried && nsegs == 1) {
      const uint64_t mn = p.segmm[0], mx = p.segmm[1];
      const uint32_t fb0 = mn == mx ? 0u : (uint32_t)__builtin_clzll(mn ^ mx) >> 3;
      hit = p.guess[9] != 0 && mn != mx && p.guess[8] == fb0;
#pragma unroll
      for (int w = 0; w < 8; ++w) {
        hit &= p.guess[w] == s_al[w];
        p.guess[w] = s_al[w];
      }
      p.guess[8] = fb0;
      p.guess[9] = mn != mx;
      p.ctr->hist_fused = hit;
    }
    p.ctr->hist_done = hit;
    p.ctr->flat_len = carry;
    p.ctr->big_next = 0;  // the coming segscan counts this pass's big children
    p.ctr->big_next_strings = 0;
  }
}

// ---------------------------------------------------------------------------
// Global pass, step 3: per-tile digit counts, digit-major per segment.

template <int RB, int TI>
__device__ __forceinline__ void histogram_tile(Plan& p, uint32_t b, uint32_t nt) {
  constexpr uint32_t NB = 1u << RB;
  __shared__ uint32_t h[NB];
  __shared__ uint8_t rank[256];
  const Tile t = p.tiles[xcd_tile(b, nt)];
  // Everything that needs only the tile is loaded before the segment's
  // min/max can end the block (a load after that return cannot be hoisted
  // above it): the keys (slots past the tile read its last key and are not
  // counted), the alphabet and the segment's count-matrix place.
  uint64_t key[TI];
#pragma unroll
  for (int j = 0; j < TI; ++j) {
    const uint32_t i = j * kTileThreads + threadIdx.x;
    key[j] = p.keys[t.begin + (i < t.len ? i : t.len - 1)];
  }
  uint32_t alpha[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) alpha[w] = p.alpha[w];
  const SegAux ax = p.aux[t.seg];
  const uint64_t mn = p.segmm[2 * t.seg], mx = p.segmm[2 * t.seg + 1];
  if (mn == mx) return;
  const DigitSpec ds = digit_spec<RB>(mn, mx, build_rank(alpha, rank));
  for (uint32_t d = threadIdx.x; d < ds.nbu; d += kTileThreads) h[d] = 0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < TI; ++j) {
    const uint32_t i = j * kTileThreads + threadIdx.x;
    if (i < t.len) atomicAdd(&h[digit_of(key[j], ds, rank)], 1u);
  }
  __syncthreads();
  uint16_t* col = p.cnt + ax.flat_base + t.tidx;
  for (uint32_t d = threadIdx.x; d < ds.nbu; d += kTileThreads)
    col[(uint64_t)d * ax.ntiles] = (uint16_t)h[d];
}

template <int RB, int TI>
__global__ __launch_bounds__(kTileThreads) void k_histogram(Plan p) {
  if (p.fuse_hist && p.ctr->hist_done) return;  // the gather counted them
  const uint32_t nt = p.ctr->ntiles;
  if (blockIdx.x < nt) histogram_tile<RB, TI>(p, blockIdx.x, nt);
}

// ---------------------------------------------------------------------------
// Global pass, step 4: exclusive scan of the 16-bit count array into the
// 32-bit flat array (block sums, their scan, block-local scans plus
// offsets).  flat[len] receives the total.

constexpr int kScanThreads = 256;
constexpr int kScanItems = kScanBlock / kScanThreads;

// The two 16-bit counts of a 32-bit word, added.
__device__ __forceinline__ uint32_t count_pair_sum(uint32_t w) { return (w & 0xFFFFu) + (w >> 16); }

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(Plan p) {
  __shared__ uint32_t s_scan[kScanThreads / kWave + 1];
  const uint32_t len = p.ctr->flat_len;
  const uint32_t blk = blockIdx.x;
  const uint64_t b0 = (uint64_t)blk * kScanBlock;
  if (b0 >= len) return;
  static_assert(kScanItems % 8 == 0, "whole 16-byte groups of counts per thread");
  uint32_t sum = 0;
  if (b0 + kScanBlock <= len) {  // 16-byte loads of eight counts
    const uint4* __restrict__ g = reinterpret_cast<const uint4*>(p.cnt + b0);
#pragma unroll
    for (int r = 0; r < kScanItems / 8; ++r) {
      const uint4 v = g[r * kScanThreads + threadIdx.x];
      sum += count_pair_sum(v.x) + count_pair_sum(v.y) + count_pair_sum(v.z) +
             count_pair_sum(v.w);
    }
  } else {
    for (int j = 0; j < kScanItems; ++j) {
      const uint64_t i = b0 + j * kScanThreads + threadIdx.x;
      if (i < len) sum += p.cnt[i];
    }
  }
  uint32_t total;
  (void)block_excl_scan<kScanThreads>(sum, s_scan, &total);
  if (threadIdx.x == 0) p.flat_blk[blk] = total;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(Plan p) {
  __shared__ uint32_t s_scan[1024 / kWave + 1];
  const uint32_t len = p.ctr->flat_len;
  const uint32_t nblk = (len + kScanBlock - 1) / kScanBlock;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nblk; b0 += 1024) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t v = b < nblk ? p.flat_blk[b] : 0u;
    uint32_t total;
    const uint32_t pre = block_excl_scan<1024>(v, s_scan, &total);
    if (b < nblk) p.flat_blk[b] = carry + pre;
    carry += total;
  }
  if (threadIdx.x == 0) p.flat[len] = carry;
}

// SELF_PREFIX: flat_blk holds the blocks' raw sums (no k_scan_blocks ran):
// each block adds up those before it itself (at most kScanSelfMax of them,
// a few reads per thread) and the block holding the end writes the total.
template <bool SELF_PREFIX>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(Plan p) {
  static_assert(kScanItems % 4 == 0, "whole 16-byte groups per thread");
  constexpr int Q = kScanItems / 4;  // 16-byte groups per thread
  __shared__ uint32_t s_scan[kScanThreads / kWave + 1];
  __shared__ __attribute__((aligned(16))) uint32_t s_v[kScanBlock];
  __shared__ __attribute__((aligned(16))) uint16_t s_c[kScanBlock];
  const uint32_t len = p.ctr->flat_len;
  const uint32_t blk = blockIdx.x;
  const uint64_t b0 = (uint64_t)blk * kScanBlock;
  if (SELF_PREFIX && len == 0 && blk == 0 && threadIdx.x == 0) p.flat[0] = 0;
  if (b0 >= len) return;
  uint32_t before = 0;
  if (SELF_PREFIX) {
    uint32_t part = 0;
    for (uint32_t b = threadIdx.x; b < blk; b += kScanThreads) part += p.flat_blk[b];
    (void)block_excl_scan<kScanThreads>(part, s_scan, &before);
    __syncthreads();  // s_scan is reused below
  }
  // The block's counts come in and its scan goes out through LDS so that
  // global loads and stores are 16-byte and coalesced while each thread
  // scans consecutive entries.
  const bool whole = b0 + kScanBlock <= len;
  uint4* const v4 = reinterpret_cast<uint4*>(s_v);
  uint4* const c4 = reinterpret_cast<uint4*>(s_c);
  constexpr int QC = kScanItems / 8;  // 16-byte groups of counts per thread
  if (whole) {
    const uint4* __restrict__ g = reinterpret_cast<const uint4*>(p.cnt + b0);
#pragma unroll
    for (int r = 0; r < QC; ++r) c4[r * kScanThreads + threadIdx.x] = g[r * kScanThreads + threadIdx.x];
  } else {
    for (uint32_t i = threadIdx.x; i < (uint32_t)kScanBlock; i += kScanThreads)
      s_c[i] = b0 + i < len ? p.cnt[b0 + i] : (uint16_t)0;
  }
  __syncthreads();
  // Thread t owns entries [t * kScanItems, (t + 1) * kScanItems).
  uint32_t w[4 * QC];  // two counts each
  uint32_t sum = 0;
#pragma unroll
  for (int r = 0; r < QC; ++r) {
    const uint4 c = c4[threadIdx.x * QC + r];
    w[4 * r] = c.x;
    w[4 * r + 1] = c.y;
    w[4 * r + 2] = c.z;
    w[4 * r + 3] = c.w;
    sum += count_pair_sum(c.x) + count_pair_sum(c.y) + count_pair_sum(c.z) + count_pair_sum(c.w);
  }
  uint32_t total;
  uint32_t run = block_excl_scan<kScanThreads>(sum, s_scan, &total) +
                 (SELF_PREFIX ? before : p.flat_blk[blk]);
  if (SELF_PREFIX && b0 + kScanBlock >= len && threadIdx.x == 0) p.flat[len] = before + total;
  static_assert(Q == 2 * QC, "four scanned entries per two count words");
#pragma unroll
  for (int r = 0; r < Q; ++r) {
    uint4 o;
    o.x = run;
    o.y = o.x + (w[2 * r] & 0xFFFFu);
    o.z = o.y + (w[2 * r] >> 16);
    o.w = o.z + (w[2 * r + 1] & 0xFFFFu);
    run = o.w + (w[2 * r + 1] >> 16);
    v4[threadIdx.x * Q + r] = o;
  }
  __syncthreads();
  if (whole) {
    uint4* __restrict__ g4 = reinterpret_cast<uint4*>(p.flat + b0);
#pragma unroll
    for (int r = 0; r < Q; ++r) g4[r * kScanThreads + threadIdx.x] = v4[r * kScanThreads + threadIdx.x];
  } else {
    for (uint32_t i = threadIdx.x; i < (uint32_t)kScanBlock; i += kScanThreads)
      if (b0 + i < len) p.flat[b0 + i] = s_v[i];
  }
}

// ---------------------------------------------------------------------------
// Global pass, step 5: per-segment buckets and child classification.

__device__ __forceinline__ void push_local(Plan& p, uint32_t begin, uint32_t size,
                                           uint32_t depth, uint32_t buf) {
  const int c = size <= kLocalClassCap[0] ? 0 : (size <= kLocalClassCap[1] ? 1 : 2);
  const uint32_t slot = (uint32_t)atomic_add_pair(&p.ctr->local[c].segs, 1u, size);
  if (slot >= p.cap_local[c]) {
    atomicOr(&p.ctr->error, 2u);
    return;
  }
  p.local[c][slot] = LocalSeg{begin, size, depth, buf};
}

__device__ __forceinline__ void push_child(Plan& p, BigSeg* next, uint32_t begin,
                                           uint32_t size, uint32_t depth, uint32_t buf) {
  if (size <= p.local_threshold) {
    push_local(p, begin, size, depth, buf);
    return;
  }
  const uint32_t slot = (uint32_t)atomic_add_pair(&p.ctr->big_next, 1u, size);
  if (slot >= p.cap_big) {
    atomicOr(&p.ctr->error, 4u);
    return;
  }
  next[slot] = BigSeg{begin, size, depth, buf};
}

__device__ __forceinline__ void push_retire(Plan& p, uint32_t begin, uint32_t size,
                                            uint32_t buf, uint32_t lcp, uint32_t copy) {
  const uint32_t nchunks = (size + kRetireChunk - 1) / kRetireChunk;
  const uint32_t slot = atomicAdd(&p.ctr->retire, nchunks);
  if (slot + nchunks > p.cap_retire) {
    atomicOr(&p.ctr->error, 8u);
    return;
  }
  for (uint32_t c = 0; c < nchunks; ++c) {
    const uint32_t b = begin + c * kRetireChunk;
    const uint32_t l = min(kRetireChunk, size - c * kRetireChunk);
    p.retire[slot + c] = Retire{b, l, begin + size - 1, buf, lcp, copy};
  }
}

template <int RB, int BS>
__device__ __forceinline__ void segscan_seg(Plan& p, const BigSeg* segs, BigSeg* next,
                                            uint32_t si) {
  constexpr uint32_t NB = 1u << RB;
  constexpr uint32_t DPT = NB / BS;
  static_assert(DPT >= 1, "at least one digit per thread");
  __shared__ uint32_t s_first;
  // The segment's count-matrix place and the alphabet are loaded with its
  // bounds, before its min/max can end the block.
  const SegAux ax = p.aux[si];
  uint32_t alpha[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) alpha[w] = p.alpha[w];
  const BigSeg sg = segs[si];
  const uint64_t mn = p.segmm[2 * si], mx = p.segmm[2 * si + 1];
  if (mn == mx) {
    if (threadIdx.x == 0) {
      if ((mn & 0xFFu) == 0)  // the key holds the NUL: all strings identical
        push_retire(p, sg.begin, sg.size, sg.buf, sg.depth + key_len(mn), 1u);
      else
        push_child(p, next, sg.begin, sg.size, sg.depth + 8, sg.buf);
    }
    return;
  }
  const uint32_t nbuf = sg.buf == 1 ? 2u : 1u;
  const bool has_nul = alpha[0] & 1u;  // NUL, if present, has rank 0
  const DigitSpec ds = digit_spec<RB>(mn, mx, alpha_size(alpha));
  const uint32_t fb = ds.fb0 + ds.k;  // bytes every child shares in full
  const uint32_t* col0 = p.flat + ax.flat_base;
  // Bucket of digit d: [start(d), start(d+1)), start(nbu) = segment size.
  // The DPT + 1 starts a thread needs are loaded before any is used (clamped
  // to a valid entry, so the strided loads overlap).
  uint32_t sv[DPT + 1];
  const uint32_t base0 = col0[0];
#pragma unroll
  for (uint32_t q = 0; q <= DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    sv[q] = col0[(uint64_t)(d < ds.nbu ? d : ds.nbu - 1) * ax.ntiles];
  }
  uint32_t st[DPT], cnt[DPT];
  uint32_t firstd = NB;
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    st[q] = 0;
    cnt[q] = 0;
    if (d < ds.nbu) {
      st[q] = sv[q] - base0;
      cnt[q] = (d + 1 < ds.nbu ? sv[q + 1] - base0 : sg.size) - st[q];
      if (cnt[q] && firstd == NB) firstd = d;
    }
  }
  if (threadIdx.x == 0) s_first = NB;
  __syncthreads();
  if (firstd < NB) atomicMin(&s_first, firstd);
  __syncthreads();
  const uint32_t dfirst = s_first;
  uint32_t* bs = p.bstart + (uint64_t)si * NB;

  // Classify every bucket.  kind: 0 final or empty, 1..3 local class 0..2,
  // 4 next pass, 5 run of identical strings (nul = where they end).
  uint32_t kind[DPT], nulp[DPT];
  uint32_t nb = 0, nr = 0, ng = 0, sg_next = 0, nl[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    const uint32_t c = cnt[q];
    kind[q] = 0;
    nulp[q] = 0;
    if (d >= ds.nbu || c == 0) continue;
    // The bucket's strings ended if one of its digit characters is NUL (the
    // common bytes before fb0 hold none: a NUL there would make min == max).
    bool ended = false;
    if (has_nul) {
      uint32_t rest = d, scale = ds.nbu;
      for (uint32_t ch = 0; ch < ds.k; ++ch) {
        scale /= ds.A;
        if (rest / scale == 0) {
          ended = true;
          nulp[q] = ds.fb0 + ch;
          break;
        }
        rest %= scale;
      }
    }
    const bool fin = ended || c == 1;
    bs[d] = fin ? kBucketFinal : (c <= p.local_threshold ? kBucketLocal : kBucketNext);
    if (d != dfirst) ++nb;
    if (ended) {
      if (c > 1) {
        kind[q] = 5;
        nr += (c + kRetireChunk - 1) / kRetireChunk;
      }
    } else if (!fin) {
      if (c <= p.local_threshold) {
        const uint32_t cls = c <= kLocalClassCap[0] ? 0 : (c <= kLocalClassCap[1] ? 1 : 2);
        kind[q] = 1 + cls;
        ++nl[cls];
        sl[cls] += c;
      } else {
        kind[q] = 4;
        ++ng;
        sg_next += c;
      }
    }
  }

  // One reservation per list for the whole segment, the lists reserved in
  // pairs (boundaries with retire chunks, next-pass segments with their
  // strings, each class's segments with theirs): same-address atomics from
  // every bucket, or every segment, would serialise.
  __shared__ uint32_t s_scan[10 * (BS / kWave + 1)];
  __shared__ uint32_t s_base[6];
  // Ten counts per thread, scanned together: boundaries, retire chunks,
  // next-pass segments, local segments and their strings per class, and the
  // next-pass segments' strings.
  uint32_t sc[10] = {nb, nr, ng, nl[0], nl[1], nl[2], sl[0], sl[1], sl[2], sg_next};
  uint32_t tot[10];
  block_excl_scan_multi<BS, 10>(sc, s_scan, tot);
  uint32_t ob = sc[0], orr = sc[1], og = sc[2];
  uint32_t ol[3] = {sc[3], sc[4], sc[5]};
  const uint32_t tb = tot[0], tr = tot[1], tg = tot[2], t0 = tot[3], t1 = tot[4], t2 = tot[5];
  const uint32_t ts0 = tot[6], ts1 = tot[7], ts2 = tot[8], tgs = tot[9];
  // The five reservations are independent: lanes 0-4 of the first wave
  // make them with one atomic instruction (one trip to L2, not five).
  if (threadIdx.x < 5) {
    const uint32_t r = threadIdx.x;  // 0 boundaries + retire chunks, 1 next pass, 2-4 local class
    uint32_t* const at = r == 0 ? &p.ctr->boundary
                                : (r == 1 ? &p.ctr->big_next : &p.ctr->local[r - 2].segs);
    const uint32_t a = r == 0 ? tb : (r == 1 ? tg : (r == 2 ? t0 : (r == 3 ? t1 : t2)));
    const uint32_t b = r == 0 ? tr : (r == 1 ? tgs : (r == 2 ? ts0 : (r == 3 ? ts1 : ts2)));
    const uint64_t old = (a || b) ? atomic_add_pair(at, a, b) : 0ull;
    const uint32_t lo = (uint32_t)old, hi = (uint32_t)(old >> 32);
    const uint32_t cap = r == 0 ? p.cap_bounds : (r == 1 ? p.cap_big : p.cap_local[r - 2]);
    uint32_t err = lo + a > cap ? (r == 0 ? 16u : (r == 1 ? 4u : 2u)) : 0u;
    if (r == 0 && hi + b > p.cap_retire) err |= 8u;
    s_base[r == 0 ? 0 : (r == 1 ? 5 : r)] = lo;
    if (r == 0) s_base[1] = hi;
    if (err) atomicOr(&p.ctr->error, err);
  }
  __syncthreads();
  ob += s_base[0];
  orr += s_base[1];
  og += s_base[5];
  for (int c = 0; c < 3; ++c) ol[c] += s_base[2 + c];
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    const uint32_t c = cnt[q];
    if (d >= ds.nbu || c == 0) continue;
    const uint32_t cb = sg.begin + st[q];
    if (d != dfirst) {
      if (ob < p.cap_bounds) p.bounds[ob] = Boundary{cb - 1, sg.depth, nbuf, 0u};
      ++ob;
    }
    const uint32_t kd = kind[q];
    if (kd == 5) {
      const uint32_t nch = (c + kRetireChunk - 1) / kRetireChunk;
      for (uint32_t ch = 0; ch < nch; ++ch, ++orr) {
        if (orr >= p.cap_retire) continue;
        const uint32_t b = cb + ch * kRetireChunk;
        p.retire[orr] = Retire{b, min(kRetireChunk, c - ch * kRetireChunk), cb + c - 1, nbuf,
                               sg.depth + nulp[q], 0u};
      }
    } else if (kd == 4) {
      if (og < p.cap_big) next[og] = BigSeg{cb, c, sg.depth + fb, nbuf};
      ++og;
    } else if (kd >= 1) {
      const uint32_t slot = kd == 1 ? ol[0]++ : (kd == 2 ? ol[1]++ : ol[2]++);
      if (slot < p.cap_local[kd - 1])
        p.local[kd - 1][slot] = LocalSeg{cb, c, sg.depth + fb, nbuf};
    }
  }
}

// One workgroup per segment.  A pass over one segment (the first pass of a
// single sort) takes 1024 threads: its 2^RB buckets are then 2 per thread
// instead of 8, and the block's latency is the pass's.
template <int RB, int BS>
__global__ __launch_bounds__(BS) void k_segscan(Plan p, const BigSeg* segs, BigSeg* next) {
  if (blockIdx.x < p.ctr->big_cur) segscan_seg<RB, BS>(p, segs, next, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Global pass, step 6: stable scatter.  The in-tile rank of each string
// among the tile's strings of its digit comes from wave match-ranking; the
// scanned flat array gives where the tile's strings of each digit start.

template <int RB, int TI, bool CARRY>
__device__ __forceinline__ void scatter_tile(Plan& p, const BigSeg* segs, uint32_t b,
                                             uint32_t nt) {
  constexpr uint32_t NB = 1u << RB;
  constexpr int W = kTileThreads / kWave;
  constexpr int IPW = kTileThreads * TI / W;  // items per wave
  constexpr uint32_t DPT = NB >= kTileThreads ? NB / kTileThreads : 1;
  __shared__ __attribute__((aligned(16))) uint16_t s_whist[NB][W];  // digit-major
  __shared__ uint32_t s_tstart[NB];
  __shared__ uint32_t s_gbase[NB];
  __shared__ uint8_t s_kind[NB];  // kBucket* kind of each digit's bucket
  // Once every item has its place, the wave counters' space is free: a pass
  // that hands on keys stages them there (a tile's keys fill it at 11-bit
  // digits); a carried pass stages its offsets and digits there instead,
  // which leaves the block about 52 KB of LDS and three blocks per CU.
  constexpr uint32_t kItems = kTileThreads * TI;
  constexpr bool kStageKeys = CARRY && NB * W * 2 >= 8 * kItems;
  constexpr bool kAliasOff = !CARRY && NB * W * 2 >= 6 * kItems;
  __shared__ uint32_t s_off_own[kAliasOff ? 1 : kItems];
  __shared__ uint16_t s_dig_own[kAliasOff ? 1 : kItems];
  __shared__ uint32_t s_scan[W + 1];
  __shared__ uint8_t rank[256];

  const Tile t = p.tiles[xcd_tile(b, nt)];
  constexpr bool carry = CARRY;  // hand the next pass its keys (a text-reading pass)
  char* const whist_bytes = reinterpret_cast<char*>(&s_whist[0][0]);
  uint32_t* const s_off = kAliasOff ? reinterpret_cast<uint32_t*>(whist_bytes) : s_off_own;
  uint16_t* const s_dig =
      kAliasOff ? reinterpret_cast<uint16_t*>(whist_bytes + 4 * kItems) : s_dig_own;
  const int wid = threadIdx.x / kWave;
  const int lane = lane_id();
  // Everything that needs only the tile is loaded before the segment's
  // min/max can end the block (a load after that return cannot be hoisted
  // above it): every item's key and next 8 bytes (slots past the tile read
  // its last string and are masked out), the alphabet and the segment's
  // count-matrix place.  The offsets follow once the segment names the
  // buffer that holds them.
  uint64_t key[TI];
  uint64_t nk[TI];  // key of the string at its child's depth (carry)
#pragma unroll
  for (int j = 0; j < TI; ++j) {
    const uint32_t i = wid * IPW + j * kWave + lane;
    const uint32_t at = t.begin + (i < t.len ? i : t.len - 1);
    key[j] = p.keys[at];
    nk[j] = carry ? p.ext[at] : 0ull;
  }
  uint32_t alpha[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) alpha[w] = p.alpha[w];
  const SegAux ax = p.aux[t.seg];
  const BigSeg sg = segs[t.seg];
  const uint64_t mn = p.segmm[2 * t.seg], mx = p.segmm[2 * t.seg + 1];
  if (mn == mx) {
    // The whole segment stays where it is and, unless its strings ended,
    // goes 8 bytes deeper: its next keys are the 8 bytes after these.
    if (carry && (mn & 0xFFu) != 0)
      for (uint32_t i = threadIdx.x; i < t.len; i += kTileThreads)
        p.keys_next[t.begin + i] = p.ext[t.begin + i];
    return;
  }
  const DigitSpec ds = digit_spec<RB>(mn, mx, build_rank(alpha, rank));
  const uint32_t nbits = digit_bits(ds.nbu);
  const uint32_t fb = ds.fb0 + ds.k;  // bytes the children share: their keys start there
  const uint32_t* __restrict__ src = p.idx[sg.buf];
  uint32_t* __restrict__ dst = p.idx[sg.buf == 1 ? 2 : 1];
  uint32_t off[TI];
#pragma unroll
  for (int j = 0; j < TI; ++j) {
    const uint32_t i = wid * IPW + j * kWave + lane;
    off[j] = src[t.begin + (i < t.len ? i : t.len - 1)];
  }

  static_assert(W == 8, "one 16-byte row of wave counters per digit");
  uint4* const rows = reinterpret_cast<uint4*>(&s_whist[0][0]);
  for (uint32_t d = threadIdx.x; d < NB; d += kTileThreads) rows[d] = make_uint4(0, 0, 0, 0);
  __syncthreads();

  const uint64_t lt = lanemask_lt();
  uint32_t dig[TI], rnk[TI];
#pragma unroll
  for (int j = 0; j < TI; ++j) {
    dig[j] = digit_of(key[j], ds, rank);
    if (carry) nk[j] = fb >= 8 ? nk[j] : (key[j] << (8 * fb)) | (nk[j] >> (64 - 8 * fb));
  }
#pragma unroll
  for (int j = 0; j < TI; ++j) {
    const uint32_t i = wid * IPW + j * kWave + lane;
    const bool v = i < t.len;
    const uint32_t d = v ? dig[j] : 0u;
    dig[j] = d;
    const uint64_t m = match_digit_n(d, __ballot(v), nbits);
    rnk[j] = 0;
    if (v) {
      const uint32_t base = s_whist[d][wid];
      rnk[j] = base + (uint32_t)__popcll(m & lt);
      if (lane == __builtin_ctzll(m)) s_whist[d][wid] = (uint16_t)(base + __popcll(m));
    }
  }
  __syncthreads();

  // This tile's place in each digit's bucket and the bucket's kind, loaded
  // before the scan below so that the strided reads overlap it.
  const uint32_t* col = p.flat + ax.flat_base;
  const uint32_t* bs = p.bstart + (uint64_t)t.seg * NB;
  const uint32_t base0 = col[0];
  uint32_t gcol[DPT], gkind[DPT];
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    const uint32_t dc = d < ds.nbu ? d : ds.nbu - 1;
    gcol[q] = col[(uint64_t)dc * ax.ntiles + t.tidx];
    gkind[q] = bs[dc];
  }

  // Per digit: exclusive prefix over waves (in place) and tile total; a
  // digit's eight wave counters are one 16-byte row of LDS.
  uint32_t tot[DPT];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    uint32_t r = 0;
    if (d < NB) {
      const uint4 c = rows[d];
      const uint32_t w[4] = {c.x, c.y, c.z, c.w};  // two 16-bit counters each
      uint32_t out[4];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const uint32_t lo = w[h] & 0xFFFFu, hi = w[h] >> 16;
        out[h] = r | ((r + lo) << 16);
        r += lo + hi;
      }
      rows[d] = make_uint4(out[0], out[1], out[2], out[3]);
    }
    tot[q] = r;
    sum += r;
  }
  uint32_t tsum;
  uint32_t run = block_excl_scan<kTileThreads>(sum, s_scan, &tsum);
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    if (d < ds.nbu) {
      s_tstart[d] = run;
      s_gbase[d] = sg.begin + (gcol[q] - base0);
      s_kind[d] = (uint8_t)gkind[q];
    }
    run += tot[q];
  }
  __syncthreads();
  // The next pass's keys follow the offsets through LDS, in the space of the
  // wave counters once every item has its place, and leave in bucket order:
  // runs of consecutive positions instead of one scattered 8-byte store per
  // string.
  uint64_t* const s_key = reinterpret_cast<uint64_t*>(whist_bytes);
  if (kStageKeys || kAliasOff) {
    uint32_t lpos[TI];
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const uint32_t d = dig[j];
      lpos[j] = s_tstart[d] + s_whist[d][wid] + rnk[j];  // the item's place in the tile
    }
    __syncthreads();  // the wave counters are read: their space is reused
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const uint32_t i = wid * IPW + j * kWave + lane;
      if (i < t.len) {
        s_off[lpos[j]] = off[j];
        s_dig[lpos[j]] = (uint16_t)dig[j];
        if (kStageKeys) s_key[lpos[j]] = nk[j];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const uint32_t i = wid * IPW + j * kWave + lane;
      if (i < t.len) {
        const uint32_t d = dig[j];
        const uint32_t within = s_whist[d][wid] + rnk[j];  // place among the tile's digit d
        const uint32_t lp = s_tstart[d] + within;
        s_off[lp] = off[j];
        s_dig[lp] = (uint16_t)d;
        // Keys for the next pass go straight to the strings' new positions
        // (only for buckets that take one).
        const uint32_t pos = s_gbase[d] + within;
        if (carry && s_kind[d] == kBucketNext && pos - sg.begin < sg.size) p.keys_next[pos] = nk[j];
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < t.len; i += kTileThreads) {
    const uint32_t d = s_dig[i];
    const uint32_t pos = s_gbase[d] + (i - s_tstart[d]);
    const uint32_t o = s_off[i];
    if (pos - sg.begin >= sg.size) {  // counts disagree: never write outside
      atomicOr(&p.ctr->error, 64u);
      continue;
    }
    dst[pos] = o;
    const uint32_t kind = s_kind[d];
    if (kind == kBucketFinal) p.out[pos] = o;
    if (kStageKeys && kind == kBucketNext) p.keys_next[pos] = s_key[i];
  }
}

template <int RB, int TI, bool CARRY>
__global__ __launch_bounds__(kTileThreads) void k_scatter(Plan p, const BigSeg* segs) {
  const uint32_t nt = p.ctr->ntiles;
  if (blockIdx.x < nt) scatter_tile<RB, TI, CARRY>(p, segs, blockIdx.x, nt);
}

// ---------------------------------------------------------------------------
// Global pass, step 7: LCP of pairs split by this pass.

__global__ void k_boundary_lcp(Plan p) {
  const uint32_t first = p.ctr->bound_done;
  const uint32_t n = min(p.ctr->boundary, p.cap_bounds);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const Boundary b = p.bounds[i];
    const uint32_t* ix = p.idx[b.buf];
    const uint64_t k1 = load_key(p.text64, p.text_bytes, (uint64_t)ix[b.pos] + b.depth);
    const uint64_t k2 = load_key(p.text64, p.text_bytes, (uint64_t)ix[b.pos + 1] + b.depth);
    p.lcp[b.pos] = (uint64_t)b.depth + key_lcp(k1, k2);
  }
}

// ---------------------------------------------------------------------------
// Runs of identical strings: copy offsets (when needed) and fill their LCPs.

__global__ void k_retire(Plan p) {
  const uint32_t count = min(p.ctr->retire, p.cap_retire);
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const Retire r = p.retire[e];
    const uint32_t* ix = p.idx[r.buf];
    for (uint32_t i = r.begin + threadIdx.x; i < r.begin + r.len; i += blockDim.x) {
      if (r.copy) p.out[i] = ix[i];
      if (p.lcp && i < r.last) p.lcp[i] = r.lcp;
    }
  }
}

// ---------------------------------------------------------------------------
// Launch wrappers

static inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 4096) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (uint32_t)g;
}

hipError_t launch_plan(const Plan& p, const BigSeg* segs, uint32_t pass, uint32_t grid_tiles,
                       uint32_t grid_segs, hipStream_t s) {
  k_plan<<<1, kPlanThreads, 0, s>>>(p, segs, pass, grid_tiles, grid_segs);
  return hipGetLastError();
}

hipError_t launch_segreduce(const Plan& p, uint32_t grid, hipStream_t s) {
  if (!grid) return hipSuccess;
  k_segreduce<<<grid, kSegReduceThreads, 0, s>>>(p);
  return hipGetLastError();
}

hipError_t launch_gather_minmax(const Plan& p, const BigSeg* segs, uint32_t grid,
                                uint32_t radix_bits, hipStream_t s) {
  if (!grid) return hipSuccess;
  if (p.tile != 2048 && p.tile != 4096) return hipErrorInvalidValue;
  const int v = (radix_bits == 11 ? 2 : (radix_bits == 8 ? 0 : 4)) + (p.tile == 4096 ? 1 : 0);
  switch (v) {
    case 0: k_gather_minmax<8, 4><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 1: k_gather_minmax<8, 8><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 2: k_gather_minmax<11, 4><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 3: k_gather_minmax<11, 8><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_layout(const Plan& p, uint32_t radix_bits, hipStream_t s) {
  if (radix_bits == 8)
    k_layout<8><<<1, 1024, 0, s>>>(p);
  else if (radix_bits == 11)
    k_layout<11><<<1, 1024, 0, s>>>(p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_histogram(const Plan& p, uint32_t grid, uint32_t radix_bits, hipStream_t s) {
  if (!grid) return hipSuccess;
  const int v = (radix_bits == 11 ? 2 : (radix_bits == 8 ? 0 : 4)) + (p.tile == 4096 ? 1 : 0);
  if (p.tile != 2048 && p.tile != 4096) return hipErrorInvalidValue;
  switch (v) {
    case 0: k_histogram<8, 4><<<grid, kTileThreads, 0, s>>>(p); break;
    case 1: k_histogram<8, 8><<<grid, kTileThreads, 0, s>>>(p); break;
    case 2: k_histogram<11, 4><<<grid, kTileThreads, 0, s>>>(p); break;
    case 3: k_histogram<11, 8><<<grid, kTileThreads, 0, s>>>(p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_flat_scan(const Plan& p, uint64_t max_len, hipStream_t s) {
  const uint32_t blocks = (uint32_t)((max_len + kScanBlock - 1) / kScanBlock);
  // Up to kScanSelfMax blocks, k_scan_down adds up the block sums before
  // each block itself, one launch fewer than with k_scan_blocks.
  constexpr uint32_t kScanSelfMax = 2048;
  if (blocks <= kScanSelfMax) {
    if (blocks) k_scan_reduce<<<blocks, kScanThreads, 0, s>>>(p);
    k_scan_down<true><<<blocks ? blocks : 1, kScanThreads, 0, s>>>(p);
    return hipGetLastError();
  }
  k_scan_reduce<<<blocks, kScanThreads, 0, s>>>(p);
  k_scan_blocks<<<1, 1024, 0, s>>>(p);
  k_scan_down<false><<<blocks, kScanThreads, 0, s>>>(p);
  return hipGetLastError();
}

hipError_t launch_segscan(const Plan& p, const BigSeg* segs, BigSeg* next, uint32_t grid,
                          uint32_t radix_bits, hipStream_t s) {
  if (!grid) return hipSuccess;
  if (radix_bits == 8) {
    if (grid == 1)
      k_segscan<8, 256><<<1, 256, 0, s>>>(p, segs, next);
    else
      k_segscan<8, 256><<<grid, 256, 0, s>>>(p, segs, next);
  } else if (radix_bits == 11) {
    if (grid == 1)
      k_segscan<11, 1024><<<1, 1024, 0, s>>>(p, segs, next);
    else
      k_segscan<11, 256><<<grid, 256, 0, s>>>(p, segs, next);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_scatter(const Plan& p, const BigSeg* segs, uint32_t grid,
                          uint32_t radix_bits, hipStream_t s) {
  if (!grid) return hipSuccess;
  const int v = (radix_bits == 11 ? 2 : (radix_bits == 8 ? 0 : 4)) + (p.tile == 4096 ? 1 : 0);
  if (p.tile != 2048 && p.tile != 4096) return hipErrorInvalidValue;
  const bool carry = !p.keys_carried;  // a text-reading pass hands on keys
  switch (v + (carry ? 8 : 0)) {
    case 0: k_scatter<8, 4, false><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 1: k_scatter<8, 8, false><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 2: k_scatter<11, 4, false><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 3: k_scatter<11, 8, false><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 8: k_scatter<8, 4, true><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 9: k_scatter<8, 8, true><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 10: k_scatter<11, 4, true><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    case 11: k_scatter<11, 8, true><<<grid, kTileThreads, 0, s>>>(p, segs); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_boundary_lcp(const Plan& p, hipStream_t s) {
  k_boundary_lcp<<<1024, 256, 0, s>>>(p);
  return hipGetLastError();
}

hipError_t launch_retire(const Plan& p, uint32_t grid, hipStream_t s) {
  if (!grid) return hipSuccess;
  k_retire<<<grid, 256, 0, s>>>(p);
  return hipGetLastError();
}

hipError_t launch_off_to_index(const uint32_t* off, const uint32_t* starts, uint32_t* idx,
                              uint64_t n, hipStream_t s) {
  k_off_to_index<<<grid_for(n, 256), 256, 0, s>>>(off, starts, idx, n);
  return hipGetLastError();
}

int add_upload(UploadBatch& b, const void* h_src, void* d_dst, size_t zero_bytes,
               size_t bytes) {
  if (bytes % 8 || zero_bytes % 8 || zero_bytes > bytes || bytes / 8 > 0xFFFFFFFFull ||
      b.count >= (uint32_t)kUploadRanges)
    return -1;
  if (!bytes) return 0;
  b.r[b.count++] = UploadRange{static_cast<const uint64_t*>(h_src), static_cast<uint64_t*>(d_dst),
                               (uint32_t)(zero_bytes / 8), (uint32_t)(bytes / 8)};
  return 0;
}

hipError_t launch_upload(UploadBatch& b, hipStream_t s) {
  uint32_t most = 0;
  for (uint32_t i = 0; i < b.count; ++i) most = b.r[i].nwords > most ? b.r[i].nwords : most;
  const uint32_t count = b.count;
  b.count = 0;
  if (!most) return hipSuccess;
  const uint32_t blocks = most >= 256u * 256u ? 256u : (most + 255) / 256;
  k_upload<<<dim3(blocks, count), 256, 0, s>>>(b);
  return hipGetLastError();
}

hipError_t launch_publish(const void* a, size_t a_bytes, const void* b, const uint32_t* count,
                          uint32_t cap, size_t rec_bytes, void* h_dst, uint64_t* h_flag,
                          uint64_t seq, hipStream_t s) {
  if (a_bytes % 8 || rec_bytes % 8 || a_bytes / 8 + (uint64_t)cap * (rec_bytes / 8) > 0xFFFFFFFFull)
    return hipErrorInvalidValue;
  k_publish<<<1, kPublishThreads, 0, s>>>(static_cast<const uint64_t*>(a), (uint32_t)(a_bytes / 8),
                               static_cast<const uint64_t*>(b), count, cap,
                               (uint32_t)(rec_bytes / 8), static_cast<uint64_t*>(h_dst), h_flag,
                               seq);
  return hipGetLastError();
}

hipError_t launch_copy_chunks(const CopyChunk* chunks, uint32_t nchunks, hipStream_t s) {
  if (!nchunks) return hipSuccess;
  k_copy_chunks<<<nchunks, 256, 0, s>>>(chunks);
  return hipGetLastError();
}

}  // namespace gss
