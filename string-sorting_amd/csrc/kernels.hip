// HIP kernels of the MI355X MSD radix string sort (gfx950, wave64).
//
// A sort runs as global MSD passes over segments too large for one
// workgroup, then finishes every remaining segment inside one workgroup.
//
// Global pass over the BigSeg list (one launch per step, tiles of kTile
// strings, never spanning two segments):
//   gather_minmax  key = the 8 bytes at each string's depth (big-endian, NUL
//                  padded), cached in `keys`; per-segment min/max key.
//   histogram      digit = the radix_bits bits that follow the common prefix
//                  of the segment's min and max key; per-segment histogram.
//   segscan        one workgroup per segment: bucket starts, child segments
//                  (next pass, local, or final), retire runs, boundaries.
//   scatter        stable partition of (offset) by digit: wave match-ranking
//                  in LDS + decoupled look-back across the segment's tiles,
//                  staged in LDS so the global writes are contiguous runs.
//   boundary_lcp   LCP of each new adjacent pair of child buckets.
// A segment whose keys are all equal costs the gather only: it keeps its
// place and buffer and moves 8 bytes deeper (or retires if the key holds the
// NUL, i.e. all its strings are identical).
//
// local<BLOCK,CAP>: one workgroup per segment of <= CAP strings, rounds of
// (gather 8-byte keys, stable LSD radix sort of (sub-segment, key) in LDS on
// the bytes that vary inside some sub-segment, split into equal-key groups,
// retire singletons and NUL-terminated groups, emit LCPs) until done.
//
// The reference's byte-at-a-time MSD (msd_CE2, src/msd_ce.cpp:143-178) and
// the 8-byte cache rules of multikey_cache (src/multikey_cache.cpp:107-125,
// 207-211) define the order and the retirement rule used here.

#include "gss_internal.h"

namespace gss {

// ---------------------------------------------------------------------------
// Small helpers

__global__ void k_init_segs(uint64_t* segmm, uint32_t* seghist, uint32_t nsegs,
                            uint32_t nbins) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t i = t0; i < 2ull * nsegs; i += stride) segmm[i] = (i & 1) ? 0ull : ~0ull;
  for (uint64_t i = t0; i < (uint64_t)nsegs * nbins; i += stride) seghist[i] = 0;
}

__global__ void k_fill_u64(uint64_t* dst, uint64_t n, uint64_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = v;
}

__global__ void k_zero_counters(Counters* c) {
  uint32_t* w = reinterpret_cast<uint32_t*>(c);
  for (uint32_t i = threadIdx.x; i < sizeof(Counters) / 4; i += blockDim.x) w[i] = 0;
}

__global__ void k_ptr_to_off(const uint64_t* __restrict__ ptrs, uint64_t base,
                             uint32_t* __restrict__ off, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    off[i] = (uint32_t)(ptrs[i] - base);
}

__global__ void k_off_to_ptr(const uint32_t* __restrict__ off, uint64_t base,
                             uint64_t* __restrict__ ptrs, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    ptrs[i] = base + off[i];
}

// Packed texts: offsets were assigned in input order, so the input index of
// a sorted offset is found by binary search in the ascending start array.
__global__ void k_off_to_orig_ptr(const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ starts,
                                  const uint64_t* __restrict__ orig,
                                  uint64_t* __restrict__ ptrs, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t o = off[i];
    uint64_t lo = 0, hi = n;  // first index with starts[idx] >= o
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (starts[mid] < o) lo = mid + 1; else hi = mid;
    }
    ptrs[i] = orig[lo];
  }
}

// Digit of `key` for a segment whose keys share their first `lzb` bits.
template <int RB>
__device__ __forceinline__ uint32_t seg_digit(uint64_t key, uint32_t lzb) {
  return (uint32_t)((key << lzb) >> (64 - RB));
}

// ---------------------------------------------------------------------------
// Global pass, step 1: gather keys, per-segment min/max.

__global__ __launch_bounds__(kTileThreads) void k_gather_minmax(Plan p, const BigSeg* segs) {
  __shared__ uint64_t s_mn[kTileThreads / kWave], s_mx[kTileThreads / kWave];
  const Tile t = p.tiles[blockIdx.x];
  const BigSeg sg = segs[t.seg];
  const uint32_t* __restrict__ src = p.idx[sg.buf];
  uint32_t o[kTileItems];
  uint64_t k[kTileItems];
#pragma unroll
  for (int j = 0; j < kTileItems; ++j) {
    const uint32_t i = j * kTileThreads + threadIdx.x;
    o[j] = i < t.len ? src[t.begin + i] : 0u;
  }
  uint64_t mn = ~0ull, mx = 0ull;
#pragma unroll
  for (int j = 0; j < kTileItems; ++j) {
    const uint32_t i = j * kTileThreads + threadIdx.x;
    if (i < t.len) {
      k[j] = load_key(p.text64, p.nwords, (uint64_t)o[j] + sg.depth);
    }
  }
#pragma unroll
  for (int j = 0; j < kTileItems; ++j) {
    const uint32_t i = j * kTileThreads + threadIdx.x;
    if (i < t.len) {
      p.keys[t.begin + i] = k[j];
      mn = k[j] < mn ? k[j] : mn;
      mx = k[j] > mx ? k[j] : mx;
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
  if (threadIdx.x == 0) {
    for (int w = 1; w < kTileThreads / kWave; ++w) {
      mn = s_mn[w] < mn ? s_mn[w] : mn;
      mx = s_mx[w] > mx ? s_mx[w] : mx;
    }
    atomicMin((unsigned long long*)&p.segmm[2 * t.seg], (unsigned long long)mn);
    atomicMax((unsigned long long*)&p.segmm[2 * t.seg + 1], (unsigned long long)mx);
  }
}

// ---------------------------------------------------------------------------
// Global pass, step 2: per-segment digit histogram.

template <int RB>
__global__ __launch_bounds__(kTileThreads) void k_histogram(Plan p) {
  constexpr uint32_t NB = 1u << RB;
  __shared__ uint32_t h[NB];
  const Tile t = p.tiles[blockIdx.x];
  const uint64_t mn = p.segmm[2 * t.seg], mx = p.segmm[2 * t.seg + 1];
  if (mn == mx) return;
  const uint32_t lzb = (uint32_t)__builtin_clzll(mn ^ mx);
  for (uint32_t d = threadIdx.x; d < NB; d += kTileThreads) h[d] = 0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTileItems; ++j) {
    const uint32_t i = j * kTileThreads + threadIdx.x;
    if (i < t.len) atomicAdd(&h[seg_digit<RB>(p.keys[t.begin + i], lzb)], 1u);
  }
  __syncthreads();
  uint32_t* gh = p.seghist + (uint64_t)t.seg * NB;
  for (uint32_t d = threadIdx.x; d < NB; d += kTileThreads)
    if (h[d]) atomicAdd(&gh[d], h[d]);
}

// ---------------------------------------------------------------------------
// Global pass, step 3: per-segment bucket starts and child classification.

__device__ __forceinline__ void push_local(Plan& p, uint32_t begin, uint32_t size,
                                           uint32_t depth, uint32_t buf) {
  const int c = size <= kLocalClassCap[0] ? 0 : (size <= kLocalClassCap[1] ? 1 : 2);
  const uint32_t slot = atomicAdd(&p.ctr->local[c], 1u);
  atomicAdd(&p.ctr->local_strings[c], size);
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
  const uint32_t slot = atomicAdd(&p.ctr->big_next, 1u);
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

template <int RB>
__global__ __launch_bounds__(256) void k_segscan(Plan p, const BigSeg* segs, BigSeg* next) {
  constexpr uint32_t NB = 1u << RB;
  constexpr uint32_t DPT = NB / 256;
  __shared__ uint32_t s_scan[256 / kWave + 1];
  __shared__ uint32_t s_first;
  const uint32_t si = blockIdx.x;
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
  const uint32_t lzb = (uint32_t)__builtin_clzll(mn ^ mx);
  const uint32_t nreal = min((uint32_t)RB, 64u - lzb);  // digit bits inside the key
  const uint32_t shared_bits = lzb + nreal;
  const uint32_t fb = shared_bits / 8;  // bytes every child shares in full
  const uint64_t hi = lzb ? (mn & (~0ull << (64 - lzb))) : 0ull;
  const uint32_t* gh = p.seghist + (uint64_t)si * NB;

  uint32_t cnt[DPT];
  uint32_t sum = 0, firstd = NB;
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    cnt[q] = gh[d];
    sum += cnt[q];
    if (cnt[q] && firstd == NB) firstd = d;
  }
  if (threadIdx.x == 0) s_first = NB;
  __syncthreads();
  if (firstd < NB) atomicMin(&s_first, firstd);
  uint32_t total;
  uint32_t run = block_excl_scan<256>(sum, s_scan, &total);
  const uint32_t dfirst = s_first;
  uint32_t* bs = p.bstart + (uint64_t)si * NB;
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    const uint32_t c = cnt[q];
    const uint32_t st = run;
    run += c;
    if (c == 0) {
      bs[d] = st;
      continue;
    }
    const uint64_t pre = hi | ((uint64_t)(d >> (RB - nreal)) << (64 - shared_bits));
    bool ended = false;
    uint32_t nul = 0;
    for (uint32_t j = 0; j < fb; ++j) {
      if (((pre >> (56 - 8 * j)) & 0xFFu) == 0) {
        ended = true;
        nul = j;
        break;
      }
    }
    const bool fin = ended || c == 1;
    bs[d] = st | (fin ? 0x80000000u : 0u);
    const uint32_t cb = sg.begin + st;
    if (d != dfirst) {
      const uint32_t slot = atomicAdd(&p.ctr->boundary, 1u);
      if (slot < p.cap_bounds)
        p.bounds[slot] = Boundary{cb - 1, sg.depth, nbuf, 0u};
      else
        atomicOr(&p.ctr->error, 16u);
    }
    if (ended) {
      if (c > 1) push_retire(p, cb, c, nbuf, sg.depth + nul, 0u);
    } else if (!fin) {
      push_child(p, next, cb, c, sg.depth + fb, nbuf);
    }
  }
}

// ---------------------------------------------------------------------------
// Global pass, step 4: stable scatter with decoupled look-back.

constexpr uint64_t kFlagA = 1ull << 46;  // aggregate of this tile only
constexpr uint64_t kFlagP = 2ull << 46;  // inclusive prefix up to this tile
constexpr uint64_t kValMask = (1ull << 46) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

template <int RB>
__global__ __launch_bounds__(kTileThreads) void k_scatter(Plan p, const BigSeg* segs,
                                                          uint32_t ticket_base,
                                                          uint32_t epoch) {
  constexpr uint32_t NB = 1u << RB;
  constexpr int W = kTileThreads / kWave;
  constexpr int IPW = kTile / W;  // items per wave
  constexpr uint32_t DPT = NB >= kTileThreads ? NB / kTileThreads : 1;
  __shared__ uint16_t s_whist[W][NB];
  __shared__ uint32_t s_tstart[NB];
  __shared__ uint32_t s_gbase[NB];
  __shared__ uint32_t s_off[kTile];
  __shared__ uint16_t s_dig[kTile];
  __shared__ uint32_t s_scan[W + 1];
  __shared__ uint32_t s_vt;

  if (threadIdx.x == 0) s_vt = atomicAdd(&p.ctr->tile_ticket, 1u) - ticket_base;
  __syncthreads();
  const uint32_t vt = s_vt;
  const Tile t = p.tiles[vt];
  const BigSeg sg = segs[t.seg];
  const uint64_t mn = p.segmm[2 * t.seg], mx = p.segmm[2 * t.seg + 1];
  if (mn == mx) return;  // whole segment skipped, no look-back partner
  const uint32_t lzb = (uint32_t)__builtin_clzll(mn ^ mx);
  const uint32_t* __restrict__ src = p.idx[sg.buf];
  uint32_t* __restrict__ dst = p.idx[sg.buf == 1 ? 2 : 1];

  for (uint32_t i = threadIdx.x; i < W * NB; i += kTileThreads) (&s_whist[0][0])[i] = 0;
  __syncthreads();

  const int wid = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();
  uint32_t off[kTileItems], dig[kTileItems], rnk[kTileItems];
#pragma unroll
  for (int j = 0; j < kTileItems; ++j) {
    const uint32_t i = wid * IPW + j * kWave + lane;
    const bool v = i < t.len;
    uint32_t d = 0;
    off[j] = 0;
    if (v) {
      off[j] = src[t.begin + i];
      d = seg_digit<RB>(p.keys[t.begin + i], lzb);
    }
    dig[j] = d;
    const uint64_t live = __ballot(v);
    const uint64_t m = match_digit<RB>(d, live);
    rnk[j] = 0;
    if (v) {
      const uint32_t base = s_whist[wid][d];
      rnk[j] = base + (uint32_t)__popcll(m & lt);
      if (lane == __builtin_ctzll(m)) s_whist[wid][d] = (uint16_t)(base + __popcll(m));
    }
  }
  __syncthreads();

  // Per digit: exclusive prefix over waves (in place) and tile total.
  uint32_t tot[DPT];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    uint32_t r = 0;
    if (d < NB) {
      for (int w = 0; w < W; ++w) {
        const uint32_t c = s_whist[w][d];
        s_whist[w][d] = (uint16_t)r;
        r += c;
      }
    }
    tot[q] = r;
    sum += r;
  }
  uint32_t tsum;
  uint32_t run = block_excl_scan<kTileThreads>(sum, s_scan, &tsum);
  const uint32_t* bs = p.bstart + (uint64_t)t.seg * NB;
#pragma unroll
  for (uint32_t q = 0; q < DPT; ++q) {
    const uint32_t d = threadIdx.x * DPT + q;
    if (d >= NB) continue;
    s_tstart[d] = run;
    const uint32_t c = tot[q];
    run += c;
    // Decoupled look-back over the earlier tiles of this segment.
    uint64_t* st = p.status + (uint64_t)vt * NB + d;
    const uint64_t tag = (uint64_t)epoch << 48;
    uint64_t prefix = 0;
    if (t.first) {
      __hip_atomic_store(st, tag | kFlagP | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(st, tag | kFlagA | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint64_t* look = st - NB;
      uint32_t spins = 0;
      while (true) {
        const uint64_t w = __hip_atomic_load(look, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool ready = (w >> 48) == epoch && (w & (kFlagA | kFlagP));
        if (!ready) {
          if (++spins > kSpinLimit) {
            atomicOr(&p.ctr->error, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        prefix += w & kValMask;
        if (w & kFlagP) break;
        look -= NB;
      }
      __hip_atomic_store(st, tag | kFlagP | (prefix + c), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gbase[d] = sg.begin + (bs[d] & 0x7FFFFFFFu) + (uint32_t)prefix;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTileItems; ++j) {
    const uint32_t i = wid * IPW + j * kWave + lane;
    if (i < t.len) {
      const uint32_t d = dig[j];
      const uint32_t lp = s_tstart[d] + s_whist[wid][d] + rnk[j];
      s_off[lp] = off[j];
      s_dig[lp] = (uint16_t)d;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < t.len; i += kTileThreads) {
    const uint32_t d = s_dig[i];
    const uint32_t pos = s_gbase[d] + (i - s_tstart[d]);
    const uint32_t o = s_off[i];
    // Positions stay inside the segment unless a look-back wait gave up
    // (error flagged above); never write outside it.
    if (pos - sg.begin >= sg.size) continue;
    dst[pos] = o;
    if (bs[d] & 0x80000000u) p.out[pos] = o;
  }
}

// ---------------------------------------------------------------------------
// Global pass, step 5: LCP of pairs split by this pass.

__global__ void k_boundary_lcp(Plan p, uint32_t first) {
  const uint32_t n = min(p.ctr->boundary, p.cap_bounds);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const Boundary b = p.bounds[i];
    const uint32_t* ix = p.idx[b.buf];
    const uint64_t k1 = load_key(p.text64, p.nwords, (uint64_t)ix[b.pos] + b.depth);
    const uint64_t k2 = load_key(p.text64, p.nwords, (uint64_t)ix[b.pos + 1] + b.depth);
    p.lcp[b.pos] = (uint64_t)b.depth + key_lcp(k1, k2);
  }
}

// ---------------------------------------------------------------------------
// Runs of identical strings: copy offsets (when needed) and fill their LCPs.

__global__ void k_retire(Plan p, uint32_t count) {
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
// Local sort: one workgroup finishes one segment of at most CAP strings.

template <int BLOCK, int CAP>
struct LocalSmem {
  static constexpr int W = BLOCK / kWave;
  uint64_t key[CAP];
  uint32_t off[CAP];
  uint16_t sid[CAP];
  uint16_t ss_start[CAP / 2 + 2];  // sub-segment start in the active list
  uint16_t ss_base[CAP / 2 + 2];   // sub-segment start in the final order
  uint16_t whist[W][256];
  uint32_t scan[W + 1];
  uint64_t red[W];
};

// One stable LSD pass over the active list by an 8-bit digit: key byte
// `which` (0 = least significant) or, for which = 8 / 9, the low / high byte
// of the sub-segment id.
template <int BLOCK, int CAP>
__device__ void local_radix_pass(LocalSmem<BLOCK, CAP>& s, uint32_t nact, int which) {
  constexpr int W = BLOCK / kWave;
  constexpr int IPW = CAP / W;
  constexpr int IPT = IPW / kWave;
  const int wid = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();

  for (int i = threadIdx.x; i < W * 256; i += BLOCK) (&s.whist[0][0])[i] = 0;
  __syncthreads();

  uint64_t k[IPT];
  uint32_t o[IPT], sd[IPT], dg[IPT], rk[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = wid * IPW + j * kWave + lane;
    const bool v = i < nact;
    k[j] = 0;
    o[j] = 0;
    sd[j] = 0;
    uint32_t d = 0;
    if (v) {
      k[j] = s.key[i];
      o[j] = s.off[i];
      sd[j] = s.sid[i];
      d = which < 8 ? (uint32_t)(k[j] >> (8 * which)) & 0xFFu
                    : (which == 8 ? sd[j] & 0xFFu : sd[j] >> 8);
    }
    dg[j] = d;
    rk[j] = 0;
    if ((uint32_t)(wid * IPW + j * kWave) < nact) {  // wave-uniform
      const uint64_t live = __ballot(v);
      const uint64_t m = match_digit<8>(d, live);
      if (v) {
        const uint32_t base = s.whist[wid][d];
        rk[j] = base + (uint32_t)__popcll(m & lt);
        if (lane == __builtin_ctzll(m)) s.whist[wid][d] = (uint16_t)(base + __popcll(m));
      }
    }
  }
  __syncthreads();

  constexpr int DPT = 256 >= BLOCK ? 256 / BLOCK : 1;
  uint32_t tot[DPT];
  uint32_t sum = 0;
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int d = threadIdx.x * DPT + q;
    uint32_t r = 0;
    if (d < 256) {
      for (int w = 0; w < W; ++w) {
        const uint32_t c = s.whist[w][d];
        s.whist[w][d] = (uint16_t)r;
        r += c;
      }
    }
    tot[q] = r;
    sum += r;
  }
  uint32_t total;
  uint32_t run = block_excl_scan<BLOCK>(sum, s.scan, &total);
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int d = threadIdx.x * DPT + q;
    if (d < 256) {
      for (int w = 0; w < W; ++w) s.whist[w][d] = (uint16_t)(s.whist[w][d] + run);
    }
    run += tot[q];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = wid * IPW + j * kWave + lane;
    if (i < nact) {
      const uint32_t dst = s.whist[wid][dg[j]] + rk[j];
      s.key[dst] = k[j];
      s.off[dst] = o[j];
      s.sid[dst] = (uint16_t)sd[j];
    }
  }
  __syncthreads();
}

template <int BLOCK, int CAP>
__global__ __launch_bounds__(BLOCK) void k_local(Plan p, const LocalSeg* segs) {
  constexpr int IPT = CAP / BLOCK;
  __shared__ LocalSmem<BLOCK, CAP> s;
  const LocalSeg sg = segs[blockIdx.x];
  const uint32_t n = sg.size;
  const uint32_t gb = sg.begin;
  const uint32_t* __restrict__ src = p.idx[sg.buf];
  uint64_t* __restrict__ lcp = p.lcp;
  uint32_t* __restrict__ out = p.out;

  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
    s.off[i] = src[gb + i];
    s.sid[i] = 0;
  }
  if (threadIdx.x == 0) {
    s.ss_start[0] = 0;
    s.ss_start[1] = (uint16_t)n;
    s.ss_base[0] = 0;
  }
  uint32_t nact = n, nss = 1, maxss = n, depth = sg.depth;
  __syncthreads();

  while (true) {
    // 1. Gather the next 8 bytes of every active string.
    {
      uint32_t o[IPT];
      uint64_t k[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = j * BLOCK + threadIdx.x;
        o[j] = i < nact ? s.off[i] : 0u;
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = j * BLOCK + threadIdx.x;
        if (i < nact) k[j] = load_key(p.text64, p.nwords, (uint64_t)o[j] + depth);
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = j * BLOCK + threadIdx.x;
        if (i < nact) s.key[i] = k[j];
      }
    }
    __syncthreads();

    // 2. Which key bytes vary inside some sub-segment.
    uint64_t vm = 0;
    for (uint32_t i = threadIdx.x; i < nact; i += BLOCK)
      vm |= s.key[i] ^ s.key[s.ss_start[s.sid[i]]];
    vm = block_or<BLOCK>(vm, s.red);

    // 3. Sort every sub-segment by key (stable).
    if (vm) {
      if (maxss <= 16) {
        for (uint32_t g = threadIdx.x; g < nss; g += BLOCK) {
          const uint32_t a = s.ss_start[g], e = s.ss_start[g + 1];
          for (uint32_t i = a + 1; i < e; ++i) {
            const uint64_t kk = s.key[i];
            const uint32_t oo = s.off[i];
            uint32_t j = i;
            while (j > a && s.key[j - 1] > kk) {
              s.key[j] = s.key[j - 1];
              s.off[j] = s.off[j - 1];
              --j;
            }
            s.key[j] = kk;
            s.off[j] = oo;
          }
        }
        __syncthreads();
      } else {
        for (int b = 0; b < 8; ++b)
          if ((vm >> (8 * b)) & 0xFFu) local_radix_pass<BLOCK, CAP>(s, nact, b);
        if (nss > 1) local_radix_pass<BLOCK, CAP>(s, nact, 8);
        if (nss > 256) local_radix_pass<BLOCK, CAP>(s, nact, 9);
      }
    }

    // 4. Split into equal-key groups; retire singletons and groups whose key
    //    holds the NUL; emit the LCP of every pair split here.
    uint32_t keep_off[IPT], keep_pos[IPT];
    uint32_t flags = 0;  // bit 2j: member of a continuing group, bit 2j+1: its head
    uint32_t nmem = 0, nhead = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = threadIdx.x * IPT + j;
      keep_off[j] = 0;
      keep_pos[j] = 0;
      if (i < nact) {
        const uint64_t k = s.key[i];
        const uint32_t g = s.sid[i];
        const uint32_t a = s.ss_start[g], e = s.ss_start[g + 1];
        const bool has_prev = i > a;
        const uint64_t kp = has_prev ? s.key[i - 1] : 0;
        const bool first = !has_prev || kp != k;
        const bool last = i + 1 >= e || s.key[i + 1] != k;
        const uint32_t pos = s.ss_base[g] + (i - a);
        const uint32_t o = s.off[i];
        if (lcp && has_prev && kp != k) lcp[gb + pos - 1] = (uint64_t)depth + key_lcp(kp, k);
        const bool ended = (k & 0xFFu) == 0;
        if (ended || (first && last)) {
          out[gb + pos] = o;
          if (lcp && ended && !first) lcp[gb + pos - 1] = (uint64_t)depth + key_len(k);
        } else {
          flags |= 1u << (2 * j);
          nmem++;
          if (first) {
            flags |= 2u << (2 * j);
            nhead++;
          }
        }
        keep_off[j] = o;
        keep_pos[j] = pos;
      }
    }
    uint32_t tmem, thead;
    uint32_t m = block_excl_scan<BLOCK>(nmem, s.scan, &tmem);
    uint32_t h = block_excl_scan<BLOCK>(nhead, s.scan, &thead);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (flags & (1u << (2 * j))) {
        if (flags & (2u << (2 * j))) {
          s.ss_start[h] = (uint16_t)m;
          s.ss_base[h] = (uint16_t)keep_pos[j];
          ++h;
        }
        s.off[m] = keep_off[j];
        s.sid[m] = (uint16_t)(h - 1);
        ++m;
      }
    }
    nact = tmem;
    nss = thead;
    if (threadIdx.x == 0) s.ss_start[nss] = (uint16_t)nact;
    __syncthreads();
    if (nact == 0) break;
    uint32_t mx = 0;
    for (uint32_t g = threadIdx.x; g < nss; g += BLOCK)
      mx = max(mx, (uint32_t)(s.ss_start[g + 1] - s.ss_start[g]));
    maxss = block_max<BLOCK>(mx, s.scan);
    depth += 8;
    // Every string ends inside the text (reads past it return NUL), so no
    // group can stay active beyond the text length; stop rather than spin.
    if ((uint64_t)depth > p.nwords * 8 + 8) {
      if (threadIdx.x == 0) atomicOr(&p.ctr->error, 32u);
      break;
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

hipError_t launch_init_segs(const Plan& p, uint32_t nsegs, uint32_t nbins, hipStream_t s) {
  k_init_segs<<<grid_for((uint64_t)nsegs * nbins, 256), 256, 0, s>>>(p.segmm, p.seghist,
                                                                      nsegs, nbins);
  return hipGetLastError();
}

hipError_t launch_gather_minmax(const Plan& p, const BigSeg* segs, uint32_t ntiles,
                                hipStream_t s) {
  k_gather_minmax<<<ntiles, kTileThreads, 0, s>>>(p, segs);
  return hipGetLastError();
}

hipError_t launch_histogram(const Plan& p, const BigSeg* segs, uint32_t ntiles,
                            uint32_t radix_bits, hipStream_t s) {
  (void)segs;
  if (radix_bits == 8)
    k_histogram<8><<<ntiles, kTileThreads, 0, s>>>(p);
  else if (radix_bits == 11)
    k_histogram<11><<<ntiles, kTileThreads, 0, s>>>(p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_segscan(const Plan& p, const BigSeg* segs, BigSeg* next, uint32_t nsegs,
                          uint32_t pass, uint32_t radix_bits, hipStream_t s) {
  (void)pass;
  if (radix_bits == 8)
    k_segscan<8><<<nsegs, 256, 0, s>>>(p, segs, next);
  else if (radix_bits == 11)
    k_segscan<11><<<nsegs, 256, 0, s>>>(p, segs, next);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_scatter(const Plan& p, const BigSeg* segs, uint32_t ntiles, uint32_t pass,
                          uint32_t epoch, uint32_t radix_bits, hipStream_t s) {
  // `pass` carries the ticket base: tickets grow monotonically over a sort.
  if (radix_bits == 8)
    k_scatter<8><<<ntiles, kTileThreads, 0, s>>>(p, segs, pass, epoch);
  else if (radix_bits == 11)
    k_scatter<11><<<ntiles, kTileThreads, 0, s>>>(p, segs, pass, epoch);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_boundary_lcp(const Plan& p, uint32_t first, hipStream_t s) {
  k_boundary_lcp<<<1024, 256, 0, s>>>(p, first);
  return hipGetLastError();
}

hipError_t launch_retire(const Plan& p, uint32_t nretire, hipStream_t s) {
  if (!nretire) return hipSuccess;
  k_retire<<<grid_for(nretire, 1, 8192), 256, 0, s>>>(p, nretire);
  return hipGetLastError();
}

hipError_t launch_local(const Plan& p, int cls, uint32_t nsegs, hipStream_t s) {
  if (!nsegs) return hipSuccess;
  switch (cls) {
    case 0:
      k_local<64, 64><<<nsegs, 64, 0, s>>>(p, p.local[0]);
      break;
    case 1:
      k_local<256, 1024><<<nsegs, 256, 0, s>>>(p, p.local[1]);
      break;
    case 2:
      k_local<1024, 8192><<<nsegs, 1024, 0, s>>>(p, p.local[2]);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_zero_counters(const Plan& p, hipStream_t s) {
  k_zero_counters<<<1, 64, 0, s>>>(p.ctr);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* dst, uint64_t n, uint64_t v, hipStream_t s) {
  if (!n) return hipSuccess;
  k_fill_u64<<<grid_for(n, 256), 256, 0, s>>>(dst, n, v);
  return hipGetLastError();
}

hipError_t launch_ptr_to_off(const uint64_t* ptrs, uint64_t base, uint32_t* off, uint64_t n,
                             hipStream_t s) {
  k_ptr_to_off<<<grid_for(n, 256), 256, 0, s>>>(ptrs, base, off, n);
  return hipGetLastError();
}

hipError_t launch_off_to_ptr(const uint32_t* off, uint64_t base, uint64_t* ptrs, uint64_t n,
                             hipStream_t s) {
  k_off_to_ptr<<<grid_for(n, 256), 256, 0, s>>>(off, base, ptrs, n);
  return hipGetLastError();
}

hipError_t launch_off_to_orig_ptr(const uint32_t* off, const uint32_t* starts,
                                  const uint64_t* orig, uint64_t* ptrs, uint64_t n,
                                  hipStream_t s) {
  k_off_to_orig_ptr<<<grid_for(n, 256), 256, 0, s>>>(off, starts, orig, ptrs, n);
  return hipGetLastError();
}

}  // namespace gss
