// Workgroup-local finish of the MI355X string sort (gfx950, wave64).
//
// One workgroup takes one segment of at most CAP strings (offsets into the
// text, all equal on their first `depth` bytes) and sorts it completely in
// LDS.  The segment is kept as an ordered list of active sub-segments (groups
// of strings still equal so far), each with its own depth, and rounds repeat:
//
//   1. gather  the next kLocalWords 8-byte keys (32 bytes) of every active
//              string at its sub-segment's depth, big-endian and NUL-padded
//              like multikey_cache's cache (src/multikey_cache.cpp:107-125);
//   2. choose  per sub-segment the first of those keys that differs between
//              two members or holds a NUL; keys before it are equal for the
//              whole group and are skipped.  A group equal on all 32 bytes
//              with no NUL just moves 32 bytes deeper;
//   3. sort    every sub-segment by its chosen key, stably: LSD radix passes
//              over the key bytes that vary inside some sub-segment, then
//              over the sub-segment id (wave match-ranking with per-wave LDS
//              digit counters); when no sub-segment exceeds insertion_max
//              strings each is insertion-sorted by one thread (16, middle
//              class), or every item is ranked inside its group by counting
//              (128 in the largest class, always in the one-wave class);
//   4. split   into equal-key groups.  A singleton, or a group whose key holds
//              the NUL (identical strings), is final -- the retirement rule
//              of multikey_cache (src/multikey_cache.cpp:207-211) and of
//              msd_CE2's bucket 0 (src/msd_ce.cpp:168-173).  Every pair split
//              here gets its LCP (src/mergesort_lcp.cpp:88-100); survivors are
//              compacted in order into the next round's list.
//
// Items are laid out per wave: each round the nact active strings are spread
// as wave w -> [w*ipw, (w+1)*ipw), lane l, slot j -> w*ipw + 64*j + l, so all
// waves share the work and per-slot LDS addresses are immediate offsets.

#include "gss_internal.h"

namespace gss {

constexpr int kLocalWords = 4;

// The largest class's first round (see k_local) reads PW words per string,
// the tuning's key_bytes: 64 by default, so a group whose strings share up to
// 56 more bytes (URL-like prefix groups) is split in that round instead of
// after one or more rounds that only find the next 32 bytes equal; 32 reads
// no more than an ordinary round.  The smaller classes keep the ordinary
// first round: with 32-byte windows the probe saves them no bytes, and its
// registers cost them a wave per SIMD.
constexpr int kProbeWordsMax = 8;
constexpr bool probes_first_round(uint32_t cap) { return cap > kLocalClassCap[1]; }

// Largest group a round still sorts without the LSD passes over all items.
// The 1024-thread class and the one-wave class rank every item inside its
// group by counting (independent LDS reads, broadcast across the lanes of a
// group): up to 128 strings in the former, cheaper than the passes up to
// that size, and always in the latter, one wave over at most 128 strings
// (DNA 32M, whose time is in that class, 5.22 -> 4.53 ms).  The middle class
// insertion-sorts groups of up to 16 strings in one thread each (counting
// there gained nothing).  10M URL strings: 1.79 ms with 16 for all classes,
// 1.72 ms with a 64-string insertion sort, 1.68 ms counting up to 128 (256
// no better, 512 slower).
constexpr uint32_t kCountMax = 128;
constexpr uint32_t insertion_max(uint32_t cap) {
  return cap > kLocalClassCap[1] ? kCountMax : (cap <= kLocalClassCap[0] ? cap : 16u);
}
// Classes that count ranks rather than insertion-sort.
constexpr bool counts_ranks(uint32_t cap) {
  return cap > kLocalClassCap[1] || cap <= kLocalClassCap[0];
}

// A segment's first sorting round in the largest class ranks its strings by
// one digit of the two most significant varying key bytes (each byte ranked
// in the alphabet found at its position) when the digit takes at most this
// many values.
constexpr uint32_t kFirstDigits = 2048;

template <int BLOCK, int CAP>
struct LocalSmem {
  static constexpr int W = BLOCK / kWave;
  static constexpr int SS = CAP / 2 + 2;
  static constexpr bool kFirst = CAP > kLocalClassCap[1];  // first-round counting pass
  // First-round counting pass: bucket counts, then starts (entry nbu = total);
  // the byte values present at the two digit positions and their ranks.
  uint32_t fcnt[kFirst ? kFirstDigits + 1 : 1];
  uint8_t fpres[kFirst ? 2 : 1][256];
  uint8_t frank[kFirst ? 2 : 1][256];
  uint64_t key[CAP];
  uint32_t off[CAP];
  uint16_t sid[CAP];
  uint32_t ss_depth[SS];  // byte depth the sub-segment's next key starts at
  uint16_t ss_start[SS];  // start in the active list (entry nss = nact)
  uint16_t ss_base[SS];   // start in the segment's final order
  uint8_t ss_stop[(SS + 3) & ~3];  // bit q: key q varies or holds a NUL in the group
  alignas(8) uint16_t whist[256][W];  // per-wave digit counters, digit-major
  uint64_t edge[W][kLocalWords];  // keys of each wave's last item
  uint64_t refw[kProbeWordsMax];  // first round: the reference string's window
  uint32_t wvm[W];   // per wave: varying key bytes | 0x100 if a group is > insertion_max
  uint32_t wtot[W];  // per wave: survivors | their group heads << 16
  uint32_t scan[W + 1];
};

// Lane i receives lane i-1's value, lane 0 receives 0 (DPP wave_shr:1).
__device__ __forceinline__ uint64_t wave_shr1_64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)v, 0x138, 0xF, 0xF, false);
  const uint32_t hi =
      __builtin_amdgcn_update_dpp(0u, (uint32_t)(v >> 32), 0x138, 0xF, 0xF, false);
  return ((uint64_t)hi << 32) | lo;
}

// Lane i receives lane i-1's value, lane 0 receives `first`.
__device__ __forceinline__ uint32_t wave_shr1_32(uint32_t v, uint32_t first) {
  return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xF, 0xF, false);
}

__device__ __forceinline__ uint64_t lane63_64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, kWave - 1);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), kWave - 1);
  return ((uint64_t)hi << 32) | lo;
}

// Bit b set when byte b (0 = least significant) of x is nonzero.
__device__ __forceinline__ uint32_t nonzero_bytes(uint64_t x) {
  const uint64_t t = ((x & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | x;
  const uint64_t h = t & 0x8080808080808080ull;  // top bit of each nonzero byte
  return (uint32_t)((h * 0x0002040810204081ull) >> 56);
}

// One stable LSD pass over the active list by an 8-bit digit: key byte
// `which` (0 = least significant) or, for which = 8 / 9, the low / high byte
// of the sub-segment id.
template <int BLOCK, int CAP>
__device__ __forceinline__ void local_radix_pass(LocalSmem<BLOCK, CAP>& s, uint32_t nact,
                                                 uint32_t base, uint32_t jn, int which) {
  constexpr int W = BLOCK / kWave;
  constexpr int IPT = CAP / BLOCK;
  const int wid = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();

  // The items are read before the counters are cleared, so the LDS reads
  // overlap the barrier; the ranking below is then one chain of counter
  // reads and leader writes per item.
  uint64_t k[IPT];
  uint32_t o[IPT], sd[IPT], rk[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    k[j] = 0;
    o[j] = 0;
    sd[j] = 0;
    rk[j] = 0;
    const uint32_t i = base + j * kWave;
    if ((uint32_t)j < jn && i < nact) {
      k[j] = s.key[i];
      o[j] = s.off[i];
      sd[j] = s.sid[i];
    }
  }

  // 256 * W counters, 4 per thread (W = BLOCK / 64).
  static_assert(256 * W == 4 * BLOCK, "four digit counters per thread");
  uint64_t* const hist4 = reinterpret_cast<uint64_t*>(&s.whist[0][0]);
  hist4[threadIdx.x] = 0;
  __syncthreads();

#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if ((uint32_t)j < jn) {
      const bool v = base + j * kWave < nact;
      const uint32_t d = which < 8 ? (uint32_t)(k[j] >> (8 * which)) & 0xFFu
                                   : (which == 8 ? sd[j] & 0xFFu : sd[j] >> 8);
      const uint64_t m = match_digit<8>(d, __ballot(v));
      if (v) {
        const uint32_t b = s.whist[d][wid];
        rk[j] = (d << 16) | (b + (uint32_t)__popcll(m & lt));
        if (lane == __builtin_ctzll(m)) s.whist[d][wid] = (uint16_t)(b + __popcll(m));
      }
    }
  }
  __syncthreads();

  // Exclusive starts of every (digit, wave) in digit-major order: each
  // thread scans its four contiguous counters, one block scan joins them.
  {
    const uint64_t c4 = hist4[threadIdx.x];
    const uint32_t c0 = (uint32_t)(c4 & 0xFFFFu), c1 = (uint32_t)((c4 >> 16) & 0xFFFFu);
    const uint32_t c2 = (uint32_t)((c4 >> 32) & 0xFFFFu), c3 = (uint32_t)(c4 >> 48);
    uint32_t total;
    const uint32_t r0 = block_excl_scan<BLOCK>(c0 + c1 + c2 + c3, s.scan, &total);
    const uint32_t r1 = r0 + c0, r2 = r1 + c1, r3 = r2 + c2;
    hist4[threadIdx.x] = (uint64_t)r0 | ((uint64_t)r1 << 16) | ((uint64_t)r2 << 32) |
                         ((uint64_t)r3 << 48);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if ((uint32_t)j < jn && base + j * kWave < nact) {
      const uint32_t dst = s.whist[rk[j] >> 16][wid] + (rk[j] & 0xFFFFu);
      s.key[dst] = k[j];
      s.off[dst] = o[j];
      s.sid[dst] = (uint16_t)sd[j];
    }
  }
  __syncthreads();
}

// A round with one group (a segment's first sorting round): stable passes
// by its two most significant varying bytes leave sub-buckets of strings
// equal on both, and every item is then ranked inside its sub-bucket by
// counting (the sub-bucket's items that order before it).  A sub-bucket
// starts at a change of the two bytes; each item finds the start by ballots
// within its wave and the last start of the waves before it.  Returns false,
// the items sorted by the two bytes only, when a sub-bucket holds more than
// kCountMax strings (the caller's LSD passes then sort the group: they are
// stable, so the order these passes left does not matter).
template <int BLOCK, int CAP>
__device__ __forceinline__ bool local_two_byte_count(LocalSmem<BLOCK, CAP>& s, uint32_t nact,
                                                     uint32_t ipw, uint32_t base, uint32_t jn,
                                                     uint32_t vm, uint32_t count_max) {
  constexpr int IPT = CAP / BLOCK;
  constexpr int W = BLOCK / kWave;
  const int top = 31 - __builtin_clz(vm);
  const uint32_t rest = vm & ~(1u << top);
  if (rest) local_radix_pass<BLOCK, CAP>(s, nact, base, jn, 31 - __builtin_clz(rest));
  local_radix_pass<BLOCK, CAP>(s, nact, base, jn, top);
  if (!rest) return true;  // one varying byte: that pass sorted the group
  const int sec = 31 - __builtin_clz(rest);
  const uint32_t wid = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  const uint32_t slot0 = wid * ipw;  // first item of this wave
  auto tag = [&](uint64_t k) -> uint32_t {
    return ((uint32_t)(k >> (8 * top)) & 0xFFu) << 8 | ((uint32_t)(k >> (8 * sec)) & 0xFFu);
  };
  // Last sub-bucket start of this wave, then of the waves before it.
  uint32_t last = 0;
  bool has = false;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * kWave;
    const bool v = (uint32_t)j < jn && i < nact;
    const bool head = v && (i == 0 || tag(s.key[i]) != tag(s.key[i > 0 ? i - 1 : 0]));
    const uint64_t bh = __ballot(head);
    if (bh) {
      last = slot0 + j * kWave + (63 - __builtin_clzll(bh));
      has = true;
    }
  }
  if (lane == 0) s.wtot[wid] = has ? last : 0xFFFFFFFFu;
  __syncthreads();
  uint32_t carry = 0;  // item 0 starts a sub-bucket, so wave 0 always has one
  for (uint32_t w = 0; w < wid; ++w) {
    const uint32_t h = s.wtot[w];
    if (h != 0xFFFFFFFFu) carry = h;
  }
  // Each item's sub-bucket start goes to sid, which holds 0 for every item of
  // a one-group round; no item may lie kCountMax or more past its start
  // (checked before any counting, so a big run of equal strings costs one
  // scan of starts, not a quadratic count).  Then each item counts its
  // sub-bucket and sid takes its destination: nothing is held in registers
  // across the items.
  bool over = false;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * kWave;
    const bool v = (uint32_t)j < jn && i < nact;
    const bool head = v && (i == 0 || tag(s.key[i]) != tag(s.key[i > 0 ? i - 1 : 0]));
    const uint64_t bh = __ballot(head);
    const uint64_t le = bh & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const uint32_t lo = le ? slot0 + j * kWave + (63 - __builtin_clzll(le)) : carry;
    if (bh) carry = slot0 + j * kWave + (63 - __builtin_clzll(bh));
    if (v) {
      over = over || i - lo >= count_max;
      s.sid[i] = (uint16_t)lo;
    }
  }
  if (!__syncthreads_or(over)) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = base + j * kWave;
      if ((uint32_t)j < jn && i < nact) {
        const uint32_t lo = s.sid[i];
        const uint64_t k = s.key[i];
        const uint32_t t = tag(k);
        uint32_t r = 0;
        for (uint32_t m = lo; m < nact; ++m) {
          const uint64_t km = s.key[m];
          if (tag(km) != t) break;
          r += (km < k || (km == k && m < i)) ? 1u : 0u;
        }
        s.sid[i] = (uint16_t)(lo + r);
      }
    }
  }
  const bool bail = __syncthreads_or(over);  // block-uniform, as above
  uint32_t cd[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * kWave;
    cd[j] = (uint32_t)j < jn && i < nact ? s.sid[i] : 0xFFFFFFFFu;
  }
  if (!bail) {
    // Keys, then offsets, each read whole before any is written.
    uint64_t ck[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) ck[j] = cd[j] != 0xFFFFFFFFu ? s.key[base + j * kWave] : 0ull;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (cd[j] != 0xFFFFFFFFu) s.key[cd[j]] = ck[j];
    uint32_t co[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) co[j] = cd[j] != 0xFFFFFFFFu ? s.off[base + j * kWave] : 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (cd[j] != 0xFFFFFFFFu) s.off[cd[j]] = co[j];
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j)
    if (cd[j] != 0xFFFFFFFFu) s.sid[base + j * kWave] = 0;  // the one group again
  __syncthreads();
  return !bail;
}

// A round with one group, at least two varying key bytes (largest class):
// each string's digit is (rank of its top varying byte, rank of the next
// one), both ranked among the byte values that occur at that position, so
// digits order strings as those two bytes do.  One counting pass (LDS
// atomics) puts every string into its digit's bucket; then each string's
// place inside its bucket is the number of bucket members with a smaller key,
// or the same key and a smaller index -- the stable order.  Inside a bucket
// the two digit bytes and every byte above them are the same, so a string
// is ranked by one 64-bit value, its key bytes below the second digit byte
// (at most six) shifted up 16 bits, OR its index before the round: distinct
// values that order as (key, index) do.  Returns 1 when the group is sorted; 0, nothing moved,
// when the digits take more than kFirstDigits values (the caller sorts by the
// two-byte passes); -1, nothing moved, when a bucket holds more than
// kCountMax strings (the caller's LSD passes then sort the group).
template <int BLOCK, int CAP>
__device__ __forceinline__ int local_first_count(LocalSmem<BLOCK, CAP>& s, uint32_t nact,
                                                 uint32_t base, uint32_t jn, uint32_t vm) {
  constexpr int IPT = CAP / BLOCK;
  static_assert(BLOCK == 1024 && kFirstDigits == 2 * BLOCK, "two digit counters per thread");
  const uint32_t wid = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  const int top = 31 - __builtin_clz(vm);
  const int sec = 31 - __builtin_clz(vm & ~(1u << top));
  // Byte values present at both positions; the counters cleared meanwhile.
  if (threadIdx.x < 128) reinterpret_cast<uint32_t*>(&s.fpres[0][0])[threadIdx.x] = 0u;
  s.fcnt[2 * threadIdx.x] = 0u;
  s.fcnt[2 * threadIdx.x + 1] = 0u;
  uint64_t k[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * kWave;
    k[j] = (uint32_t)j < jn && i < nact ? s.key[i] : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if ((uint32_t)j < jn && base + j * kWave < nact) {
      s.fpres[0][(uint32_t)(k[j] >> (8 * top)) & 0xFFu] = 1;
      s.fpres[1][(uint32_t)(k[j] >> (8 * sec)) & 0xFFu] = 1;
    }
  }
  __syncthreads();
  // Ranks: wave 0 for the top byte, wave 1 for the second.
  if (wid < 2) {
    const uint64_t lt = lanemask_lt();
    uint32_t cnt = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint64_t bal = __ballot(s.fpres[wid][64 * c + lane] != 0);
      s.frank[wid][64 * c + lane] = (uint8_t)(cnt + (uint32_t)__popcll(bal & lt));
      cnt += (uint32_t)__popcll(bal);
    }
    if (lane == 0) s.wtot[wid] = cnt;
  }
  __syncthreads();
  const uint32_t a1 = s.wtot[0], a2 = s.wtot[1];
  if (a1 * a2 > kFirstDigits) return 0;  // block-uniform
  uint32_t d[IPT], r[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    d[j] = 0;
    r[j] = 0;
    if ((uint32_t)j < jn && base + j * kWave < nact) {
      d[j] = (uint32_t)s.frank[0][(uint32_t)(k[j] >> (8 * top)) & 0xFFu] * a2 +
             s.frank[1][(uint32_t)(k[j] >> (8 * sec)) & 0xFFu];
      r[j] = atomicAdd(&s.fcnt[d[j]], 1u);
    }
  }
  __syncthreads();
  // Bucket starts; a bucket over kCountMax strings sends the group to the
  // LSD passes before anything moves.
  const uint32_t c0 = s.fcnt[2 * threadIdx.x], c1 = s.fcnt[2 * threadIdx.x + 1];
  if (__syncthreads_or(c0 > kCountMax || c1 > kCountMax)) return -1;
  uint32_t total;
  const uint32_t st = block_excl_scan<BLOCK>(c0 + c1, s.scan, &total);
  s.fcnt[2 * threadIdx.x] = st;
  s.fcnt[2 * threadIdx.x + 1] = st + c0;
  if (threadIdx.x == 0) s.fcnt[kFirstDigits] = total;
  uint32_t o[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * kWave;
    o[j] = (uint32_t)j < jn && i < nact ? s.off[i] : 0u;
  }
  __syncthreads();
  // Into digit order, as rank values (the offsets stay in registers).
  static_assert(CAP <= 65536, "indices fit 16 bits");
  const uint64_t low = sec ? (1ull << (8 * sec)) - 1ull : 0ull;  // sec <= 6
  uint32_t lo[IPT], hi[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    lo[j] = 0;
    hi[j] = 0;
    const uint32_t i = base + j * kWave;
    if ((uint32_t)j < jn && i < nact) {
      lo[j] = s.fcnt[d[j]];
      hi[j] = s.fcnt[d[j] + 1];
      s.key[lo[j] + r[j]] = ((k[j] & low) << 16) | i;
    }
  }
  __syncthreads();
  // Place inside the bucket: the members with a smaller rank value, four
  // reads in flight at a time.
  uint32_t fin[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * kWave;
    fin[j] = 0xFFFFFFFFu;
    if ((uint32_t)j < jn && i < nact) {
      const uint64_t v = ((k[j] & low) << 16) | i;
      uint32_t c = 0, m = lo[j];
      for (; m + 4 <= hi[j]; m += 4) {
        const uint64_t a0 = s.key[m], a1 = s.key[m + 1], a2 = s.key[m + 2], a3 = s.key[m + 3];
        c += (a0 < v ? 1u : 0u) + (a1 < v ? 1u : 0u) + (a2 < v ? 1u : 0u) + (a3 < v ? 1u : 0u);
      }
      for (; m < hi[j]; ++m) c += s.key[m] < v ? 1u : 0u;
      fin[j] = lo[j] + c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if (fin[j] != 0xFFFFFFFFu) {
      s.key[fin[j]] = k[j];
      s.off[fin[j]] = o[j];
      s.sid[fin[j]] = 0;  // the one group again
    }
  }
  __syncthreads();
  return 1;
}

// COUNT: also count the 8-byte key levels the strings advance through (for
// profiling only: the count costs registers).
// Waves per SIMD each class is compiled for (the second launch bound): the
// small classes run many workgroups per CU and keep their occupancy.
constexpr int local_min_waves(uint32_t cap) {
  return cap <= kLocalClassCap[0] ? 7 : (cap <= kLocalClassCap[1] ? 6 : 4);
}

// Block b finishes segment first + b of the class's list; PW: the first
// round's words per string (largest class only).
template <int BLOCK, int CAP, bool COUNT, int PW = kProbeWordsMax>
__global__ __launch_bounds__(BLOCK, local_min_waves(CAP)) void k_local(Plan p, uint32_t first) {
  constexpr int W = BLOCK / kWave;
  constexpr int IPT = CAP / BLOCK;
  constexpr int KW = kLocalWords;
  constexpr int kClass = CAP <= kLocalClassCap[0] ? 0 : (CAP <= kLocalClassCap[1] ? 1 : 2);
  constexpr uint64_t kSkipKey = ~0ull;  // key of a group that only moves deeper
  static_assert(CAP % BLOCK == 0 && (CAP / W) % kWave == 0, "layout");
  __shared__ LocalSmem<BLOCK, CAP> s;
  uint64_t* __restrict__ lcp = p.lcp;
  uint32_t* __restrict__ out = p.out;
  const uint32_t wid = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  const uint64_t lt = lanemask_lt();
  unsigned long long round_strings = 0;
  // Bytes this wave's strings advanced through (eight per 8-byte key level):
  // a round's group that stops at key e consumed e + 1 levels, a skipping
  // group KW, a first round the bytes to the end of its key.  Built from
  // ballots, so the count is wave-uniform (scalar registers).
  unsigned long long wave_levels = 0;
  // A launch with the top bit of `first` set was sized before the host knew
  // the class's segment count (the largest class, launched right after the
  // passes): blocks past the device count have no segment.  Both loads are
  // issued before the test.
  const bool bounded = (first & 0x80000000u) != 0;
  first &= 0x7FFFFFFFu;
  const LocalSeg sg = p.local[kClass][first + blockIdx.x];
  if (bounded && first + blockIdx.x >= min(p.ctr->local[kClass].segs, p.cap_local[kClass])) return;
  const uint32_t n = sg.size;
  const uint32_t gb = sg.begin;
  const uint32_t* __restrict__ src = p.idx[sg.buf];

  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
    s.off[i] = src[gb + i];
    s.sid[i] = 0;
  }
  if (threadIdx.x == 0) {
    s.ss_start[0] = 0;
    s.ss_start[1] = (uint16_t)n;
    s.ss_base[0] = 0;
    s.ss_depth[0] = sg.depth;
    s.ss_stop[0] = 0;
  }
  uint32_t nact = n, nss = 1, rounds = 0;
  __syncthreads();

  while (true) {
    ++rounds;
    round_strings += nact;
    // Per-wave share of this round's active strings, a multiple of 64.
    const uint32_t ipw = (((nact + W - 1) / W) + kWave - 1) & ~(uint32_t)(kWave - 1);
    const uint32_t jn = ipw / kWave;
    const uint32_t base = wid * ipw + lane;

    uint32_t gv[IPT];  // sub-segment of the item (0xFFFF when inactive)
    bool wave_big;     // some sub-segment holds more than insertion_max strings
    if (probes_first_round(CAP) && rounds == 1) {
      // 0. First round: one group, every string active.  Each string's PW
      //    words at the group depth are compared with those of the
      //    segment's first string (the reference): l_i = lcp(s_i, ref)
      //    within the window, the equal bytes before the first mismatch or
      //    NUL.  The group's LCP is G = min l_i (the LCP of a set is the
      //    minimum over its members' LCPs with any one of them, and a NUL
      //    ends every comparison), so the group's first word that differs
      //    or holds a NUL is at byte G.  The round sorts by the 8 bytes at G
      //    (at most PW * 8 - 8, so that the window holds them), which all
      //    start where the group first differs: the aligned word holding G
      //    would leave only 8 - G % 8 varying bytes in the key.  A string
      //    keeps only l_i and its own 8 bytes at l_i: its bytes [G, l_i) are
      //    the reference's, the rest its own.  A window with no mismatch and
      //    no NUL moves the group PW words deeper.
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        gv[j] = (uint32_t)j < jn && base + j * kWave < nact ? 0u : 0xFFFFu;
      wave_big = n > insertion_max(CAP);
      // l_i and the string's 8 bytes at l_i wait in s.sid and s.key (sid
      // is 0 for every string of a first round and is restored below).
      // The reference window passes through LDS (registers are this class's
      // limit).
      uint32_t depth = sg.depth;
      uint32_t G;
      for (;;) {
        if (threadIdx.x == 0) {
          uint64_t rw[PW];
          load_keys<PW>(p.text64, p.text_bytes, (uint64_t)s.off[0] + depth, rw);
#pragma unroll
          for (int q = 0; q < PW; ++q) s.refw[q] = rw[q];
        }
        __syncthreads();
        uint32_t lmin = 8 * PW;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          if (gv[j] != 0xFFFFu) {
            const uint32_t i = base + j * kWave;
            uint64_t w[PW];
            load_keys<PW>(p.text64, p.text_bytes, (uint64_t)s.off[i] + depth, w);
            uint32_t l = 8 * PW;
            uint64_t o = 0;
#pragma unroll
            for (int q = PW - 1; q >= 0; --q) {  // the first stop wins
              const uint64_t r = s.refw[q];  // the same word for every lane
              if (w[q] != r || (w[q] & 0xFFu) == 0) {
                const uint32_t lb = w[q] != r ? key_lcp(w[q], r) : key_len(w[q]);  // 0..7
                const uint64_t after = q + 1 < PW ? w[q + 1 < PW ? q + 1 : q] : 0ull;
                l = 8 * q + lb;
                o = lb ? (w[q] << (8 * lb)) | (after >> (64 - 8 * lb)) : w[q];
              }
            }
            s.sid[i] = (uint16_t)l;
            s.key[i] = o;
            lmin = min(lmin, l);
          }
        }
        lmin = wave_min(lmin);
        if (lane == 0) s.wtot[wid] = lmin;
        __syncthreads();
        G = 8 * PW;
#pragma unroll
        for (int w = 0; w < W; ++w) G = min(G, s.wtot[w]);
        if (G < 8u * PW) break;
        // Every string equals the reference on all PW words, with no NUL.
        if (COUNT) {
#pragma unroll
          for (int j = 0; j < IPT; ++j)
            if ((uint32_t)j < jn) wave_levels += 8 * PW * (uint64_t)__popcll(__ballot(gv[j] != 0xFFFFu));
        }
        depth += 8 * PW;
        ++rounds;
        round_strings += nact;
        __syncthreads();  // s.refw and s.wtot are rewritten
        if ((uint64_t)rounds > (p.text_bytes >> 3) + 3) break;
      }
      if (G >= 8u * PW) {  // the text ended without a NUL: cannot happen
        if (threadIdx.x == 0) atomicOr(&p.ctr->error, 32u);
        break;
      }
      const uint32_t gk = min(G, 8u * PW - 8);  // the key's byte in the window
      const uint32_t qk = gk / 8, rk = gk % 8;
      const uint64_t ref_k =
          rk ? (s.refw[qk] << (8 * rk)) | (s.refw[qk + 1 < PW ? qk + 1 : qk] >> (64 - 8 * rk))
             : s.refw[qk];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        if (gv[j] != 0xFFFFu) {
          const uint32_t i = base + j * kWave;
          const uint32_t l = s.sid[i];
          if (l >= gk + 8) {
            s.key[i] = ref_k;
          } else if (l > gk) {  // l - gk reference bytes, then its own
            const uint32_t m = l - gk;
            s.key[i] = (ref_k & (~0ull << (64 - 8 * m))) | (s.key[i] >> (8 * m));
          }  // l == gk: its own 8 bytes
          s.sid[i] = 0;
        }
        if (COUNT && (uint32_t)j < jn)
          wave_levels += (gk + 8) * (uint64_t)__popcll(__ballot(gv[j] != 0xFFFFu));
      }
      if (threadIdx.x == 0) {
        s.ss_depth[0] = depth + gk;  // the key is key 0 at this depth
        s.ss_stop[0] = 1;
      }
      __syncthreads();
    } else {
    // 1. Gather the next KW keys of every active string.
    uint64_t kw[IPT][KW];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
#pragma unroll
      for (int q = 0; q < KW; ++q) kw[j][q] = 0;
      gv[j] = 0xFFFFu;
      if ((uint32_t)j < jn) {
        const uint32_t i = base + j * kWave;
        const bool v = i < nact;
        const uint32_t g = v ? s.sid[i] : 0u;
        const uint32_t o = v ? s.off[i] : 0u;
        gv[j] = v ? g : 0xFFFFu;
        load_keys<KW>(p.text64, p.text_bytes, (uint64_t)o + s.ss_depth[g], kw[j]);
      }
    }
    if (lane == kWave - 1) {
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        if ((uint32_t)j + 1 == jn) {
#pragma unroll
          for (int q = 0; q < KW; ++q) s.edge[wid][q] = kw[j][q];
        }
    }
    __syncthreads();

    // 2. Per sub-segment: which keys differ between adjacent members (two
    //    members differ somewhere only if some adjacent pair does) or hold a
    //    NUL.  Also note whether any sub-segment exceeds insertion_max strings.
    bool big = false;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if ((uint32_t)j < jn) {
        const uint32_t i = base + j * kWave;
        uint32_t stop = 0;
        const bool same = gv[j] != 0xFFFFu && i > 0 && s.sid[i > 0 ? i - 1 : 0] == gv[j];
#pragma unroll
        for (int q = 0; q < KW; ++q) {
          const uint64_t up = wave_shr1_64(kw[j][q]);
          const uint64_t wrap = j > 0 ? lane63_64(kw[j > 0 ? j - 1 : 0][q])
                                      : (wid > 0 ? s.edge[wid > 0 ? wid - 1 : 0][q] : 0ull);
          const uint64_t prev = lane ? up : wrap;
          if ((same && kw[j][q] != prev) || (kw[j][q] & 0xFFu) == 0) stop |= 1u << q;
        }
        // Neighbouring lanes mostly share a group and its stop bits: only
        // the first lane of each run of equal (group, bits) issues the LDS
        // atomic, which would otherwise serialise on one word.
        const uint32_t tag = (gv[j] << 4) | stop;
        const uint32_t prev_tag = wave_shr1_32(tag, 0xFFFFFFFFu);
        if (gv[j] != 0xFFFFu) {
          if (stop && tag != prev_tag)
            atomicOr((uint32_t*)&s.ss_stop[0] + (gv[j] >> 2), stop << (8 * (gv[j] & 3)));
          big = big || (s.ss_start[gv[j] + 1] - s.ss_start[gv[j]]) > insertion_max(CAP);
        }
      }
    }
    wave_big = __ballot(big) != 0;
    __syncthreads();

    // A round in which no group found a differing key or a NUL only moves
    // every group KW keys deeper: nothing to sort, split or compact.
    bool stopped = false;
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if ((uint32_t)j < jn && gv[j] != 0xFFFFu && s.ss_stop[gv[j]] != 0) stopped = true;
    if (!__syncthreads_or(stopped)) {
      if (COUNT) {
#pragma unroll
        for (int j = 0; j < IPT; ++j)
          if ((uint32_t)j < jn) wave_levels += 8 * KW * (uint64_t)__popcll(__ballot(gv[j] != 0xFFFFu));
      }
      for (uint32_t g = threadIdx.x; g < nss; g += BLOCK) s.ss_depth[g] += 8 * KW;
      __syncthreads();
      if ((uint64_t)rounds > (p.text_bytes >> 3) + 3) {
        if (threadIdx.x == 0) atomicOr(&p.ctr->error, 32u);
        break;
      }
      continue;
    }

    // 3. Each string's key: the first key its group stops at, still in
    //    registers (reloading it would miss L2: a round's lines do not stay
    //    there); a group that does not stop gets kSkipKey.
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if ((uint32_t)j < jn) {
        uint32_t lv = 0;
        if (gv[j] != 0xFFFFu) {
          const uint32_t i = base + j * kWave;
          const uint32_t e = __builtin_ctz(s.ss_stop[gv[j]] | (1u << KW));
          uint64_t k = kSkipKey;
#pragma unroll
          for (int q = 0; q < KW; ++q)
            if ((uint32_t)q == e) k = kw[j][q];
          s.key[i] = k;
          lv = e < (uint32_t)KW ? e + 1 : (uint32_t)KW;
        }
        if (COUNT) {
#pragma unroll
          for (int q = 1; q <= KW; ++q)
            wave_levels += 8 * (uint64_t)__popcll(__ballot(lv >= (uint32_t)q));
        }
      }
    }
    __syncthreads();
    }  // rounds after the first

    // A segment's first round (largest class) keys every string at the
    // group's first differing byte, so the key's top byte varies: it counts
    // by bytes 7 and 6 at once, without the scan for varying bytes below
    // (that scan and the other sorts remain for the cases it declines).
    int pre_first = 2;  // 2: not tried
    if constexpr (kClass == 2) {
      if (rounds == 1) pre_first = local_first_count<BLOCK, CAP>(s, nact, base, jn, 0xC0u);
    }
    if (pre_first != 1) {
    // 4. Bytes of the chosen keys that vary inside some sub-segment.
    uint32_t vm = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if ((uint32_t)j < jn) {
        const uint32_t i = base + j * kWave;
        uint32_t nz = 0;
        if (gv[j] != 0xFFFFu && i > 0 && s.sid[i - 1] == gv[j])
          nz = nonzero_bytes(s.key[i] ^ s.key[i - 1]);
#pragma unroll
        for (int b = 0; b < 8; ++b)
          if (__ballot((nz >> b) & 1u)) vm |= 1u << b;
      }
    }
    if (lane == 0) s.wvm[wid] = vm | (wave_big ? 0x100u : 0u);
    __syncthreads();
    vm = 0;
    for (int w = 0; w < W; ++w) vm |= s.wvm[w];
    const bool big = (vm & 0x100u) != 0;
    vm &= 0xFFu;

    // 5. Sort every sub-segment by its key (stable).
    if (vm) {
      if (!big && counts_ranks(CAP)) {
        // Every item's stable rank inside its group by counting: it reads
        // the group's keys in order (the lanes of one group read the same
        // word at each step, which the LDS broadcasts), then all items move.
        // The reads are independent, unlike an insertion sort's chain.
        uint64_t ck[IPT];
        uint32_t co[IPT], cd[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          ck[j] = 0;
          co[j] = 0;
          cd[j] = 0xFFFFFFFFu;
          const uint32_t i = base + j * kWave;
          if ((uint32_t)j < jn && i < nact) {
            const uint32_t g = s.sid[i];
            const uint32_t a = s.ss_start[g], en = s.ss_start[g + 1];
            const uint64_t k = s.key[i];
            uint32_t r = 0;
            for (uint32_t m = a; m < en; ++m) {
              const uint64_t km = s.key[m];
              r += (km < k || (km == k && m < i)) ? 1u : 0u;
            }
            ck[j] = k;
            co[j] = s.off[i];
            cd[j] = a + r;
          }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          if (cd[j] != 0xFFFFFFFFu) {
            s.key[cd[j]] = ck[j];
            s.off[cd[j]] = co[j];
          }
        }
        __syncthreads();
      } else if (!big) {
        for (uint32_t g = threadIdx.x; g < nss; g += BLOCK) {
          const uint32_t a = s.ss_start[g], en = s.ss_start[g + 1];
          for (uint32_t i = a + 1; i < en; ++i) {
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
        // One group with two or more varying bytes: the counting pass, else
        // the two-byte passes; the LSD passes when a bucket is too large.
        int first = 0;
        if constexpr (kClass == 2) {
          // The two most significant varying bytes (the digit's bytes).
          const uint32_t hb = 1u << (31 - __builtin_clz(vm));
          const uint32_t rest = vm & ~hb;
          const uint32_t top2 = rest ? hb | (1u << (31 - __builtin_clz(rest))) : hb;
          if (pre_first != 2 && top2 == 0xC0u)
            first = pre_first;  // the same digit bytes: the same answer
          else if (nss == 1 && (vm & (vm - 1u)))
            first = local_first_count<BLOCK, CAP>(s, nact, base, jn, vm);
          if (first == 0 && nss == 1 &&
              local_two_byte_count<BLOCK, CAP>(s, nact, ipw, base, jn, vm, kCountMax))
            first = 1;
        }
        if (first != 1) {
          for (int b = 0; b < 8; ++b)
            if ((vm >> b) & 1u) local_radix_pass<BLOCK, CAP>(s, nact, base, jn, b);
          if (nss > 1) local_radix_pass<BLOCK, CAP>(s, nact, base, jn, 8);
          if (nss > 256) local_radix_pass<BLOCK, CAP>(s, nact, base, jn, 9);
        }
      }
    }
    }  // pre_first != 1

    // 6. Split into equal-key groups; retire singletons and groups whose key
    //    holds the NUL; emit the LCP of every pair split here.  Survivors are
    //    compacted in order: wave-local ranks from ballots, wave bases from
    //    the per-wave totals.
    uint32_t keep_off[IPT], keep_pos[IPT], keep_dep[IPT], rank[IPT];
    uint32_t wm = 0, wh = 0;  // members / heads of this wave so far
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      keep_off[j] = 0;
      keep_pos[j] = 0;
      keep_dep[j] = 0;
      rank[j] = 0x8000u;
      if ((uint32_t)j < jn) {
        const uint32_t i = base + j * kWave;
        bool member = false, head = false;
        if (i < nact) {
          const uint64_t k = s.key[i];
          const uint32_t g = s.sid[i];
          const uint32_t a = s.ss_start[g], en = s.ss_start[g + 1];
          const uint32_t e = __builtin_ctz(s.ss_stop[g] | (1u << KW));
          const uint32_t dep = s.ss_depth[g] + 8 * e;  // depth of key k
          const bool has_prev = i > a;
          const uint64_t kp = has_prev ? s.key[i - 1] : 0;
          const bool first = !has_prev || kp != k;
          const bool last = i + 1 >= en || s.key[i + 1] != k;
          const uint32_t pos = s.ss_base[g] + (i - a);
          const uint32_t o = s.off[i];
          if (lcp && has_prev && kp != k) lcp[gb + pos - 1] = (uint64_t)dep + key_lcp(kp, k);
          const bool ended = e < (uint32_t)KW && (k & 0xFFu) == 0;
          if (ended || (first && last)) {
            out[gb + pos] = o;
            if (lcp && ended && !first) lcp[gb + pos - 1] = (uint64_t)dep + key_len(k);
          } else {
            member = true;
            head = first;
          }
          keep_off[j] = o;
          keep_pos[j] = pos;
          keep_dep[j] = e < (uint32_t)KW ? dep + 8 : dep;  // e == KW: dep = depth + 32
        }
        const uint64_t bm = __ballot(member);
        const uint64_t bh = __ballot(head);
        // Heads of this wave up to and including this item's group head.
        const uint32_t hr = wh + (uint32_t)__popcll(bh & lt) + (head ? 1u : 0u);
        rank[j] = (wm + (uint32_t)__popcll(bm & lt)) | (head ? 0x4000u : 0u) |
                  (member ? 0u : 0x8000u) | (hr << 16);
        wm += (uint32_t)__popcll(bm);
        wh += (uint32_t)__popcll(bh);
      }
    }
    if (lane == 0) s.wtot[wid] = wm | (wh << 16);
    __syncthreads();
    uint32_t mbase = 0, tot = 0;
    for (uint32_t w = 0; w < (uint32_t)W; ++w) {
      const uint32_t t = s.wtot[w];
      if (w < wid) mbase += t;
      tot += t;
    }
    const uint32_t hbase = mbase >> 16;
    mbase &= 0xFFFFu;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if ((uint32_t)j < jn && !(rank[j] & 0x8000u)) {
        const uint32_t mi = mbase + (rank[j] & 0x3FFFu);
        // A member whose head sits in an earlier wave has hr == 0: its group
        // is the last one started before this wave (index hbase - 1).
        const uint32_t g = hbase + (rank[j] >> 16) - 1;
        s.off[mi] = keep_off[j];
        s.sid[mi] = (uint16_t)g;
        if (rank[j] & 0x4000u) {
          s.ss_start[g] = (uint16_t)mi;
          s.ss_base[g] = (uint16_t)keep_pos[j];
          s.ss_depth[g] = keep_dep[j];
          s.ss_stop[g] = 0;
        }
      }
    }
    nact = tot & 0xFFFFu;
    nss = tot >> 16;
    if (threadIdx.x == 0) s.ss_start[nss] = (uint16_t)nact;
    __syncthreads();
    if (nact == 0) break;
    // Every round moves each group >= 8 bytes deeper and no string reaches
    // past the text, so the rounds are bounded; stop rather than spin.
    if ((uint64_t)rounds > (p.text_bytes >> 3) + 3) {
      if (threadIdx.x == 0) atomicOr(&p.ctr->error, 32u);
      break;
    }
  }
  unsigned long long* slot = p.lstats->v[blockIdx.x % kStatSlots];
  if (COUNT && lane == 0 && wave_levels) atomicAdd(&slot[4 + kClass], wave_levels);
  if (threadIdx.x == 0) {
    atomicAdd(&slot[kClass], round_strings);
    atomicMax(&slot[3], (unsigned long long)rounds);
  }
}

template <bool COUNT>
hipError_t launch_local_t(const Plan& p, int cls, uint32_t first, uint32_t nsegs, hipStream_t s) {
  switch (cls) {
    case 0:
      k_local<64, 128, COUNT><<<nsegs, 64, 0, s>>>(p, first);
      break;
    case 1:
      k_local<256, 1024, COUNT><<<nsegs, 256, 0, s>>>(p, first);
      break;
    case 2:
      if (p.key_bytes == 32)
        k_local<1024, 6144, COUNT, 4><<<nsegs, 1024, 0, s>>>(p, first);
      else
        k_local<1024, 6144, COUNT, 8><<<nsegs, 1024, 0, s>>>(p, first);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_local(const Plan& p, int cls, uint32_t first, uint32_t nsegs, hipStream_t s,
                        bool count_levels) {
  if (!nsegs) return hipSuccess;
  return count_levels ? launch_local_t<true>(p, cls, first, nsegs, s)
                      : launch_local_t<false>(p, cls, first, nsegs, s);
}

}  // namespace gss
