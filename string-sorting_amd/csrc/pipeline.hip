// The steps either side of the sort, on the device: ingest (a delimited text
// becomes NUL-terminated strings and their offsets), suffix offsets, and the
// --write formatting of a sorted offset array (gfx950, wave64).
//
//   ingest   create_strings_delim (src/sortstring.c:310-335): one string per
//            delimiter byte, starting after the previous delimiter (or at 0);
//            each delimiter becomes the string's NUL (unless it is NUL); bytes
//            after the last delimiter belong to no string.  Two launches over
//            the text: per-chunk delimiter counts, then (after a scan of the
//            counts) every delimiter writes the offset of the string after it.
//   merge    merge_lcp_2way (src/mergesort_lcp.cpp:133-211) as a parallel
//            merge path: tile splits by binary search, then every thread
//            merges its outputs with the LCP rule of the reference.
//   suffixes create_suffixes (src/sortstring.c:349-358): offset i for every
//            text position i.
//   write    write_result (src/sortstring.c:360-391): the strings in array
//            order, each followed by '\n'.  Lengths, a 64-bit scan of
//            length+1, then every 16-byte block of the output is assembled
//            in registers from the strings overlapping it and stored once.

#include <mutex>
#include <string>
#include <vector>

#include "gss_internal.h"
#include "host.h"

namespace gss {

namespace {

constexpr int kIngThreads = 256;
constexpr int kIngWords = 8;  // 8-byte words per thread
constexpr uint64_t kIngChunk = (uint64_t)kIngThreads * kIngWords * 8;  // 16 KiB per block

constexpr int kScanThreads64 = 256;
constexpr int kScanItems64 = 16;
constexpr uint64_t kScanBlock64 = (uint64_t)kScanThreads64 * kScanItems64;

// Exact per-byte equality flags: bit 8b+7 set when byte b of x equals byte
// b of `pattern` (no false positives: the additions never carry).
__device__ __forceinline__ uint64_t eq_bytes(uint64_t x, uint64_t pattern) {
  const uint64_t v = x ^ pattern;
  const uint64_t t = ((v & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | v;
  return ~t & 0x8080808080808080ull;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t t = __shfl_up(v, o, kWave);
    if ((int)lane_id() >= o) v += t;
  }
  return v;
}

// Block-wide exclusive scan of 64-bit values; scratch holds BLOCK/64 + 1.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* scratch,
                                                      uint64_t* total) {
  constexpr int W = BLOCK / kWave;
  const uint64_t inc = wave_incl_scan64(v);
  const int wid = threadIdx.x / kWave;
  if ((int)lane_id() == kWave - 1) scratch[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    const uint64_t t = (int)lane_id() < W ? scratch[lane_id()] : 0ull;
    const uint64_t ti = wave_incl_scan64(t);
    if ((int)lane_id() < W) scratch[lane_id()] = ti - t;
    if ((int)lane_id() == W - 1) scratch[W] = ti;
  }
  __syncthreads();
  const uint64_t r = scratch[wid] + inc - v;
  *total = scratch[W];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Ingest

// Delimiter flags of the 8 words [w0, w0+8) (one thread's 64 contiguous
// bytes, loaded 16 bytes at a time); bytes at or beyond len never match.
__device__ __forceinline__ void ingest_flags(const uint64_t* __restrict__ t64, uint64_t len,
                                             uint64_t pat, uint64_t w0, uint64_t (&f)[kIngWords]) {
  uint64_t raw[kIngWords];
  if (w0 * 8 >= len) {
#pragma unroll
    for (int q = 0; q < kIngWords; ++q) f[q] = 0;
    return;
  }
  text_words<kIngWords>(t64, len, w0, raw);
#pragma unroll
  for (int q = 0; q < kIngWords; ++q) {
    const uint64_t b = (w0 + q) * 8;
    uint64_t m = b < len ? eq_bytes(raw[q], pat) : 0ull;
    if (b < len && len - b < 8) m &= (1ull << (8 * (len - b))) - 1ull;
    f[q] = m;
  }
}

// Block b covers the 16 KiB chunk of words [b*2048, (b+1)*2048); thread t
// owns its words [8t, 8t+8), so a thread's delimiters are in text order and
// one block scan ranks them.
__global__ __launch_bounds__(kIngThreads) void k_ingest_count(const uint64_t* __restrict__ t64,
                                                             uint64_t len, uint64_t pat,
                                                             uint64_t* __restrict__ counts) {
  __shared__ uint64_t scratch[kIngThreads / kWave + 1];
  const uint64_t w0 = (uint64_t)blockIdx.x * (kIngChunk / 8) + (uint64_t)threadIdx.x * kIngWords;
  uint64_t f[kIngWords];
  ingest_flags(t64, len, pat, w0, f);
  uint64_t c = 0;
#pragma unroll
  for (int q = 0; q < kIngWords; ++q) c += __popcll(f[q]);
  uint64_t total;
  block_excl_scan64<kIngThreads>(c, scratch, &total);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Every delimiter j at byte i writes off[j+1] = i+1 (j+1 < n) and, unless
// the delimiter is NUL, becomes a NUL.  Block bases come from the scanned
// counts, ranks inside the block from one scan of per-thread counts.  n is
// the scanned total (*total): when it is 0 or exceeds cap nothing is
// written, the text included.
__global__ __launch_bounds__(kIngThreads) void k_ingest_emit(uint64_t* __restrict__ t64,
                                                            uint64_t len, uint64_t pat,
                                                            bool clear, const uint64_t* __restrict__ base,
                                                            const uint64_t* __restrict__ total,
                                                            uint64_t cap, uint32_t* __restrict__ off) {
  const uint64_t n = *total;
  if (n == 0 || n > cap) return;
  __shared__ uint64_t scratch[kIngThreads / kWave + 1];
  const uint64_t w0 = (uint64_t)blockIdx.x * (kIngChunk / 8) + (uint64_t)threadIdx.x * kIngWords;
  uint64_t f[kIngWords];
  ingest_flags(t64, len, pat, w0, f);
  uint64_t c = 0;
#pragma unroll
  for (int q = 0; q < kIngWords; ++q) c += __popcll(f[q]);
  uint64_t tot;
  uint64_t j = base[blockIdx.x] + block_excl_scan64<kIngThreads>(c, scratch, &tot);
  if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0) off[0] = 0;
#pragma unroll
  for (int q = 0; q < kIngWords; ++q) {
    const uint64_t w = w0 + q;
    for (uint64_t m = f[q]; m; m &= m - 1) {
      const uint64_t i = w * 8 + (__builtin_ctzll(m) >> 3);
      if (j + 1 < n) off[j + 1] = (uint32_t)(i + 1);
      ++j;
    }
    if (clear && f[q]) {  // delimiters -> NUL
      if (w * 8 + 8 <= len) {
        t64[w] &= ~((f[q] >> 7) * 0xFFull);
      } else {  // the text's last, partial word: touch only its own bytes
        unsigned char* t8 = reinterpret_cast<unsigned char*>(t64);
        for (uint64_t mm = f[q]; mm; mm &= mm - 1) t8[w * 8 + (__builtin_ctzll(mm) >> 3)] = 0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 64-bit scan of per-item values (reduce, scan of block sums, down-sweep).
// Values: counts[i] (ingest) or len[i] + 1 (write).

template <bool PLUS_ONE, typename T>
__global__ __launch_bounds__(kScanThreads64) void k_scan64_reduce(const T* __restrict__ in,
                                                                uint64_t n,
                                                                uint64_t* __restrict__ bsum) {
  __shared__ uint64_t scratch[kScanThreads64 / kWave + 1];
  const uint64_t b0 = (uint64_t)blockIdx.x * kScanBlock64;
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems64; ++q) {
    const uint64_t i = b0 + (uint64_t)q * kScanThreads64 + threadIdx.x;
    if (i < n) s += (uint64_t)in[i] + (PLUS_ONE ? 1u : 0u);
  }
  uint64_t total;
  block_excl_scan64<kScanThreads64>(s, scratch, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan64_blocks(uint64_t* __restrict__ bsum, uint64_t nb,
                                                       uint64_t* __restrict__ total_out) {
  __shared__ uint64_t scratch[1024 / kWave + 1];
  const uint64_t per = (nb + 1023) / 1024;
  const uint64_t a = threadIdx.x * per;
  uint64_t s = 0;
  for (uint64_t i = a; i < a + per && i < nb; ++i) s += bsum[i];
  uint64_t total;
  uint64_t r = block_excl_scan64<1024>(s, scratch, &total);
  for (uint64_t i = a; i < a + per && i < nb; ++i) {
    const uint64_t v = bsum[i];
    bsum[i] = r;
    r += v;
  }
  if (threadIdx.x == 0) *total_out = total;
}

// out[i] = exclusive prefix of the values; out[n] = total (when WRITE_END).
template <bool PLUS_ONE, typename T>
__global__ __launch_bounds__(kScanThreads64) void k_scan64_down(const T* __restrict__ in,
                                                              uint64_t n,
                                                              const uint64_t* __restrict__ bsum,
                                                              uint64_t* __restrict__ out) {
  __shared__ uint64_t scratch[kScanThreads64 / kWave + 1];
  const uint64_t b0 = (uint64_t)blockIdx.x * kScanBlock64 + (uint64_t)threadIdx.x * kScanItems64;
  uint64_t v[kScanItems64];
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems64; ++q) {
    const uint64_t i = b0 + q;
    v[q] = i < n ? (uint64_t)in[i] + (PLUS_ONE ? 1u : 0u) : 0ull;
    s += v[q];
  }
  uint64_t total;
  uint64_t r = bsum[blockIdx.x] + block_excl_scan64<kScanThreads64>(s, scratch, &total);
#pragma unroll
  for (int q = 0; q < kScanItems64; ++q) {
    const uint64_t i = b0 + q;
    if (i < n) out[i] = r;
    r += v[q];
    if (i + 1 == n) out[n] = r;
  }
}

// ---------------------------------------------------------------------------
// Suffix offsets and --write

__global__ void k_iota32(uint32_t* __restrict__ off, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    off[i] = (uint32_t)i;
}

// Length of every string: bytes before its NUL or the end of the text.
// Reads 64 bytes at a time (four 16-byte loads in flight), which covers most
// strings in one round trip.
__global__ __launch_bounds__(256) void k_strlen(const uint64_t* __restrict__ t64, uint64_t len,
                                               const uint32_t* __restrict__ off, uint64_t n,
                                               uint32_t* __restrict__ lens) {
  const uint64_t nwords = (len + 7) / 8;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t o = off[i];
    uint64_t w = o >> 3;
    uint64_t before = (1ull << (8 * (o & 7))) - 1ull;  // bytes before the string: not NUL
    uint64_t end = len;
    while (w < nwords) {
      uint64_t raw[8];
      text_words<8>(t64, len, w, raw);
      bool found = false;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint64_t b = (w + q) * 8;
        uint64_t x = raw[q];
        if (b + 8 > len) x = b >= len ? 0ull : (x & ((1ull << (8 * (len - b))) - 1ull));
        if (q == 0) x |= before;
        const uint64_t z = eq_bytes(x, 0);
        if (z && !found) {
          end = b + (__builtin_ctzll(z) >> 3);
          found = true;
        }
      }
      if (found) break;
      before = 0;
      w += 8;
    }
    lens[i] = (uint32_t)(end - o);
  }
}

// --write, string-parallel: block b takes sorted strings [256 b, 256 b +
// 256), whose output is one contiguous range [pos[s0], pos[s_end]).  Each
// thread reads its string (72 bytes in flight per load round) and writes
// it, and its newline, into LDS laid out on the output's 16-byte grid; the
// block then stores the range with aligned 16-byte stores (byte stores only
// for the two chunks it shares with its neighbours, whose bytes differ).
// A range over the stage's capacity (long strings) is written straight from
// the threads.  Nothing is written when the output (pos[n]) exceeds cap.
constexpr int kWrThreads = 256;
constexpr uint32_t kWrStage = 32768;
__global__ __launch_bounds__(kWrThreads) void k_write_staged(const uint64_t* __restrict__ t64,
                                                           uint64_t len,
                                                           const uint32_t* __restrict__ off,
                                                           const uint64_t* __restrict__ pos,
                                                           uint64_t n, uint64_t cap,
                                                           unsigned char* __restrict__ out) {
  __shared__ __align__(16) unsigned char s_buf[kWrStage];
  if (pos[n] > cap) return;
  const uint64_t s0 = (uint64_t)blockIdx.x * kWrThreads;
  const uint64_t s_end = s0 + kWrThreads < n ? s0 + kWrThreads : n;
  const uint64_t p0 = pos[s0], p1 = pos[s_end];
  const uint64_t base = p0 & ~15ull;  // LDS byte 0 is output byte `base`
  const bool staged = p1 - base <= kWrStage;
  const uint64_t s = s0 + threadIdx.x;
  if (s < s_end) {
    const uint64_t ps = pos[s], pe = pos[s + 1];  // bytes [ps, pe - 1), '\n' at pe - 1
    const uint64_t slen = pe - 1 - ps;
    const uint64_t o = off[s];
    for (uint64_t k = 0; k < slen; k += 64) {
      uint64_t raw[9];
      text_words<9>(t64, len, (o + k) >> 3, raw);
      const uint32_t sh = (uint32_t)((o + k) & 7) * 8;
      const uint64_t m = slen - k < 64 ? slen - k : 64;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if ((uint64_t)(8 * q) >= m) break;
        const uint64_t w = funnel(raw[q], raw[q + 1], sh);
        const uint64_t dst = ps + k + 8 * q;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          if ((uint64_t)(8 * q + b) >= m) break;
          const unsigned char c = (unsigned char)(w >> (8 * b));
          if (staged)
            s_buf[dst + b - base] = c;
          else
            out[dst + b] = c;
        }
      }
    }
    if (staged)
      s_buf[pe - 1 - base] = '\n';
    else
      out[pe - 1] = '\n';
  }
  if (!staged) return;  // the same for the whole block
  __syncthreads();
  for (uint64_t c = base + (uint64_t)threadIdx.x * 16; c < p1; c += (uint64_t)kWrThreads * 16) {
    if (c >= p0 && c + 16 <= p1) {
      *reinterpret_cast<TextPair*>(out + c) = *reinterpret_cast<const TextPair*>(s_buf + (c - base));
    } else {
      const uint64_t a = c > p0 ? c : p0, e = c + 16 < p1 ? c + 16 : p1;
      for (uint64_t q = a; q < e; ++q) out[q] = s_buf[q - base];
    }
  }
}

// ---------------------------------------------------------------------------
// LCP-aware merge of two sorted runs (merge_lcp_2way, src/mergesort_lcp.cpp:
// 133-211, restated for a parallel merge path).  Ties take A first, so the
// merge of two stable sorts is the stable sort of A followed by B.

constexpr int kMergeThreads = 256;
constexpr int kMergeItems = 8;
static_assert((uint64_t)kMergeThreads * kMergeItems == kMergeTile, "merge tile");

struct StrCmp {
  int sign;      // <0, 0, >0 as strcmp
  uint64_t lcp;  // equal bytes before the first mismatch or NUL
};

// Compares the strings at byte offsets x and y, known equal on their first
// h bytes, as 8-byte keys (the key packing of multikey_cache), four keys of
// each string per step: the loads of a step are issued together, so a
// comparison waits for one memory round trip per 32 bytes.
__device__ __forceinline__ StrCmp str_compare(const uint64_t* __restrict__ t64, uint64_t tlen,
                                              uint64_t x, uint64_t y, uint64_t h) {
  for (uint64_t d = h;; d += 32) {
    uint64_t kx[4], ky[4];
    load_keys<4>(t64, tlen, x + d, kx);
    load_keys<4>(t64, tlen, y + d, ky);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (kx[q] != ky[q]) return StrCmp{kx[q] < ky[q] ? -1 : 1, d + 8 * q + key_lcp(kx[q], ky[q])};
      if ((kx[q] & 0xFFu) == 0) return StrCmp{0, d + 8 * q + key_len(kx[q])};
    }
  }
}

// Smallest i in [lo, hi] with A[i] > B[d-1-i] (ties A first): the number of
// A strings among the first d outputs.  Every string compared is known to
// share its first `depth` bytes with the others.
__device__ __forceinline__ uint64_t merge_path(const uint64_t* __restrict__ t64, uint64_t tlen,
                                               const uint32_t* __restrict__ a,
                                               const uint32_t* __restrict__ b, uint64_t d,
                                               uint64_t lo, uint64_t hi, uint64_t depth = 0) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (str_compare(t64, tlen, a[mid], b[d - 1 - mid], depth).sign <= 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The merges of one launch: up to kMergePairsMax independent 2-way merges
// (a k-way merge's round), each with its own runs, output, split entries
// and tiles; every kernel finds its pair by its split entry or tile.
constexpr int kMergePairsMax = 32;
struct MergePair {
  const uint32_t* a;
  const uint64_t* alcp;
  uint64_t na;
  const uint32_t* b;
  const uint64_t* blcp;
  uint64_t nb;
  uint32_t* out;
  uint64_t* olcp;
  uint64_t split0;  // first split entry (merge_splits(na + nb) of them)
  uint64_t tile0;   // first tile (merge_splits(na + nb) - 1 of them)
};
struct MergeBatch {
  MergePair p[kMergePairsMax];
  uint32_t count;
  uint64_t splits, tiles;  // totals over the pairs
};

// The pair whose range [first(q), first(q) + size(q)) holds g (the first
// pair's range starts at 0 and they follow each other).
template <typename Batch, typename F>
__device__ __forceinline__ uint32_t merge_pair_of(const Batch& mb, uint64_t g, F first) {
  uint32_t q = 0;
  while (q + 1 < mb.count && first(mb.p[q + 1]) <= g) ++q;
  return q;
}

// split[p] = A strings among the first min(p * kMergeTile, na + nb) outputs
// of the pair that owns split entry p.
__global__ __launch_bounds__(256) void k_merge_partition(const uint64_t* __restrict__ t64,
                                                        uint64_t tlen, const MergeBatch mb,
                                                        uint64_t* __restrict__ split) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= mb.splits) return;
  const MergePair& mp = mb.p[merge_pair_of(mb, g, [](const MergePair& x) { return x.split0; })];
  const uint64_t p = g - mp.split0;
  const uint64_t total = mp.na + mp.nb;
  const uint64_t d = p * kMergeTile < total ? p * kMergeTile : total;
  const uint64_t lo = d > mp.nb ? d - mp.nb : 0, hi = d < mp.na ? d : mp.na;
  split[g] = merge_path(t64, tlen, mp.a, mp.b, d, lo, hi);
}

// Each thread merges kMergeItems outputs from its own merge-path split and
// takes one more step to learn the LCP of its last output with the next.
// LCP merge state: s = last output, ha = lcp(s, A[i]), hb = lcp(s, B[j]).
// ha > hb means A[i] < B[j] (and lcp(A[i], B[j]) = hb), ha < hb the
// reverse; only equal values need a comparison, starting at that depth.
//
// Comparisons inside a block go through two pivots, P1 and P2, the block's
// first and last A string (first and last B string when it has none).
// Every staged string e keeps, for each pivot P, its side of P (-1, 0 for
// equal, +1) and d = lcp(e, P), and its 8-byte key at the larger of the
// two d.  Against one pivot, two strings on different sides order by side
// and share min(d) bytes; on the same side the one that leaves P first
// (smaller d) is the farther from P; at the same d their keys at that depth
// decide when both are staged there.  Only a pair that neither pivot
// decides (same side and depth for both, keys equal without a NUL) reads
// the text, from the deepest byte the pivots show equal.  A block that
// spans two groups of strings with long shared prefixes (URL prefix
// groups) has a pivot in each: every string leaves the pivot of its own
// group deep, where one pivot in the other group would see the whole group
// leave it at one depth with one key.  The relations come from the runs'
// LCP arrays: an A string's d to P1 is the minimum of A's LCPs from P1 to
// it (A is sorted: it is above P1 or equal to it), its d to P2 the minimum
// from it to P2 (below P2 or equal); a B string's follow from its LCP y
// with the block's first (last) B string B0 (Bl), the minimum of B's LCPs
// between them, and from (z, c), the LCP and order of P1 and B0 (P2 and
// Bl): y > z puts it where B0 (Bl) is, at depth z; y < z puts it above P1
// (below P2) at depth y; y = z compares its key at z with the pivot's.
__global__ __launch_bounds__(kMergeThreads) void k_merge(const uint64_t* __restrict__ t64,
                                                        uint64_t tlen, const MergeBatch mb,
                                                        const uint64_t* __restrict__ split_all) {
  const MergePair& mp = mb.p[merge_pair_of(mb, blockIdx.x, [](const MergePair& x) { return x.tile0; })];
  const uint64_t tile = blockIdx.x - mp.tile0;
  const uint32_t* __restrict__ a = mp.a;
  const uint64_t* __restrict__ alcp = mp.alcp;
  const uint64_t na = mp.na;
  const uint32_t* __restrict__ b = mp.b;
  const uint64_t* __restrict__ blcp = mp.blcp;
  const uint64_t nb = mp.nb;
  const uint64_t* __restrict__ split = split_all + mp.split0;
  uint32_t* __restrict__ out = mp.out;
  uint64_t* __restrict__ olcp = mp.olcp;
  constexpr uint32_t kInf = 0xFFFFFFFFu;
  constexpr int W = kMergeThreads / kWave;
  // Relation bits: side to P1 + 1 (bits 0-1), side to P2 + 1 (bits 2-3),
  // key staged at d1 (bit 4), at d2 (bit 5).
  constexpr uint32_t kKeyAtD1 = 16, kKeyAtD2 = 32;
  __shared__ uint64_t s_key[kMergeTile];  // A's strings at [0, ia), then B's
  __shared__ uint32_t s_off[kMergeTile];
  __shared__ uint32_t s_d1[kMergeTile];
  __shared__ uint32_t s_d2[kMergeTile];
  __shared__ uint32_t s_prev[kMergeTile];  // LCP with the previous string of its run
  __shared__ uint8_t s_rel[kMergeTile];
  __shared__ uint32_t s_wmin[4][W];
  __shared__ uint64_t s_pz[6];  // z, B end's side, pivot's key at z: for P1 / B0, P2 / Bl
  __shared__ uint64_t s_len2;   // length of P2
  const uint64_t total = na + nb;
  const uint64_t k0 = tile * kMergeTile;
  const uint64_t k1 = k0 + kMergeTile < total ? k0 + kMergeTile : total;
  const uint64_t i0 = split[tile], i1 = split[tile + 1];
  const uint64_t j0 = k0 - i0, j1 = k1 - i1;
  const uint32_t ia = (uint32_t)(i1 - i0), jb = (uint32_t)(j1 - j0);  // ia + jb = k1 - k0
  const uint32_t m = ia + jb;
  const uint32_t p1 = ia ? a[i0] : b[j0];
  const uint32_t p2 = ia ? a[i1 - 1] : b[j1 - 1];
  const uint32_t own_end = ia ? ia : m;  // the pivots' run is items [0, own_end)
  const uint32_t wid = threadIdx.x / kWave, lane = lane_id();

  // The pivots against B's ends and P2's length: three text comparisons per
  // block, by threads of different waves, overlapping the scans.
  if (ia && jb && threadIdx.x == 0) {
    const StrCmp c = str_compare(t64, tlen, p1, b[j0], 0);
    s_pz[0] = c.lcp;
    s_pz[1] = (uint64_t)(int64_t)(-c.sign);
    s_pz[2] = load_key(t64, tlen, (uint64_t)p1 + c.lcp);
  }
  if (ia && jb && threadIdx.x == kWave) {
    const StrCmp c = str_compare(t64, tlen, p2, b[j1 - 1], 0);
    s_pz[3] = c.lcp;
    s_pz[4] = (uint64_t)(int64_t)(-c.sign);
    s_pz[5] = load_key(t64, tlen, (uint64_t)p2 + c.lcp);
  }
  if (threadIdx.x == 2 * kWave) s_len2 = str_compare(t64, tlen, p2, p2, 0).lcp;
  // Blocked items: thread t holds items [t * kMergeItems, (t + 1) *
  // kMergeItems) of [A part | B part].  v = the LCP with the previous item
  // of the same run (kInf at each part's first); each item's LCP with its
  // part's first item is the prefix minimum of v, with its part's last item
  // the minimum of v over the items after it.
  uint32_t v[kMergeItems], o[kMergeItems];
  const uint32_t x0 = threadIdx.x * kMergeItems;
  uint32_t tot_a = kInf, tot_b = kInf;  // minima of this thread's v per part
#pragma unroll
  for (int q = 0; q < kMergeItems; ++q) {
    const uint32_t x = x0 + q;
    v[q] = kInf;
    o[q] = 0;
    if (x < ia) {
      o[q] = a[i0 + x];
      if (x) v[q] = (uint32_t)alcp[i0 + x - 1];
      tot_a = min(tot_a, v[q]);
    } else if (x < m) {
      o[q] = b[j0 + (x - ia)];
      if (x > ia) v[q] = (uint32_t)blcp[j0 + (x - ia) - 1];
      tot_b = min(tot_b, v[q]);
    }
  }
  // Minima of the threads before (pre_*) and after (suf_*) this one, per part.
  uint32_t pre_a, pre_b, suf_a, suf_b;
  {
    uint32_t ea = tot_a, eb = tot_b, fa = tot_a, fb = tot_b;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t ta = __shfl_up(ea, off, kWave), tb = __shfl_up(eb, off, kWave);
      const uint32_t ua = __shfl_down(fa, off, kWave), ub = __shfl_down(fb, off, kWave);
      if ((int)lane >= off) {
        ea = min(ea, ta);
        eb = min(eb, tb);
      }
      if ((int)lane + off < kWave) {
        fa = min(fa, ua);
        fb = min(fb, ub);
      }
    }
    if (lane == kWave - 1) {
      s_wmin[0][wid] = ea;
      s_wmin[1][wid] = eb;
    }
    if (lane == 0) {
      s_wmin[2][wid] = fa;
      s_wmin[3][wid] = fb;
    }
    __syncthreads();
    pre_a = __shfl_up(ea, 1, kWave);
    pre_b = __shfl_up(eb, 1, kWave);
    suf_a = __shfl_down(fa, 1, kWave);
    suf_b = __shfl_down(fb, 1, kWave);
    if (lane == 0) pre_a = pre_b = kInf;
    if (lane == kWave - 1) suf_a = suf_b = kInf;
    for (uint32_t w = 0; w < (uint32_t)W; ++w) {
      if (w < wid) {
        pre_a = min(pre_a, s_wmin[0][w]);
        pre_b = min(pre_b, s_wmin[1][w]);
      } else if (w > wid) {
        suf_a = min(suf_a, s_wmin[2][w]);
        suf_b = min(suf_b, s_wmin[3][w]);
      }
    }
  }
  // LCP of each item with its part's last item (kInf for that item).
  uint32_t to_last[kMergeItems];
  {
    uint32_t ra = suf_a, rb = suf_b;
#pragma unroll
    for (int q = kMergeItems - 1; q >= 0; --q) {
      const uint32_t x = x0 + q;
      to_last[q] = x < ia ? ra : rb;
      if (x < ia)
        ra = min(ra, v[q]);
      else if (x < m)
        rb = min(rb, v[q]);
    }
  }
  const uint32_t len2 = (uint32_t)s_len2;
  const bool cross = ia && jb;
  const uint32_t z1 = cross ? (uint32_t)s_pz[0] : 0, z2 = cross ? (uint32_t)s_pz[3] : 0;
  const int c1 = cross ? (int)(int64_t)s_pz[1] : 0, c2 = cross ? (int)(int64_t)s_pz[4] : 0;
  const uint64_t pk1 = cross ? s_pz[2] : 0, pk2 = cross ? s_pz[5] : 0;
  // A B string's relation to a pivot P from its LCP y with the B end E next
  // to P and (z, c, pk) of P against E; `up` when E is B's first string
  // (the string lies at or above E), else B's last (at or below E).
  auto via_end = [&](uint32_t off, uint32_t y, uint32_t z, int c, uint64_t pk, uint32_t piv,
                     bool up, int* side, uint32_t* d) {
    if (y > z) {
      *side = c;
      *d = z;
    } else if (y < z) {
      *side = up ? 1 : -1;
      *d = y;
    } else {
      const uint64_t kb = load_key(t64, tlen, (uint64_t)off + z);
      if (kb != pk) {
        *side = kb > pk ? 1 : -1;
        *d = z + key_lcp(kb, pk);
      } else if ((kb & 0xFFu) == 0) {
        *side = 0;
        *d = z + key_len(kb);
      } else {
        const StrCmp r = str_compare(t64, tlen, piv, off, (uint64_t)z + 8);
        *side = -r.sign;
        *d = (uint32_t)r.lcp;
      }
    }
  };
  uint32_t run_a = pre_a, run_b = pre_b;
#pragma unroll
  for (int q = 0; q < kMergeItems; ++q) {
    const uint32_t x = x0 + q;
    if (x >= m) continue;
    if (x < ia)
      run_a = min(run_a, v[q]);
    else
      run_b = min(run_b, v[q]);
    const uint32_t to_first = x < ia ? run_a : run_b;  // kInf at the part's first
    int s1, s2;
    uint32_t d1, d2;
    const bool own = x < own_end;
    if (own) {
      d1 = to_first;
      s1 = x == 0 ? 0 : 1;
      d2 = to_last[q];
      s2 = x + 1 == own_end ? 0 : -1;
    } else {
      via_end(o[q], to_first, z1, c1, pk1, p1, true, &s1, &d1);
      via_end(o[q], to_last[q], z2, c2, pk2, p2, false, &s2, &d2);
    }
    // The key at the deeper of the two depths (both when they are equal).
    uint32_t rel = 0;
    uint64_t key = 0;
    const uint32_t kd = d1 == kInf ? d2 : (d2 == kInf ? d1 : max(d1, d2));
    if (kd != kInf) {
      key = load_key(t64, tlen, (uint64_t)o[q] + kd);
      if (kd == d1) rel |= kKeyAtD1;
      if (kd == d2) rel |= kKeyAtD2;
      // A string of the pivots' run that ends where it leaves P1 equals
      // P1 (it is not below it); one that ends where it leaves P2 equals P2
      // when P2 ends there too.
      if (own && (key >> 56) == 0) {
        if (kd == d1 && x != 0) s1 = 0;
        if (kd == d2 && d2 == len2 && x + 1 != own_end) s2 = 0;
      }
    }
    rel |= (uint32_t)(s1 + 1) | ((uint32_t)(s2 + 1) << 2);
    s_off[x] = o[q];
    s_prev[x] = v[q];
    s_d1[x] = d1;
    s_d2[x] = d2;
    s_key[x] = key;
    s_rel[x] = (uint8_t)rel;
  }
  __syncthreads();

  // A[i] against B[j], both staged, through the pivots.  `known` (bytes the
  // two are known to share) only moves the text comparison's start.
  auto cmp_block = [&](uint64_t i, uint64_t j, uint64_t known) -> StrCmp {
    const uint32_t e = (uint32_t)(i - i0), f = ia + (uint32_t)(j - j0);
    const uint32_t re = s_rel[e], rf = s_rel[f];
    uint64_t lower = known;
#pragma unroll
    for (int pv = 0; pv < 2; ++pv) {
      const int se = (int)((re >> (2 * pv)) & 3u) - 1, sf = (int)((rf >> (2 * pv)) & 3u) - 1;
      const uint32_t de = pv ? s_d2[e] : s_d1[e], df = pv ? s_d2[f] : s_d1[f];
      const uint32_t lo = min(de, df);
      if (se != sf) return StrCmp{se < sf ? -1 : 1, lo};
      if (se == 0) return StrCmp{0, lo};  // both equal the pivot
      if (de != df) {
        const bool e_less = se > 0 ? de > df : de < df;
        return StrCmp{e_less ? -1 : 1, lo};
      }
      const uint32_t at = pv ? kKeyAtD2 : kKeyAtD1;
      if ((re & at) && (rf & at)) {
        const uint64_t ke = s_key[e], kf = s_key[f];
        if (ke != kf) return StrCmp{ke < kf ? -1 : 1, de + key_lcp(ke, kf)};
        if ((ke & 0xFFu) == 0) return StrCmp{0, de + key_len(ke)};
        lower = max(lower, (uint64_t)de + 8);
      } else {
        lower = max(lower, (uint64_t)de);
      }
    }
    return str_compare(t64, tlen, s_off[e], s_off[f], lower);
  };
  const uint64_t d = k0 + (uint64_t)threadIdx.x * kMergeItems;
  if (d >= k1) return;
  // Merge-path split of this thread inside the block: i in [lo, hi], with
  // A[mid] and B[d - 1 - mid] inside the block's ranges for every probe.
  uint64_t lo = (d > j1 && d - j1 > i0) ? d - j1 : i0;
  uint64_t hi = d - j0 < i1 ? d - j0 : i1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (cmp_block(mid, d - 1 - mid, 0).sign <= 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  uint64_t i = lo;
  uint64_t j = d - i;
  const uint64_t end = d + kMergeItems < k1 ? d + kMergeItems : k1;
  bool have_s = false;
  uint64_t ha = 0, hb = 0;
  // ha / hb hold lcp(s, head) only while these are set: a head passed over
  // without a comparison (its run's part of the tile is used up) loses it,
  // and the text gives it again if it is ever needed (at the tile's end).
  bool ha_ok = false, hb_ok = false;
  uint32_t s_last = 0;  // offset of the last output
  auto fix_heads = [&]() {
    if (!ha_ok && i < na) ha = str_compare(t64, tlen, s_last, a[i], 0).lcp;
    if (!hb_ok && j < nb) hb = str_compare(t64, tlen, s_last, b[j], 0).lcp;
    ha_ok = hb_ok = true;
  };
  for (uint64_t k = d; k < total && k <= end; ++k) {
    bool take_a;
    uint64_t lprev = 0;  // lcp(previous output, this output)
    // Heads past the block's ranges (only on the extra step) are compared
    // on the text.
    const bool inside = i < i1 && j < j1;
    if (have_s && (i == na || j == nb)) fix_heads();
    if (i == na) {
      take_a = false;
      lprev = hb;
    } else if (j == nb) {
      take_a = true;
      lprev = ha;
    } else if (k < k1 && (i == i1 || j == j1)) {
      // The tile's outputs are A[i0, i1) and B[j0, j1): with one part used
      // up, the next output is the other's head, no comparison needed.  (Its
      // LCP with the last output is known: that head followed an output of
      // its own run or won a comparison against it.)
      take_a = j == j1;
      lprev = take_a ? ha : hb;
      if (take_a) hb_ok = false; else ha_ok = false;
    } else if (!have_s) {
      const StrCmp c = inside ? cmp_block(i, j, 0) : str_compare(t64, tlen, a[i], b[j], 0);
      take_a = c.sign <= 0;
      if (take_a) hb = c.lcp; else ha = c.lcp;
      ha_ok = hb_ok = true;
    } else {
      fix_heads();
      if (ha != hb) {
        take_a = ha > hb;
        lprev = take_a ? ha : hb;
      } else {
        // Both share ha bytes with the last output, hence with each other.
        const StrCmp c = inside ? cmp_block(i, j, ha) : str_compare(t64, tlen, a[i], b[j], ha);
        take_a = c.sign <= 0;
        lprev = ha;
        if (take_a) hb = c.lcp; else ha = c.lcp;
      }
    }
    if (have_s && olcp) olcp[k - 1] = lprev;
    if (k == end) break;  // the extra step: only the LCP above
    // The next head's LCP with this output: staged inside the block (the
    // merge's chain of decisions then waits on no global load).
    if (take_a) {
      s_last = i < i1 ? s_off[i - i0] : a[i];
      ha = i + 1 < i1 ? s_prev[i + 1 - i0] : (i + 1 < na ? alcp[i] : 0);
      ha_ok = true;
      ++i;
    } else {
      s_last = j < j1 ? s_off[ia + (j - j0)] : b[j];
      hb = j + 1 < j1 ? s_prev[ia + (j + 1 - j0)] : (j + 1 < nb ? blcp[j] : 0);
      hb_ok = true;
      ++j;
    }
    out[k] = s_last;
    have_s = true;
  }
}

uint32_t grid_cap(uint64_t items, uint32_t block, uint32_t cap = 65536) {
  uint64_t g = (items + block - 1) / block;
  if (g < 1) g = 1;
  return (uint32_t)(g > cap ? cap : g);
}

// ---------------------------------------------------------------------------
// K-way merge by 8-byte key groups.  In a sorted run the strings that share
// their first 8 bytes, none of them NUL, lie together, and the run's LCP
// array marks them: a group starts where the LCP with the previous string is
// below 8.  Only a group's first string is read, for its key (the 8 bytes
// big-endian, zero from a NUL on).  The groups of all runs merge by key as
// integers, stably (on equal keys the earlier run's group first); then each
// group's strings go out as they lie in their run, with the run's LCPs
// inside the group and the LCP of the two keys across groups.  Equal keys
// of groups from different runs (a tie) belong to strings that share 8
// bytes: each tie's span of the output is sorted in place from byte 8, by a
// stable sort (equal strings keep run order), and the rest is final.  Equal
// keys holding a NUL are equal strings, final in run order.  Strings are
// read where the merge needs them: once per group, and again only inside a
// tie, where 2-way rounds would compare every string log2(k) times.

// A run of a k-way merge: offsets, LCP array (len - 1 entries), length,
// and its first string's index in the concatenation of the runs.
struct KwRun {
  const uint32_t* off;
  const uint64_t* lcp;
  uint64_t len;
  uint64_t base;
};
constexpr uint32_t kKwRunsMax = 64;  // the run table sits in LDS
// The run table, passed by value (kernel arguments) and staged in LDS.
struct KwRuns {
  KwRun r[kKwRunsMax];
};
// More than n / kKwSortRatio groups: the merge sorts the runs side by side.
constexpr uint64_t kKwSortRatio = 8;
constexpr int kKwThreads = 256;
constexpr int kKwItems = 16;
constexpr uint64_t kKwBlock = (uint64_t)kKwThreads * kKwItems;  // strings per block

// The run holding concatenated index g: the last r with base[r] <= g (the
// runs are nonempty, so their bases ascend).
__device__ __forceinline__ uint32_t kw_run_of(const uint64_t* s_base, uint32_t k, uint64_t g) {
  uint32_t lo = 0, hi = k - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (s_base[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Group starts of the concatenated runs, block b covering strings
// [b * kKwBlock, (b + 1) * kKwBlock), unless the sample (*sampled of nsamp
// strings start groups) shows many groups.  Without EMIT: bsum[b] = the
// block's group count.  With EMIT (bsum scanned): group h (in concatenated order)
// gets hgid[h] = its first string's index, hkey[h] = its key; runhead[r] =
// run r's first group; hgid[H] = n and runhead[k] = H for the H groups.
template <bool EMIT>
__global__ __launch_bounds__(kKwThreads) void k_kw_heads(
    const KwRuns runs, uint32_t k, uint64_t n, const uint64_t* __restrict__ t64,
    uint64_t tlen, uint64_t* __restrict__ bsum, uint32_t* __restrict__ hgid,
    uint64_t* __restrict__ hkey, uint32_t* __restrict__ runhead,
    const uint64_t* __restrict__ sampled, uint64_t nsamp) {
  if (*sampled * kKwSortRatio > nsamp) return;  // many groups: the runs are sorted instead
  __shared__ KwRun s_runs[kKwRunsMax];
  __shared__ uint64_t s_base[kKwRunsMax + 1];
  __shared__ uint8_t s_flag[kKwBlock];
  __shared__ uint64_t scratch[kKwThreads / kWave + 1];
  for (uint32_t r = threadIdx.x; r < k; r += kKwThreads) {
    s_runs[r] = runs.r[r];
    s_base[r] = runs.r[r].base;
  }
  if (threadIdx.x == 0) s_base[k] = n;
  __syncthreads();
  const uint64_t b0 = (uint64_t)blockIdx.x * kKwBlock;
  // Flags in LDS from coalesced reads of the LCP arrays.
#pragma unroll 4
  for (int q = 0; q < kKwItems; ++q) {
    const uint64_t x = (uint64_t)q * kKwThreads + threadIdx.x;
    const uint64_t g = b0 + x;
    uint8_t f = 0;
    if (g < n) {
      const uint32_t r = kw_run_of(s_base, k, g);
      const uint64_t i = g - s_base[r];
      f = (i == 0 || s_runs[r].lcp[i - 1] < 8) ? 1 : 0;
    }
    s_flag[x] = f;
  }
  __syncthreads();
  // Thread t owns items [t * kKwItems, (t + 1) * kKwItems) of the block.
  const uint32_t x0 = threadIdx.x * kKwItems;
  uint32_t bits = 0;
#pragma unroll
  for (int q = 0; q < kKwItems; ++q) bits |= (uint32_t)s_flag[x0 + q] << q;
  uint64_t total;
  const uint64_t ex = block_excl_scan64<kKwThreads>((uint64_t)__popc(bits), scratch, &total);
  if (!EMIT) {
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
    return;
  }
  uint64_t h = bsum[blockIdx.x] + ex;
  for (int q = 0; q < kKwItems; ++q) {
    const uint64_t g = b0 + x0 + q;
    if (g >= n) break;
    if ((bits >> q) & 1u) {
      const uint32_t r = kw_run_of(s_base, k, g);
      const uint64_t i = g - s_base[r];
      hgid[h] = (uint32_t)g;
      hkey[h] = load_key(t64, tlen, s_runs[r].off[i]);
      if (i == 0) runhead[r] = (uint32_t)h;
      ++h;
    }
    if (g + 1 == n) {
      hgid[h] = (uint32_t)n;
      runhead[k] = (uint32_t)h;
    }
  }
}

// Of kKwSamples evenly spaced strings of the concatenated runs (all of
// them when n is smaller), those that start a group: bsum[b] for block b's
// 1024 samples (the estimate of the group count times kKwSamples / n).
constexpr uint32_t kKwSampleLog = 16;
constexpr uint64_t kKwSamples = 1ull << kKwSampleLog;
__global__ __launch_bounds__(1024) void k_kw_sample(const KwRuns runs, uint32_t k,
                                                   uint64_t n, uint64_t* __restrict__ bsum) {
  __shared__ KwRun s_runs[kKwRunsMax];
  __shared__ uint64_t s_base[kKwRunsMax + 1];
  __shared__ uint64_t scratch[1024 / kWave + 1];
  for (uint32_t r = threadIdx.x; r < k; r += 1024) {
    s_runs[r] = runs.r[r];
    s_base[r] = runs.r[r].base;
  }
  if (threadIdx.x == 0) s_base[k] = n;
  __syncthreads();
  const uint64_t j = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  uint32_t c = 0;
  if (n <= kKwSamples ? j < n : j < kKwSamples) {
    const uint64_t g = n <= kKwSamples ? j : (j * n) >> kKwSampleLog;
    const uint32_t r = kw_run_of(s_base, k, g);
    const uint64_t i = g - s_base[r];
    c = (i == 0 || s_runs[r].lcp[i - 1] < 8) ? 1u : 0u;
  }
  uint64_t total;
  block_excl_scan64<1024>(c, scratch, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// hval[h] = group h's first string index | its size << 32.
__global__ __launch_bounds__(256) void k_kw_pack(const uint32_t* __restrict__ hgid, uint64_t ngroups,
                                                uint64_t* __restrict__ hval) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < ngroups; h += stride)
    hval[h] = (uint64_t)hgid[h] | ((uint64_t)(hgid[h + 1] - hgid[h]) << 32);
}

// The runs' group lists merged in one launch when they are few: group h
// of run r goes to its index in r plus, in every other run, the groups with
// a smaller key (in an earlier run, also those with an equal key: equal
// keys keep run order).  Its value is its first string's index | its size
// << 32, and its output position the strings of those groups and of the
// groups before it in r (hgid of a run's bound group less hgid of its first
// group: hgid[H] = n ends the last run).
__global__ __launch_bounds__(256) void k_kw_rank(const uint64_t* __restrict__ hkey,
                                                const uint32_t* __restrict__ hgid,
                                                const uint32_t* __restrict__ runhead, uint32_t k,
                                                uint64_t ngroups, uint64_t* __restrict__ mkey,
                                                uint64_t* __restrict__ mval,
                                                uint64_t* __restrict__ mpos) {
  __shared__ uint32_t s_rh[kKwRunsMax + 1];
  for (uint32_t r = threadIdx.x; r <= k; r += 256) s_rh[r] = runhead[r];
  __syncthreads();
  const uint64_t h = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (h >= ngroups) return;
  uint32_t lo = 0, hi = k - 1;  // the last run whose first group is at or before h
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (s_rh[mid] <= h)
      lo = mid;
    else
      hi = mid - 1;
  }
  const uint32_t r = lo;
  const uint64_t x = hkey[h];
  const uint32_t gh = hgid[h];
  uint64_t rank = h - s_rh[r];
  uint64_t at = gh - hgid[s_rh[r]];
  for (uint32_t q = 0; q < k; ++q) {
    if (q == r) continue;
    uint32_t a = s_rh[q], e = s_rh[q + 1];
    while (a < e) {
      const uint32_t mid = (a + e) >> 1;
      const uint64_t y = hkey[mid];
      if (q < r ? y <= x : y < x)
        a = mid + 1;
      else
        e = mid;
    }
    rank += a - s_rh[q];
    at += hgid[a] - hgid[s_rh[q]];
  }
  mkey[rank] = x;
  mval[rank] = (uint64_t)gh | ((uint64_t)(hgid[h + 1] - gh) << 32);
  mpos[rank] = at;
  if (h == 0) mpos[ngroups] = hgid[ngroups];
}

// The runs' offsets side by side: cat[g] for every concatenated index g.
__global__ __launch_bounds__(kKwThreads) void k_kw_concat(const KwRuns runs,
                                                         uint32_t k, uint64_t n,
                                                         uint32_t* __restrict__ cat) {
  __shared__ uint64_t s_base[kKwRunsMax + 1];
  __shared__ const uint32_t* s_off[kKwRunsMax];
  for (uint32_t r = threadIdx.x; r < k; r += kKwThreads) {
    s_base[r] = runs.r[r].base;
    s_off[r] = runs.r[r].off;
  }
  if (threadIdx.x == 0) s_base[k] = n;
  __syncthreads();
  const uint64_t b0 = (uint64_t)blockIdx.x * kKwBlock;
  for (int q = 0; q < kKwItems; ++q) {
    const uint64_t g = b0 + (uint64_t)q * kKwThreads + threadIdx.x;
    if (g >= n) break;
    const uint32_t r = kw_run_of(s_base, k, g);
    cat[g] = s_off[r][g - s_base[r]];
  }
}

// Stable 2-way merges of (key, value) lists by key, several per launch:
// merge-path splits every kKwmTile outputs, then one block per tile merges
// its share through LDS.
constexpr int kKwmThreads = 256;
constexpr int kKwmItems = 8;
constexpr uint64_t kKwmTile = (uint64_t)kKwmThreads * kKwmItems;
struct KwPair {
  const uint64_t* ak;
  const uint64_t* av;
  uint64_t na;
  const uint64_t* bk;
  const uint64_t* bv;
  uint64_t nb;
  uint64_t* ok;
  uint64_t* ov;
  uint64_t split0;  // first split entry (tiles + 1 of them)
  uint64_t tile0;   // first tile
};
struct KwBatch {
  KwPair p[kMergePairsMax];
  uint32_t count;
  uint64_t splits, tiles;
};

// The number of A items among the first d outputs: the smallest i in
// [lo, hi] with a[i] > b[d - 1 - i] (equal keys: A first).
__device__ __forceinline__ uint64_t kw_path(const uint64_t* __restrict__ a,
                                            const uint64_t* __restrict__ b, uint64_t d,
                                            uint64_t lo, uint64_t hi) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] <= b[d - 1 - mid])
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_kw_mpart(const KwBatch mb, uint64_t* __restrict__ split) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= mb.splits) return;
  const KwPair& mp = mb.p[merge_pair_of(mb, g, [](const KwPair& x) { return x.split0; })];
  const uint64_t p = g - mp.split0;
  const uint64_t total = mp.na + mp.nb;
  const uint64_t d = p * kKwmTile < total ? p * kKwmTile : total;
  const uint64_t lo = d > mp.nb ? d - mp.nb : 0, hi = d < mp.na ? d : mp.na;
  split[g] = kw_path(mp.ak, mp.bk, d, lo, hi);
}

__global__ __launch_bounds__(kKwmThreads) void k_kw_merge(const KwBatch mb,
                                                         const uint64_t* __restrict__ split_all) {
  const KwPair& mp = mb.p[merge_pair_of(mb, blockIdx.x, [](const KwPair& x) { return x.tile0; })];
  const uint64_t tile = blockIdx.x - mp.tile0;
  const uint64_t* __restrict__ split = split_all + mp.split0;
  __shared__ uint64_t s_k[kKwmTile];
  __shared__ uint64_t s_v[kKwmTile];
  const uint64_t total = mp.na + mp.nb;
  const uint64_t k0 = tile * kKwmTile;
  const uint64_t k1 = k0 + kKwmTile < total ? k0 + kKwmTile : total;
  const uint64_t i0 = split[tile], i1 = split[tile + 1];
  const uint64_t j0 = k0 - i0;
  const uint32_t ia = (uint32_t)(i1 - i0), m = (uint32_t)(k1 - k0), jb = m - ia;
  for (uint32_t x = threadIdx.x; x < m; x += kKwmThreads) {
    if (x < ia) {
      s_k[x] = mp.ak[i0 + x];
      s_v[x] = mp.av[i0 + x];
    } else {
      s_k[x] = mp.bk[j0 + (x - ia)];
      s_v[x] = mp.bv[j0 + (x - ia)];
    }
  }
  __syncthreads();
  const uint64_t* sa = s_k;
  const uint64_t* sb = s_k + ia;
  const uint32_t dd = threadIdx.x * kKwmItems;
  uint64_t ok[kKwmItems], ov[kKwmItems];
  if (dd < m) {
    uint32_t ai = (uint32_t)kw_path(sa, sb, dd, dd > jb ? dd - jb : 0, dd < ia ? dd : ia);
    uint32_t bj = dd - ai;
#pragma unroll
    for (int q = 0; q < kKwmItems; ++q) {
      const bool take_a = ai < ia && (bj >= jb || sa[ai] <= sb[bj]);
      const uint32_t src = take_a ? ai : ia + bj;
      ok[q] = s_k[src < m ? src : 0];
      ov[q] = s_v[src < m ? src : 0];
      ai += take_a ? 1u : 0u;
      bj += take_a ? 0u : 1u;
    }
  }
  __syncthreads();
  if (dd < m) {
#pragma unroll
    for (int q = 0; q < kKwmItems; ++q) {
      if (dd + q < m) {
        s_k[dd + q] = ok[q];
        s_v[dd + q] = ov[q];
      }
    }
  }
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < m; x += kKwmThreads) {
    mp.ok[k0 + x] = s_k[x];
    mp.ov[k0 + x] = s_v[x];
  }
}

// The size of a merged group as a scan input: the value's high word.
struct KwSize {
  uint64_t v;
  __device__ __forceinline__ explicit operator uint64_t() const { return v >> 32; }
};

// first[b] = the merged group whose output range holds position 256 * b.
__global__ __launch_bounds__(256) void k_kw_owner(const uint64_t* __restrict__ pos, uint64_t ngroups,
                                                 uint32_t* __restrict__ first) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < ngroups; m += stride) {
    const uint64_t a = pos[m], e = pos[m + 1];
    for (uint64_t b = (a + 255) / 256; b * 256 < e; ++b) first[b] = (uint32_t)m;
  }
}

// Output position p (block b covers [256 b, 256 b + 256)): its offset from
// its group's run, and its LCP with position p + 1: the run's LCP inside
// the group; across groups the keys' LCP when they differ, the strings'
// length when they are equal with a NUL (equal strings); nothing inside a
// tie (its sort writes those).
__global__ __launch_bounds__(256) void k_kw_expand(const KwRuns runs, uint32_t k,
                                                  uint64_t n, const uint64_t* __restrict__ mkey,
                                                  const uint64_t* __restrict__ mval,
                                                  const uint64_t* __restrict__ pos,
                                                  uint64_t ngroups,
                                                  const uint32_t* __restrict__ first,
                                                  uint32_t* __restrict__ out,
                                                  uint64_t* __restrict__ olcp) {
  __shared__ KwRun s_runs[kKwRunsMax];
  __shared__ uint64_t s_base[kKwRunsMax + 1];
  for (uint32_t r = threadIdx.x; r < k; r += 256) {
    s_runs[r] = runs.r[r];
    s_base[r] = runs.r[r].base;
  }
  if (threadIdx.x == 0) s_base[k] = n;
  __syncthreads();
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const uint64_t nblk = (n + 255) / 256;
  uint64_t lo = first[blockIdx.x];
  uint64_t hi = blockIdx.x + 1 < nblk ? first[blockIdx.x + 1] : ngroups - 1;
  while (lo < hi) {  // the last group m in [lo, hi] with pos[m] <= p
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (pos[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  const uint64_t m = lo;
  const uint64_t v = mval[m];
  const uint64_t g = (v & 0xFFFFFFFFull) + (p - pos[m]);
  const uint32_t r = kw_run_of(s_base, k, g);
  const uint64_t i = g - s_base[r];
  out[p] = s_runs[r].off[i];
  if (olcp && p + 1 < n) {
    const uint64_t a = mkey[m];
    if (p + 1 < pos[m + 1]) {
      // Inside a group: the run's LCP, unless the group is part of a tie.
      const bool tied = (a & 0xFFu) != 0 && ((m > 0 && mkey[m - 1] == a) ||
                                             (m + 1 < ngroups && mkey[m + 1] == a));
      if (!tied) olcp[p] = s_runs[r].lcp[i];
    } else {
      const uint64_t b = mkey[m + 1];
      if (a != b)
        olcp[p] = key_lcp(a, b);
      else if ((a & 0xFFu) == 0)
        olcp[p] = key_len(a);
    }
  }
}

// tflag[m] = 1 where a tie starts: group m's key holds no NUL, the next
// group has the same key and the previous one another.
__global__ __launch_bounds__(256) void k_kw_tieflag(const uint64_t* __restrict__ mkey,
                                                   uint64_t ngroups, uint32_t* __restrict__ tflag) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < ngroups; m += stride) {
    const uint64_t x = mkey[m];
    tflag[m] = ((x & 0xFFu) != 0 && m + 1 < ngroups && mkey[m + 1] == x &&
                (m == 0 || mkey[m - 1] != x))
                   ? 1u
                   : 0u;
  }
}

// tie[tpos[m]] = the output span [start, end) of the tie starting at group
// m, packed start | end << 32 (a tie has one group per run at most, so the
// walk to its end is short).
__global__ __launch_bounds__(256) void k_kw_ties(const uint64_t* __restrict__ mkey,
                                                uint64_t ngroups, const uint32_t* __restrict__ tflag,
                                                const uint64_t* __restrict__ tpos,
                                                const uint64_t* __restrict__ pos,
                                                uint64_t* __restrict__ tie) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < ngroups; m += stride) {
    if (!tflag[m]) continue;
    const uint64_t x = mkey[m];
    uint64_t e = m + 1;
    while (e < ngroups && mkey[e] == x) ++e;
    tie[tpos[m]] = pos[m] | (pos[e] << 32);
  }
}

// The span sort's request table from the T = *ntie tie spans: for tie t,
// request 2t keeps [end of tie t - 1, its start) (possibly empty) and
// request 2t + 1 sorts the tie from byte 8; request 2T keeps the rest.
// req: 2T + 2 bounds, dep: 2T + 1 depths.
__global__ __launch_bounds__(256) void k_kw_reqs(const uint64_t* __restrict__ tie,
                                                const uint64_t* __restrict__ ntie, uint64_t n,
                                                uint64_t* __restrict__ req,
                                                uint32_t* __restrict__ dep) {
  const uint64_t nt = *ntie;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
    req[2 * t] = t ? tie[t - 1] >> 32 : 0;
    dep[2 * t] = kSpanKeep;
    req[2 * t + 1] = tie[t] & 0xFFFFFFFFull;
    dep[2 * t + 1] = 8;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    req[2 * nt] = nt ? tie[nt - 1] >> 32 : 0;
    dep[2 * nt] = kSpanKeep;
    req[2 * nt + 1] = n;
  }
}

// Per-device grow-only scratch and a pinned word for read-backs; calls that
// use it hold its mutex for their whole duration.
struct Scratch {
  std::mutex mu;
  char* dev = nullptr;
  size_t bytes = 0;
  uint64_t* host = nullptr;
  uint64_t* hlist = nullptr;  // pinned, grow-only: lists read back
  size_t hlist_bytes = 0;
};

constexpr int kMaxDev = 16;
Scratch g_scratch[kMaxDev];

#define PL_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(GSS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int scratch_for(Scratch** out) {
  int dev = 0;
  PL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(GSS_E_DEVICE, "device id out of range");
  *out = &g_scratch[dev];
  return GSS_OK;
}

// Ensures `bytes` of device scratch (called with the mutex held).
int scratch_reserve(Scratch& s, size_t bytes, hipStream_t st) {
  if (!s.host) PL_HIP(hipHostMalloc((void**)&s.host, 64, hipHostMallocDefault));
  if (bytes <= s.bytes) return GSS_OK;
  PL_HIP(hipStreamSynchronize(st));
  if (s.dev) (void)hipFree(s.dev);
  s.dev = nullptr;
  s.bytes = 0;
  if (hipMalloc((void**)&s.dev, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(GSS_E_NOMEM, "device scratch of " + std::to_string(bytes) + " bytes");
  }
  s.bytes = bytes;
  return GSS_OK;
}

size_t al256(size_t v) { return (v + 255) / 256 * 256; }

// Ensures `bytes` of the pinned list buffer (called with the mutex held,
// nothing in flight that writes it).
int hlist_reserve(Scratch& s, size_t bytes) {
  if (bytes <= s.hlist_bytes) return GSS_OK;
  if (s.hlist) (void)hipHostFree(s.hlist);
  s.hlist = nullptr;
  s.hlist_bytes = 0;
  const size_t want = std::max<size_t>(bytes, 1 << 16);
  PL_HIP(hipHostMalloc((void**)&s.hlist, want, hipHostMallocDefault));
  s.hlist_bytes = want;
  return GSS_OK;
}

// Exclusive 64-bit scan of values (+1 when PLUS_ONE) into out[0..n]; writes
// the total to *host_total after a stream sync.  bsum: scratch of
// ceil(n / kScanBlock64) + 1 words.
template <bool PLUS_ONE, typename T>
int scan64(const T* in, uint64_t n, uint64_t* bsum, uint64_t* out, uint64_t* d_total,
           uint64_t* host_total, hipStream_t st) {
  const uint64_t nb = (n + kScanBlock64 - 1) / kScanBlock64;
  if (nb) k_scan64_reduce<PLUS_ONE, T><<<(uint32_t)nb, kScanThreads64, 0, st>>>(in, n, bsum);
  k_scan64_blocks<<<1, 1024, 0, st>>>(bsum, nb, d_total);
  if (out && nb) k_scan64_down<PLUS_ONE, T><<<(uint32_t)nb, kScanThreads64, 0, st>>>(in, n, bsum, out);
  PL_HIP(hipGetLastError());
  PL_HIP(hipMemcpyAsync(host_total, d_total, 8, hipMemcpyDeviceToHost, st));
  PL_HIP(hipStreamSynchronize(st));
  return GSS_OK;
}

}  // namespace

}  // namespace gss

using namespace gss;

extern "C" {

int gss_ingest_device32(unsigned char* d_text, uint64_t text_bytes, int delim, uint32_t* d_off,
                        uint64_t max_strings, uint64_t* n_strings, void* stream) {
  if (n_strings) *n_strings = 0;
  if (!d_text || !n_strings) return fail(GSS_E_INVALID, "NULL text or count pointer");
  if (reinterpret_cast<uintptr_t>(d_text) & 7)
    return fail(GSS_E_INVALID, "the device text must be 8-byte aligned");
  if (delim < 0 || delim > 255) return fail(GSS_E_INVALID, "the delimiter must be a byte value");
  if (text_bytes >= (1ull << 32))
    return fail(GSS_E_TOO_LARGE, "text of 4 GiB or more: offsets are 32-bit");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  Scratch* sc = nullptr;
  int rc = scratch_for(&sc);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(sc->mu);
  const uint64_t nblk = (text_bytes + kIngChunk - 1) / kIngChunk;
  const uint64_t nsb = (nblk + kScanBlock64 - 1) / kScanBlock64 + 1;
  rc = scratch_reserve(*sc, al256(8 * (nblk + 1)) * 2 + al256(8 * nsb) + 256, st);
  if (rc) return rc;
  uint64_t* counts = reinterpret_cast<uint64_t*>(sc->dev);
  uint64_t* bases = reinterpret_cast<uint64_t*>(sc->dev + al256(8 * (nblk + 1)));
  uint64_t* bsum = reinterpret_cast<uint64_t*>(sc->dev + 2 * al256(8 * (nblk + 1)));
  uint64_t* d_total = reinterpret_cast<uint64_t*>(sc->dev + 2 * al256(8 * (nblk + 1)) +
                                                  al256(8 * nsb));
  const uint64_t pat = 0x0101010101010101ull * (uint64_t)(uint8_t)delim;
  uint64_t* t64 = reinterpret_cast<uint64_t*>(d_text);
  if (nblk) {
    k_ingest_count<<<(uint32_t)nblk, kIngThreads, 0, st>>>(t64, text_bytes, pat, counts);
    PL_HIP(hipGetLastError());
  }
  if (!d_off) {  // size query
    rc = scan64<false, uint64_t>(counts, nblk, bsum, bases, d_total, sc->host, st);
    if (rc) return rc;
    *n_strings = *sc->host;
    if (!*n_strings) return fail(GSS_E_INVALID, "unable to read any lines from the input");
    return GSS_OK;
  }
  // The offsets with no read-back in between: the emit kernel takes the
  // count from the scan and writes nothing (text included) when it is 0 or
  // over max_strings.
  const uint64_t nsb2 = (nblk + kScanBlock64 - 1) / kScanBlock64;
  if (nblk)
    k_scan64_reduce<false, uint64_t><<<(uint32_t)nsb2, kScanThreads64, 0, st>>>(counts, nblk, bsum);
  k_scan64_blocks<<<1, 1024, 0, st>>>(bsum, nsb2, d_total);
  if (nblk) {
    k_scan64_down<false, uint64_t><<<(uint32_t)nsb2, kScanThreads64, 0, st>>>(counts, nblk, bsum,
                                                                              bases);
    k_ingest_emit<<<(uint32_t)nblk, kIngThreads, 0, st>>>(t64, text_bytes, pat, delim != 0, bases,
                                                          d_total, max_strings, d_off);
  }
  PL_HIP(hipGetLastError());
  PL_HIP(hipMemcpyAsync(sc->host, d_total, 8, hipMemcpyDeviceToHost, st));
  PL_HIP(hipStreamSynchronize(st));
  const uint64_t n = *sc->host;
  *n_strings = n;
  if (n == 0) return fail(GSS_E_INVALID, "unable to read any lines from the input");
  if (n > max_strings)
    return fail(GSS_E_TOO_LARGE, std::to_string(n) + " strings exceed max_strings");
  return GSS_OK;
}

}  // extern "C"

namespace gss {

uint64_t merge_splits(uint64_t total) { return (total + kMergeTile - 1) / kMergeTile + 1; }

// Queues the merges of a batch (their splits in split[0, mb.splits)).
hipError_t merge_batch_launch(const uint64_t* t64, uint64_t tlen, const MergeBatch& mb,
                              uint64_t* split, hipStream_t st) {
  if (!mb.count || !mb.tiles) return hipSuccess;
  k_merge_partition<<<(uint32_t)((mb.splits + 255) / 256), 256, 0, st>>>(t64, tlen, mb, split);
  k_merge<<<(uint32_t)mb.tiles, kMergeThreads, 0, st>>>(t64, tlen, mb, split);
  return hipGetLastError();
}

// Adds a merge to the batch (false when the batch is full).
bool merge_batch_add(MergeBatch& mb, const uint32_t* a, const uint64_t* alcp, uint64_t na,
                     const uint32_t* b, const uint64_t* blcp, uint64_t nb, uint32_t* out,
                     uint64_t* olcp) {
  if (mb.count >= (uint32_t)kMergePairsMax) return false;
  const uint64_t ns = merge_splits(na + nb);
  mb.p[mb.count++] = MergePair{a, alcp, na, b, blcp, nb, out, olcp, mb.splits, mb.tiles};
  mb.splits += ns;
  mb.tiles += ns - 1;
  return true;
}

hipError_t merge_launch(const uint64_t* t64, uint64_t tlen, const uint32_t* a,
                        const uint64_t* alcp, uint64_t na, const uint32_t* b,
                        const uint64_t* blcp, uint64_t nb, uint32_t* out, uint64_t* olcp,
                        uint64_t* split, hipStream_t st) {
  MergeBatch mb{};
  merge_batch_add(mb, a, alcp, na, b, blcp, nb, out, olcp);
  return merge_batch_launch(t64, tlen, mb, split, st);
}

int merge_check(const unsigned char* d_text, uint64_t text_bytes, const uint32_t* d_a,
                const uint64_t* d_a_lcp, uint64_t na, const uint32_t* d_b,
                const uint64_t* d_b_lcp, uint64_t nb, const uint32_t* d_out) {
  const uint64_t total = na + nb;
  if (!d_text || !d_out || (na && !d_a) || (nb && !d_b))
    return fail(GSS_E_INVALID, "NULL text, run or output pointer");
  if ((na > 1 && !d_a_lcp) || (nb > 1 && !d_b_lcp))
    return fail(GSS_E_INVALID, "each run of two or more strings needs its LCP array");
  if (text_bytes == 0) return fail(GSS_E_INVALID, "empty text: every string starts inside it");
  if (reinterpret_cast<uintptr_t>(d_text) & 7)
    return fail(GSS_E_INVALID, "the device text must be 8-byte aligned");
  if (total >= (1ull << 32)) return fail(GSS_E_TOO_LARGE, "at most 2^32-1 strings");
  if (text_bytes >= (1ull << 32))
    return fail(GSS_E_TOO_LARGE, "text of 4 GiB or more: offsets and LCPs are 32-bit here");
  return GSS_OK;
}

// K-way merge by key groups (see k_kw_heads): runs are the nonempty runs in
// run order (2 <= k <= kKwRunsMax), total their strings (< 2^31).
constexpr uint64_t kKwTieFirst = 8192;  // tie spans read back with the count
constexpr uint64_t kKwRankMax = 1ull << 18;  // groups merged by ranks in one launch

int kway_by_keys(const unsigned char* d_text, uint64_t text_bytes, const std::vector<KwRun>& hr,
                 uint64_t total, uint32_t* d_out, uint64_t* d_out_lcp, hipStream_t st) {
  // Every step runs on one stream, the span sort included: a sort given the
  // null stream runs on the library's own (non-blocking) stream, which would
  // not wait for kernels queued on the null stream.  So the caller's null
  // stream is drained first (its runs are ready) and the library's stream
  // takes everything.
  if (!st) {
    PL_HIP(hipStreamSynchronize(nullptr));
    const int lr = library_stream(&st);
    if (lr) return lr;
  }
  const uint32_t k = (uint32_t)hr.size();
  const uint64_t n = total;
  const uint64_t nblk = (n + kKwBlock - 1) / kKwBlock;
  const uint64_t nexp = (n + 255) / 256;
  const uint64_t max_splits = n / kKwmTile + 2ull * k + 2;
  Scratch* sc = nullptr;
  int rc = scratch_for(&sc);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(sc->mu);
  const size_t b_bsum = al256(8 * (nblk + 2)), b_ssum = al256(8 * (kKwSamples / 1024 + 2));
  const size_t b_gid = al256(4 * (n + 1)), b_rh = al256(4ull * (k + 1));
  const size_t b_kv = al256(8 * (n + 2)), b_split = al256(8 * max_splits);
  const size_t b_first = al256(4 * (nexp + 1));
  rc = scratch_reserve(*sc, b_bsum + b_ssum + 256 + b_gid + b_rh + 4 * b_kv + b_split + b_kv +
                                b_first,
                       st);
  if (rc) return rc;
  char* at = sc->dev;
  auto take = [&](size_t bytes) {
    char* q = at;
    at += bytes;
    return q;
  };
  uint64_t* bsum = reinterpret_cast<uint64_t*>(take(b_bsum));
  uint64_t* ssum = reinterpret_cast<uint64_t*>(take(b_ssum));
  uint64_t* d_total = reinterpret_cast<uint64_t*>(take(256));  // [0] groups, [1] sampled starts
  uint32_t* hgid = reinterpret_cast<uint32_t*>(take(b_gid));
  uint32_t* runhead = reinterpret_cast<uint32_t*>(take(b_rh));
  uint64_t* key[2] = {reinterpret_cast<uint64_t*>(take(b_kv)), reinterpret_cast<uint64_t*>(take(b_kv))};
  uint64_t* val[2] = {reinterpret_cast<uint64_t*>(take(b_kv)), reinterpret_cast<uint64_t*>(take(b_kv))};
  uint64_t* split = reinterpret_cast<uint64_t*>(take(b_split));
  uint64_t* pos = reinterpret_cast<uint64_t*>(take(b_kv));
  uint32_t* first = reinterpret_cast<uint32_t*>(take(b_first));
  const uint64_t* t64 = reinterpret_cast<const uint64_t*>(d_text);
  KwRuns runs{};
  for (uint32_t r = 0; r < k; ++r) runs.r[r] = hr[r];

  // 1. Whether the groups are few, from a sample of the LCP arrays, and
  //    (when they are) every group: its first string and key; one read-back
  //    for both.  When they are many (most strings differ from their
  //    neighbours within 8 bytes: short or diverse strings), each would be
  //    read and merged on its own, and a sort of the runs side by side does
  //    less: it gives the same order (stable: equal strings in run order).
  //    The group kernels then exit at once.
  const uint64_t nsamp = std::min<uint64_t>(n, kKwSamples);
  const uint32_t sblk = (uint32_t)((nsamp + 1023) / 1024);
  k_kw_sample<<<sblk, 1024, 0, st>>>(runs, k, n, ssum);
  k_scan64_blocks<<<1, 1024, 0, st>>>(ssum, sblk, d_total + 1);
  k_kw_heads<false><<<(uint32_t)nblk, kKwThreads, 0, st>>>(runs, k, n, t64, text_bytes, bsum,
                                                           nullptr, nullptr, nullptr,
                                                           d_total + 1, nsamp);
  k_scan64_blocks<<<1, 1024, 0, st>>>(bsum, nblk, d_total);
  k_kw_heads<true><<<(uint32_t)nblk, kKwThreads, 0, st>>>(runs, k, n, t64, text_bytes, bsum, hgid,
                                                          key[0], runhead, d_total + 1, nsamp);
  PL_HIP(hipGetLastError());
  PL_HIP(hipMemcpyAsync(sc->host, d_total, 16, hipMemcpyDeviceToHost, st));
  if ((rc = wait_stream(st))) return rc;
  if (sc->host[1] * kKwSortRatio > nsamp) {
    uint32_t* cat = reinterpret_cast<uint32_t*>(key[0]);
    k_kw_concat<<<(uint32_t)nblk, kKwThreads, 0, st>>>(runs, k, n, cat);
    PL_HIP(hipGetLastError());
    return sort_device32(d_text, text_bytes, cat, d_out, d_out_lcp, n, st);
  }
  const uint64_t ngroups = sc->host[0];
  if (ngroups < k || ngroups > n) return fail(GSS_E_INTERNAL, "k-way merge: bad group count");

  // 2. The runs' group lists merged by key, stably: by ranks in one launch
  //    when the groups are few, else pairwise in rounds (adjacent lists, so
  //    equal keys keep run order), each round in one pair of launches per
  //    kMergePairsMax merges.
  int cb = 0;
  if (ngroups <= kKwRankMax) {
    k_kw_rank<<<(uint32_t)((ngroups + 255) / 256), 256, 0, st>>>(key[0], hgid, runhead, k, ngroups,
                                                                 key[1], val[1], pos);
    PL_HIP(hipGetLastError());
    cb = 1;
  } else {
    std::vector<uint32_t> hrh(k + 1);
    PL_HIP(hipMemcpyAsync(hrh.data(), runhead, 4ull * (k + 1), hipMemcpyDeviceToHost, st));
    if ((rc = wait_stream(st))) return rc;
    if (hrh[0] != 0 || hrh[k] != ngroups) return fail(GSS_E_INTERNAL, "k-way merge: bad group lists");
    k_kw_pack<<<grid_cap(ngroups, 256), 256, 0, st>>>(hgid, ngroups, val[0]);
    PL_HIP(hipGetLastError());
    struct List {
      uint64_t at, len;
    };
    std::vector<List> cur;
    for (uint32_t r = 0; r < k; ++r) cur.push_back(List{hrh[r], (uint64_t)hrh[r + 1] - hrh[r]});
    auto launch = [&](KwBatch& mb) -> int {
      if (!mb.count) return GSS_OK;
      k_kw_mpart<<<(uint32_t)((mb.splits + 255) / 256), 256, 0, st>>>(mb, split);
      k_kw_merge<<<(uint32_t)mb.tiles, kKwmThreads, 0, st>>>(mb, split);
      PL_HIP(hipGetLastError());
      mb = KwBatch{};
      return GSS_OK;
    };
    while (cur.size() > 1) {
      std::vector<List> next;
      KwBatch mb{};
      for (size_t r = 0; r < cur.size(); r += 2) {
        const List a = cur[r];
        const List b = r + 1 < cur.size() ? cur[r + 1] : List{a.at + a.len, 0};
        const uint64_t len = a.len + b.len;
        const uint64_t tiles = (len + kKwmTile - 1) / kKwmTile;
        if (mb.count == (uint32_t)kMergePairsMax || mb.splits + tiles + 1 > max_splits)
          if ((rc = launch(mb))) return rc;
        KwPair& q = mb.p[mb.count++];
        q = KwPair{key[cb] + a.at, val[cb] + a.at, a.len, key[cb] + b.at, val[cb] + b.at, b.len,
                   key[cb ^ 1] + a.at, val[cb ^ 1] + a.at, mb.splits, mb.tiles};
        mb.splits += tiles + 1;
        mb.tiles += tiles;
        next.push_back(List{a.at, len});
      }
      if ((rc = launch(mb))) return rc;
      cur.swap(next);
      cb ^= 1;
    }
  }
  const uint64_t* mkey = key[cb];
  const uint64_t* mval = val[cb];

  // 3. Output positions of the merged groups (the rank merge wrote them),
  //    then every string's offset and LCP.
  const uint64_t nsb = (ngroups + kScanBlock64 - 1) / kScanBlock64;
  if (ngroups > kKwRankMax) {
    const KwSize* sizes = reinterpret_cast<const KwSize*>(mval);
    k_scan64_reduce<false, KwSize><<<(uint32_t)nsb, kScanThreads64, 0, st>>>(sizes, ngroups, bsum);
    k_scan64_blocks<<<1, 1024, 0, st>>>(bsum, nsb, d_total);
    k_scan64_down<false, KwSize><<<(uint32_t)nsb, kScanThreads64, 0, st>>>(sizes, ngroups, bsum,
                                                                           pos);
  }
  k_kw_owner<<<grid_cap(ngroups, 256), 256, 0, st>>>(pos, ngroups, first);
  k_kw_expand<<<(uint32_t)nexp, 256, 0, st>>>(runs, k, n, mkey, mval, pos, ngroups, first, d_out,
                                              d_out_lcp);
  PL_HIP(hipGetLastError());

  // 4. Ties: their spans, in output order, read back for the span sort (the
  //    buffers the merge no longer needs hold the flags and spans).
  uint32_t* tflag = hgid;
  uint64_t* tpos = val[cb ^ 1];
  uint64_t* tie = key[cb ^ 1];
  k_kw_tieflag<<<grid_cap(ngroups, 256), 256, 0, st>>>(mkey, ngroups, tflag);
  k_scan64_reduce<false, uint32_t><<<(uint32_t)nsb, kScanThreads64, 0, st>>>(tflag, ngroups, bsum);
  k_scan64_blocks<<<1, 1024, 0, st>>>(bsum, nsb, d_total);
  k_scan64_down<false, uint32_t><<<(uint32_t)nsb, kScanThreads64, 0, st>>>(tflag, ngroups, bsum,
                                                                           tpos);
  k_kw_ties<<<grid_cap(ngroups, 256), 256, 0, st>>>(mkey, ngroups, tflag, tpos, pos, tie);
  PL_HIP(hipGetLastError());
  // 5. The span sort's request table, built on the device (the spans
  //    between ties kept as they are, every tie sorted in place from byte
  //    8) and read back with the tie count (a second copy when there are
  //    more than kKwTieFirst ties).
  uint64_t* dreq = val[cb ^ 1];
  uint32_t* ddep = hgid;
  k_kw_reqs<<<grid_cap(ngroups / 2 + 1, 256), 256, 0, st>>>(tie, d_total, n, dreq, ddep);
  PL_HIP(hipGetLastError());
  auto table_bytes = [](uint64_t t) { return al256(8 * (2 * t + 2)) + 4 * (2 * t + 1); };
  const uint64_t first_n = std::min<uint64_t>(kKwTieFirst, ngroups / 2);
  if ((rc = hlist_reserve(*sc, table_bytes(first_n)))) return rc;
  PL_HIP(hipMemcpyAsync(sc->host, d_total, 8, hipMemcpyDeviceToHost, st));
  PL_HIP(hipMemcpyAsync(sc->hlist, dreq, 8 * (2 * first_n + 2), hipMemcpyDeviceToHost, st));
  PL_HIP(hipMemcpyAsync(reinterpret_cast<char*>(sc->hlist) + al256(8 * (2 * first_n + 2)), ddep,
                        4 * (2 * first_n + 1), hipMemcpyDeviceToHost, st));
  if ((rc = wait_stream(st))) return rc;
  const uint64_t nties = sc->host[0];
  if (nties > ngroups / 2) return fail(GSS_E_INTERNAL, "k-way merge: bad tie count");
  if (nties == 0) return wait_stream(st);
  uint64_t m = first_n;  // ties whose requests the pinned table holds
  if (nties > first_n) {
    m = nties;
    if ((rc = hlist_reserve(*sc, table_bytes(m)))) return rc;
    PL_HIP(hipMemcpyAsync(sc->hlist, dreq, 8 * (2 * m + 2), hipMemcpyDeviceToHost, st));
    PL_HIP(hipMemcpyAsync(reinterpret_cast<char*>(sc->hlist) + al256(8 * (2 * m + 2)), ddep,
                          4 * (2 * m + 1), hipMemcpyDeviceToHost, st));
    if ((rc = wait_stream(st))) return rc;
  }
  const uint64_t* hreq = sc->hlist;
  const uint32_t* hdep =
      reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(sc->hlist) + al256(8 * (2 * m + 2)));
  // The table covers [0, n) with 2 * nties + 1 requests (the span sort
  // checks that the bounds ascend from 0).
  if (hreq[2 * nties + 1] != n) return fail(GSS_E_INTERNAL, "k-way merge: bad tie spans");
  rc = sort_spans_device32(d_text, text_bytes, d_out, d_out_lcp, hreq, hdep,
                           (uint32_t)(2 * nties + 1), st);
  if (rc) return rc;
  return wait_stream(st);
}

}  // namespace gss

extern "C" {

int gss_merge_lcp_device32(const unsigned char* d_text, uint64_t text_bytes, const uint32_t* d_a,
                           const uint64_t* d_a_lcp, uint64_t na, const uint32_t* d_b,
                           const uint64_t* d_b_lcp, uint64_t nb, uint32_t* d_out,
                           uint64_t* d_out_lcp, void* stream) {
  const uint64_t total = na + nb;
  if (total == 0) return GSS_OK;
  const int vrc = merge_check(d_text, text_bytes, d_a, d_a_lcp, na, d_b, d_b_lcp, nb, d_out);
  if (vrc) return vrc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  Scratch* sc = nullptr;
  int rc = scratch_for(&sc);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(sc->mu);
  rc = scratch_reserve(*sc, al256(8 * merge_splits(total)), st);
  if (rc) return rc;
  PL_HIP(merge_launch(reinterpret_cast<const uint64_t*>(d_text), text_bytes, d_a,
                      d_a_lcp, na, d_b, d_b_lcp, nb, d_out, d_out_lcp,
                      reinterpret_cast<uint64_t*>(sc->dev), st));
  PL_HIP(hipStreamSynchronize(st));
  return GSS_OK;
}

int gss_merge_lcp_kway_device32(const unsigned char* d_text, uint64_t text_bytes,
                                const uint32_t* const* d_runs, const uint64_t* const* d_run_lcps,
                                const uint64_t* run_len, uint32_t k, uint32_t* d_out,
                                uint64_t* d_out_lcp, void* stream) {
  if (k && (!d_runs || !d_run_lcps || !run_len))
    return fail(GSS_E_INVALID, "NULL run table");
  uint64_t total = 0;
  for (uint32_t r = 0; r < k; ++r) {
    if (run_len[r] && !d_runs[r]) return fail(GSS_E_INVALID, "NULL run pointer");
    if (run_len[r] > 1 && !d_run_lcps[r])
      return fail(GSS_E_INVALID, "each run of two or more strings needs its LCP array");
    total += run_len[r];
  }
  if (total == 0) return GSS_OK;
  if (!d_text || !d_out) return fail(GSS_E_INVALID, "NULL text or output pointer");
  if (text_bytes == 0) return fail(GSS_E_INVALID, "empty text: every string starts inside it");
  if (reinterpret_cast<uintptr_t>(d_text) & 7)
    return fail(GSS_E_INVALID, "the device text must be 8-byte aligned");
  if (total >= (1ull << 32)) return fail(GSS_E_TOO_LARGE, "at most 2^32-1 strings");
  if (text_bytes >= (1ull << 32))
    return fail(GSS_E_TOO_LARGE, "text of 4 GiB or more: offsets and LCPs are 32-bit here");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // A run in the current round: offsets, LCPs (len - 1 of them), length.
  struct Run {
    const uint32_t* off;
    const uint64_t* lcp;
    uint64_t len;
  };
  std::vector<Run> cur;
  for (uint32_t r = 0; r < k; ++r)
    if (run_len[r]) cur.push_back(Run{d_runs[r], d_run_lcps[r], run_len[r]});
  if (cur.size() == 1) {  // nothing to merge: copy the run
    PL_HIP(hipMemcpyAsync(d_out, cur[0].off, 4 * total, hipMemcpyDeviceToDevice, st));
    if (d_out_lcp && total > 1)
      PL_HIP(hipMemcpyAsync(d_out_lcp, cur[0].lcp, 8 * (total - 1), hipMemcpyDeviceToDevice, st));
    PL_HIP(hipStreamSynchronize(st));
    return GSS_OK;
  }
  // Three or more runs: merged by key groups (one pass over the runs, and
  // strings read once per group and inside ties, where the 2-way rounds
  // would compare every string log2(k) times).
  if (cur.size() >= 3 && cur.size() <= kKwRunsMax && total < (1ull << 31)) {
    std::vector<KwRun> hr;
    uint64_t base = 0;
    for (const Run& r : cur) {
      hr.push_back(KwRun{r.off, r.lcp, r.len, base});
      base += r.len;
    }
    return kway_by_keys(d_text, text_bytes, hr, total, d_out, d_out_lcp, st);
  }
  // Two ping-pong buffers for the intermediate rounds (a run at position p
  // of the merged order keeps its offsets and LCPs at [p, p + len)) and the
  // merge-path splits of one round, in the device's grow-only scratch.
  // Every merge is queued on the stream; one sync at the end.
  Scratch* sc = nullptr;
  int rc = scratch_for(&sc);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(sc->mu);
  const size_t b_off = al256(4 * total), b_lcp = al256(8 * total);
  const size_t b_split = al256(8 * (merge_splits(total) + cur.size() + 2));
  rc = scratch_reserve(*sc, 2 * b_off + 2 * b_lcp + b_split, st);
  if (rc) return rc;
  char* base = sc->dev;
  uint32_t* t_off[2] = {reinterpret_cast<uint32_t*>(base), reinterpret_cast<uint32_t*>(base + b_off)};
  uint64_t* t_lcp[2] = {reinterpret_cast<uint64_t*>(base + 2 * b_off),
                        reinterpret_cast<uint64_t*>(base + 2 * b_off + b_lcp)};
  uint64_t* split = reinterpret_cast<uint64_t*>(base + 2 * b_off + 2 * b_lcp);
  const uint64_t* t64 = reinterpret_cast<const uint64_t*>(d_text);
  const uint64_t tlen = text_bytes;
  // Rounds of pairwise merges of adjacent runs (run 2i with 2i + 1), so
  // that equal strings keep the order of their runs; the last round writes
  // the output.  All merges of a round go out in one pair of launches (up
  // to kMergePairsMax at a time); an odd run is a merge with an empty run,
  // which copies it with its LCPs.  The rounds run in stream order, so one
  // round's buffer is free again two rounds later and the split table after
  // each launch.
  for (int round = 0; cur.size() > 1; ++round) {
    const bool last = cur.size() == 2;
    uint32_t* o_off = last ? d_out : t_off[round & 1];
    uint64_t* o_lcp = last ? d_out_lcp : t_lcp[round & 1];
    std::vector<Run> next;
    uint64_t p = 0;
    MergeBatch mb{};
    for (size_t r = 0; r < cur.size(); r += 2) {
      const Run& x = cur[r];
      const Run y = r + 1 < cur.size() ? cur[r + 1] : Run{nullptr, nullptr, 0};
      const uint64_t len = x.len + y.len;
      if (mb.count == (uint32_t)kMergePairsMax) {
        PL_HIP(merge_batch_launch(t64, tlen, mb, split, st));
        mb = MergeBatch{};
      }
      merge_batch_add(mb, x.off, x.lcp, x.len, y.off, y.lcp, y.len, o_off + p,
                      o_lcp ? o_lcp + p : nullptr);
      next.push_back(Run{o_off + p, o_lcp ? o_lcp + p : nullptr, len});
      p += len;
    }
    PL_HIP(merge_batch_launch(t64, tlen, mb, split, st));
    cur.swap(next);
  }
  PL_HIP(hipStreamSynchronize(st));
  return GSS_OK;
}

int gss_suffixes_device32(uint64_t text_bytes, uint32_t* d_off, void* stream) {
  if (!d_off && text_bytes) return fail(GSS_E_INVALID, "NULL offset array");
  if (text_bytes >= (1ull << 31))
    return fail(GSS_E_TOO_LARGE, "at most 2^31-1 suffixes on the 32-bit path");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (text_bytes) k_iota32<<<grid_cap(text_bytes, 256), 256, 0, st>>>(d_off, text_bytes);
  PL_HIP(hipGetLastError());
  PL_HIP(hipStreamSynchronize(st));
  return GSS_OK;
}

int gss_write_device32(const unsigned char* d_text, uint64_t text_bytes, const uint32_t* d_off,
                       uint64_t n, unsigned char* d_out, uint64_t out_cap, uint64_t* out_bytes,
                       void* stream) {
  if (out_bytes) *out_bytes = 0;
  if (!out_bytes) return fail(GSS_E_INVALID, "NULL out_bytes pointer");
  if (n == 0) return GSS_OK;
  if (!d_text || !d_off) return fail(GSS_E_INVALID, "NULL text or offsets");
  if (text_bytes == 0) return fail(GSS_E_INVALID, "empty text: every string starts inside it");
  if (reinterpret_cast<uintptr_t>(d_text) & 7)
    return fail(GSS_E_INVALID, "the device text must be 8-byte aligned");
  if (d_out && (reinterpret_cast<uintptr_t>(d_out) & 15))
    return fail(GSS_E_INVALID, "the output buffer must be 16-byte aligned");
  if (n >= (1ull << 32)) return fail(GSS_E_TOO_LARGE, "at most 2^32-1 strings");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  Scratch* sc = nullptr;
  int rc = scratch_for(&sc);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(sc->mu);
  const uint64_t nsb = (n + kScanBlock64 - 1) / kScanBlock64 + 1;
  const size_t head = al256(4 * n) + al256(8 * (n + 1)) + al256(8 * nsb) + 256;
  rc = scratch_reserve(*sc, head, st);
  if (rc) return rc;
  uint32_t* lens = reinterpret_cast<uint32_t*>(sc->dev);
  uint64_t* pos = reinterpret_cast<uint64_t*>(sc->dev + al256(4 * n));
  uint64_t* bsum = reinterpret_cast<uint64_t*>(sc->dev + al256(4 * n) + al256(8 * (n + 1)));
  uint64_t* d_total = reinterpret_cast<uint64_t*>(sc->dev + head - 256);
  const uint64_t* t64 = reinterpret_cast<const uint64_t*>(d_text);
  k_strlen<<<grid_cap(n, 256), 256, 0, st>>>(t64, text_bytes, d_off, n, lens);
  PL_HIP(hipGetLastError());
  if (!d_out) {  // size query
    rc = scan64<true, uint32_t>(lens, n, bsum, pos, d_total, sc->host, st);
    if (rc) return rc;
    *out_bytes = *sc->host;
    return GSS_OK;
  }
  // The positions, then the strings, with no read-back in between: the
  // write kernel finds the total (pos[n]) itself and writes nothing when it
  // exceeds out_cap.
  const uint64_t nsb2 = (n + kScanBlock64 - 1) / kScanBlock64;
  k_scan64_reduce<true, uint32_t><<<(uint32_t)nsb2, kScanThreads64, 0, st>>>(lens, n, bsum);
  k_scan64_blocks<<<1, 1024, 0, st>>>(bsum, nsb2, d_total);
  k_scan64_down<true, uint32_t><<<(uint32_t)nsb2, kScanThreads64, 0, st>>>(lens, n, bsum, pos);
  k_write_staged<<<(uint32_t)((n + kWrThreads - 1) / kWrThreads), kWrThreads, 0, st>>>(
      t64, text_bytes, d_off, pos, n, out_cap, d_out);
  PL_HIP(hipGetLastError());
  PL_HIP(hipMemcpyAsync(sc->host, d_total, 8, hipMemcpyDeviceToHost, st));
  PL_HIP(hipStreamSynchronize(st));
  const uint64_t total = *sc->host;
  *out_bytes = total;
  if (total > out_cap)
    return fail(GSS_E_TOO_LARGE, "output of " + std::to_string(total) + " bytes exceeds out_cap");
  return GSS_OK;
}

}  // extern "C"
