// Internal declarations of the MI355X string sort: device records, device
// primitives and the launch wrappers shared by the host orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace gss {

constexpr int kWave = 64;

// Largest segment a single workgroup finishes (the local kernel's capacity).
constexpr uint32_t kLocalCapMax = 6144;

// Global MSD pass geometry: a tile is kTileThreads threads with 4 or 8
// strings each, 2048 or 4096 strings (the tuning's tile); kTile is the
// default.
constexpr int kTileThreads = 512;
constexpr uint32_t kTile = 4096;
constexpr uint32_t kTileMin = 2048;

// Retire ranges are split into chunks of at most this many strings so that a
// huge run of identical strings spreads over many workgroups.
constexpr uint32_t kRetireChunk = 16384;

// A segment still too large for one workgroup: strings [begin, begin+size)
// of index buffer `buf`, all equal on their first `depth` bytes.
struct BigSeg {
  uint32_t begin;
  uint32_t size;
  uint32_t depth;
  uint32_t buf;
};

// One workgroup's share of a global pass over a BigSeg.
struct Tile {
  uint32_t seg;    // index into the pass's BigSeg list
  uint32_t begin;  // absolute position of the first string
  uint32_t len;    // strings in this tile (<= the pass's tile size)
  uint32_t tidx;   // index of this tile inside its segment
};

// Where a BigSeg's digit-major histogram sits in the pass's flat count
// array: entry flat_base + d * ntiles + t counts digit d in tile t, for the
// nbu digit values the segment can take (0 when its keys are all equal).
struct SegAux {
  uint32_t flat_base;
  uint32_t ntiles;
  uint32_t nbu;
  uint32_t tile0;  // index of the segment's first tile in the pass
};

// Kinds of a bucket in Plan::bstart: it goes to the next global pass, its
// order is final, or a workgroup finishes it.
constexpr uint32_t kBucketNext = 0, kBucketFinal = 1, kBucketLocal = 2;

// Entries per block of the flat exclusive scan.
constexpr uint32_t kScanBlock = 4096;

// A segment finished inside one workgroup.
struct LocalSeg {
  uint32_t begin;
  uint32_t size;
  uint32_t depth;
  uint32_t buf;
};

// Strings whose order is final but that still need their LCPs written (and,
// when copy != 0, their offsets copied from index buffer `buf` to the output).
struct Retire {
  uint32_t begin;  // first position of this chunk
  uint32_t len;    // positions in this chunk
  uint32_t last;   // last position of the whole run (no pair starts there)
  uint32_t buf;
  uint32_t lcp;    // LCP of every adjacent pair inside the run
  uint32_t copy;
};

// An adjacent pair (pos, pos+1) split by a global pass at `depth`.
struct Boundary {
  uint32_t pos;
  uint32_t depth;
  uint32_t buf;
  uint32_t pad;
};

// Device-side counters, set by the host at the start of every sort.  A
// later pass's set-up kernel takes its segment count from big_next and sets
// big_cur, ntiles and active_cur, which the pass's other kernels (and the
// host, when it reads them) use; the first pass's come from the host.
struct ClassCount {
  uint32_t segs;     // LocalSegs of the size class
  uint32_t strings;  // their strings
};
struct Counters {
  // Counters reserved together are adjacent, 8-byte aligned pairs, taken
  // with one 64-bit atomic (atomic_add_pair): thousands of workgroups
  // reserving list slots serialise on these few addresses.
  uint32_t big_next;     // BigSegs emitted for the next pass (initially: the first list)
  uint32_t big_next_strings;  // strings of the BigSegs emitted for the next pass
  uint32_t boundary;     // Boundaries emitted so far
  uint32_t retire;       // Retire chunks
  ClassCount local[3];   // LocalSegs per size class and their strings
  uint32_t flat_len;     // entries of the current pass's flat count array
  uint32_t error;        // nonzero: a list overflowed or a loop bound was hit
  uint32_t big_cur;      // BigSegs of the current pass
  uint32_t ntiles;       // tiles of the current pass
  uint32_t bound_done;   // boundaries whose LCPs earlier passes wrote
  uint32_t active_cur;   // strings of the current pass's segments
  uint32_t hist_done;    // the gather counted the pass's digits (set by k_layout)
  uint32_t hist_fused;   // the first pass's digits were counted by its gather
};
static_assert(sizeof(Counters) % 8 == 0, "Counters is copied in 8-byte words");
static_assert(offsetof(Counters, big_next) % 8 == 0 && offsetof(Counters, boundary) % 8 == 0 &&
                  offsetof(Counters, local) % 8 == 0,
              "pairs of counters are 8-byte aligned");

// Adds a to *lo and b to the word after it (an 8-byte aligned pair) in one
// 64-bit atomic; returns the old value of the pair (low word: *lo).  Neither
// word may overflow.
__device__ __forceinline__ uint64_t atomic_add_pair(uint32_t* lo, uint32_t a, uint32_t b) {
  return atomicAdd(reinterpret_cast<unsigned long long*>(lo),
                   ((unsigned long long)b << 32) | (unsigned long long)a);
}

// Strings active in each global pass (the statistics of a sort), written by
// the pass's set-up kernel.
constexpr int kMaxPasses = 64;
static_assert(sizeof(BigSeg) % 8 == 0, "BigSeg is copied in 8-byte words");

// Statistics of the local kernels, spread over kStatSlots slots (block b
// updates slot b % kStatSlots, each slot in its own 128-byte cache line) so
// that thousands of blocks do not serialise on one address; the host
// combines the slots.  v[slot][c] for c < 3: sum over class-c rounds of the
// active strings; v[slot][3]: most rounds one segment needed; v[slot][4 + c]:
// the 8-byte key levels class-c strings advanced through (up to 4 per round).
constexpr int kStatSlots = 64;
struct LocalStats {
  unsigned long long v[kStatSlots][16];
};

// Size classes of the local kernel: (threads, capacity).
constexpr uint32_t kLocalClassCap[3] = {128, 1024, 6144};
constexpr uint32_t kLocalClassThreads[3] = {64, 256, 1024};

// Pointers into the workspace for one sort.
// Words of Plan::guess: the alphabet (8), the first differing key byte, valid.
constexpr int kGuessWords = 10;

struct Plan {
  const uint64_t* text64;
  uint64_t text_bytes;   // text length: bytes at or beyond it read as 0 (>= 1)
  uint32_t* idx[3];      // index buffers: [0] = input, [1], [2] ping-pong
  uint32_t* out;         // sorted offsets
  uint64_t* lcp;         // NULL or n-1 LCPs
  uint64_t* keys;        // n keys of the current global pass
  // A pass that reads its keys from the text also keeps the 8 bytes after
  // each key (`ext`) and its scatter writes every string bound for the next
  // global pass its key there (`keys_next`, at the string's new position):
  // the next pass then takes its keys as they lie (keys_carried != 0) and
  // reads no text.  Passes alternate between the two kinds.
  uint64_t* keys_next;
  uint64_t* ext;
  uint32_t keys_carried;
  BigSeg* big[2];        // BigSeg lists (current, next)
  uint64_t* segmm;       // per BigSeg: min, max key
  uint32_t* bstart;      // per BigSeg and digit: kBucket* kind of the bucket
  uint32_t* alpha;       // 256-bit set of byte values seen in the pass's keys
  Tile* tiles;
  SegAux* aux;           // per BigSeg: layout of its counts in `flat`
  uint16_t* cnt;         // digit-major per-tile counts (a tile's count fits 16 bits)
  uint32_t* flat;        // their exclusive scan, in the same places (+1 total)
  uint32_t* flat_blk;    // block sums of the flat scan
  LocalSeg* local[3];
  Retire* retire;
  Boundary* bounds;
  Counters* ctr;
  LocalStats* lstats;    // [0]: main stream, [1]: side stream
  unsigned long long* pass_active;  // kMaxPasses entries
  uint64_t* tile_mm;     // per tile: min, max key
  uint32_t* tile_alpha;  // per tile: 256-bit byte-presence set
  uint32_t* seg_alpha;   // per BigSeg: 256-bit byte-presence set
  uint32_t local_threshold;
  uint32_t tile;         // strings per tile of a global pass (2048 or 4096)
  uint32_t key_bytes;    // first-round window of the largest local class (32 or 64)
  uint32_t cap_big, cap_retire, cap_bounds, cap_tiles;
  uint32_t cap_local[3];
  // A sort whose first pass has one segment: the gather also counts its
  // tiles' digits as the last such sort's digit would be formed (`guess`:
  // that pass's alphabet, 8 words, its first differing key byte and a valid
  // flag; k_layout checks the guess against the pass's own alphabet and
  // min/max, sets hist_done on a match and stores the new guess).
  uint32_t* guess;
  uint32_t fuse_hist;
};

// ---------------------------------------------------------------------------
// Device primitives

// Aligned text word `w` (little-endian) of a text of tlen >= 1 bytes; bytes
// at or beyond tlen read as 0, also inside the last, partial word, so a
// string that runs off the end is cut there as if a NUL followed.  The
// address is clamped instead of branched around, so a run of these loads
// issues back to back and waits once.
__device__ __forceinline__ uint64_t text_word(const uint64_t* __restrict__ text64,
                                              uint64_t tlen, uint64_t w) {
  const uint64_t nfull = tlen >> 3, last = (tlen - 1) >> 3;
  const uint64_t v = text64[w < last ? w : last];
  if (w < nfull) return v;
  return w == nfull ? v & ((1ull << (8 * (tlen & 7))) - 1ull) : 0ull;
}

// Clears every byte after the first NUL of little-endian `v`; *ended tells
// whether `v` holds a NUL.  The lowest flagged bit of the zero-byte test is
// the top bit of the first zero byte (false positives only sit above it).
__device__ __forceinline__ uint64_t cut_at_nul(uint64_t v, bool* ended) {
  const uint64_t z = (v - 0x0101010101010101ull) & ~v & 0x8080808080808080ull;
  *ended = z != 0;
  return z ? (v & ((z & (0ull - z)) - 1ull)) : v;
}

// Little-endian 8 bytes at byte `sh/8` of the 16 bytes lo:hi (sh in 0..56).
__device__ __forceinline__ uint64_t funnel(uint64_t lo, uint64_t hi, uint32_t sh) {
  return (lo >> sh) | ((hi << 1) << (63 - sh));
}

// Two text words as one 16-byte load (8-byte alignment suffices on gfx950).
struct __attribute__((aligned(8))) TextPair {
  uint64_t lo, hi;
};

// Aligned text words [w, w+NW) of a text of tlen bytes: 16-byte loads when
// all are whole words inside the text, word by word (bytes at or beyond
// tlen read as 0) only near the end of the text.
template <int NW>
__device__ __forceinline__ void text_words(const uint64_t* __restrict__ text64, uint64_t tlen,
                                           uint64_t w, uint64_t (&raw)[NW]) {
  if (w + NW <= (tlen >> 3)) {
#pragma unroll
    for (int q = 0; q + 1 < NW; q += 2) {
      const TextPair pr = *reinterpret_cast<const TextPair*>(text64 + w + q);
      raw[q] = pr.lo;
      raw[q + 1] = pr.hi;
    }
    if (NW & 1) raw[NW - 1] = text64[w + NW - 1];
  } else {
#pragma unroll
    for (int q = 0; q < NW; ++q) raw[q] = text_word(text64, tlen, w + q);
  }
}

// 8 bytes of the text starting at byte `pos`, big-endian (first byte most
// significant), zero after the first NUL; words beyond the text read as 0.
// Same packing as fill_cache in src/multikey_cache.cpp:107-125.
__device__ __forceinline__ uint64_t load_key(const uint64_t* __restrict__ text64,
                                             uint64_t tlen, uint64_t pos) {
  const uint32_t sh = (uint32_t)(pos & 7) * 8;
  uint64_t raw[2];
  text_words<2>(text64, tlen, pos >> 3, raw);
  bool ended;
  return __builtin_bswap64(cut_at_nul(funnel(raw[0], raw[1], sh), &ended));
}

// KW consecutive keys of the string bytes at `pos`: key q covers bytes
// [pos+8q, pos+8q+8); every byte after the first NUL is zero, so the keys of
// a string that ended are 0.
template <int KW>
__device__ __forceinline__ void load_keys(const uint64_t* __restrict__ text64, uint64_t tlen,
                                          uint64_t pos, uint64_t (&out)[KW]) {
  const uint32_t sh = (uint32_t)(pos & 7) * 8;
  uint64_t raw[KW + 1];
  text_words<KW + 1>(text64, tlen, pos >> 3, raw);
  bool done = false;
#pragma unroll
  for (int q = 0; q < KW; ++q) {
    bool e;
    const uint64_t v = cut_at_nul(funnel(raw[q], raw[q + 1], sh), &e);
    out[q] = done ? 0ull : __builtin_bswap64(v);
    done = done || e;
  }
}

// Number of string bytes before the NUL in a key whose low byte is zero.
__device__ __forceinline__ uint32_t key_len(uint64_t k) {
  return k ? 8u - ((uint32_t)__builtin_ctzll(k) >> 3) : 0u;
}

// Leading equal bytes of two different keys.
__device__ __forceinline__ uint32_t key_lcp(uint64_t a, uint64_t b) {
  return (uint32_t)__builtin_clzll(a ^ b) >> 3;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Tile taken by block b of an n-block launch over the pass's tiles.  Blocks
// are dispatched to the eight XCDs in turn (block b on XCD b % 8), each XCD
// with its own L2; giving every XCD a contiguous range of tiles lets
// neighbouring tiles, which share cache lines of the count matrix and of
// each bucket's output, meet in the same L2.  A bijection on [0, n).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  const uint32_t x = b % kXcds, i = b / kXcds, per = n / kXcds, rem = n % kXcds;
  return x * per + (x < rem ? x : rem) + i;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}

// Lanes of the wave whose `bits`-bit digit equals this lane's, among `live`.
template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t live) {
  uint64_t m = live;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint64_t bal = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, kWave);
    if ((int)lane_id() >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, kWave));
  return v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o, kWave));
  return v;
}

__device__ __forceinline__ uint64_t wave_min64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t t = __shfl_xor(v, o, kWave);
    v = t < v ? t : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_max64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t t = __shfl_xor(v, o, kWave);
    v = t > v ? t : v;
  }
  return v;
}

// Block-wide exclusive scan; `scratch` holds BLOCK/64 + 1 words.  Every
// thread of the block must call it.  Returns the exclusive prefix and the
// block total in *total.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* scratch,
                                                    uint32_t* total) {
  constexpr int W = BLOCK / kWave;
  const uint32_t inc = wave_incl_scan(v);
  const int wid = threadIdx.x / kWave;
  if (W == 1) {
    *total = __shfl(inc, kWave - 1, kWave);
    return inc - v;
  }
  if ((int)lane_id() == kWave - 1) scratch[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    uint32_t t = (int)lane_id() < W ? scratch[lane_id()] : 0u;
    const uint32_t ti = wave_incl_scan(t);
    if ((int)lane_id() < W) scratch[lane_id()] = ti - t;
    if ((int)lane_id() == W - 1) scratch[W] = ti;
  }
  __syncthreads();
  const uint32_t r = scratch[wid] + inc - v;
  *total = scratch[W];
  __syncthreads();
  return r;
}

// K block-wide exclusive scans at once (one pair of barriers for all of
// them): v[k] becomes its exclusive prefix, total[k] the block total.
// scratch holds K * (BLOCK/64 + 1) words.  Every thread must call it.
template <int BLOCK, int K>
__device__ __forceinline__ void block_excl_scan_multi(uint32_t (&v)[K], uint32_t* scratch,
                                                      uint32_t (&total)[K]) {
  constexpr int W = BLOCK / kWave;
  const int wid = threadIdx.x / kWave;
  const int lane = (int)lane_id();
  uint32_t inc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) inc[k] = wave_incl_scan(v[k]);
  if (lane == kWave - 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) scratch[k * (W + 1) + wid] = inc[k];
  }
  __syncthreads();
  if (wid == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t t = lane < W ? scratch[k * (W + 1) + lane] : 0u;
      const uint32_t ti = wave_incl_scan(t);
      if (lane < W) scratch[k * (W + 1) + lane] = ti - t;
      if (lane == W - 1) scratch[k * (W + 1) + W] = ti;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = scratch[k * (W + 1) + wid] + inc[k] - v[k];
    total[k] = scratch[k * (W + 1) + W];
  }
  __syncthreads();
}

template <int BLOCK>
__device__ __forceinline__ uint64_t block_or(uint64_t v, uint64_t* scratch) {
  constexpr int W = BLOCK / kWave;
  v = wave_or(v);
  if (W == 1) return v;
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) scratch[wid] = v;
  __syncthreads();
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) r |= scratch[i];
  __syncthreads();
  return r;
}

template <int BLOCK>
__device__ __forceinline__ uint32_t block_max(uint32_t v, uint32_t* scratch) {
  constexpr int W = BLOCK / kWave;
  v = wave_max(v);
  if (W == 1) return v;
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) scratch[wid] = v;
  __syncthreads();
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) r = max(r, scratch[i]);
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Launch wrappers (kernels.hip).  All return hipError_t of the launch.
//
// A global pass needs no host hand-over: launch_plan reads the pass's
// segment count from the device and writes the tile table, and the other
// kernels are launched with `grid` blocks, an upper bound of their work
// (tiles or segments); blocks beyond the device count exit at once.

// grid_tiles / grid_segs: the blocks the pass's tile and segment kernels are
// launched with; a pass that needs more is flagged as an error.
hipError_t launch_plan(const Plan& p, const BigSeg* segs, uint32_t pass, uint32_t grid_tiles,
                       uint32_t grid_segs, hipStream_t s);
hipError_t launch_gather_minmax(const Plan& p, const BigSeg* segs, uint32_t grid,
                                uint32_t radix_bits, hipStream_t s);
hipError_t launch_segreduce(const Plan& p, uint32_t grid, hipStream_t s);
hipError_t launch_layout(const Plan& p, uint32_t radix_bits, hipStream_t s);
hipError_t launch_histogram(const Plan& p, uint32_t grid, uint32_t radix_bits, hipStream_t s);
// Scans flat[0, ctr->flat_len) in place; max_len bounds the launch size.
hipError_t launch_flat_scan(const Plan& p, uint64_t max_len, hipStream_t s);
hipError_t launch_segscan(const Plan& p, const BigSeg* segs, BigSeg* next, uint32_t grid,
                          uint32_t radix_bits, hipStream_t s);
hipError_t launch_scatter(const Plan& p, const BigSeg* segs, uint32_t grid,
                          uint32_t radix_bits, hipStream_t s);
// LCPs of the boundaries the current pass emitted.
hipError_t launch_boundary_lcp(const Plan& p, hipStream_t s);
// Every retire chunk of the counters (grid: workgroups, each looping).
hipError_t launch_retire(const Plan& p, uint32_t grid, hipStream_t s);
// Finishes local segments [first, first + nsegs) of class cls;
// count_levels also counts the 8-byte key levels into the statistics
// (slower: profiling).
hipError_t launch_local(const Plan& p, int cls, uint32_t first, uint32_t nsegs, hipStream_t s,
                        bool count_levels);
// Host-to-device uploads from fine-grained pinned host memory, launched
// together: each range copies nwords 8-byte words from src to dst, the first
// nzero of them cleared instead of read.  A host block must stay unchanged
// until its launch has run.
struct UploadRange {
  const uint64_t* src;
  uint64_t* dst;
  uint32_t nzero, nwords;
};
constexpr int kUploadRanges = 6;
struct UploadBatch {
  UploadRange r[kUploadRanges];
  uint32_t count = 0;
};
// Adds bytes [zero_bytes, bytes) of h_src (to the same offsets of d_dst)
// and clears d_dst's first zero_bytes; both must be multiples of 8.
// Returns -1 for a bad size or a full batch.
int add_upload(UploadBatch& b, const void* h_src, void* d_dst, size_t zero_bytes, size_t bytes);
// Launches the batch's ranges (if any) and empties it.
hipError_t launch_upload(UploadBatch& b, hipStream_t s);
// Writes a_bytes from device address a, then min(*count, cap) records of
// rec_bytes from b (b may be null), to the fine-grained pinned block h_dst,
// then seq to *h_flag once all of it is visible to the host.
hipError_t launch_publish(const void* a, size_t a_bytes, const void* b, const uint32_t* count,
                          uint32_t cap, size_t rec_bytes, void* h_dst, uint64_t* h_flag,
                          uint64_t seq, hipStream_t s);
// idx[i] = the position of off[i] in the ascending array starts[0..n).
hipError_t launch_off_to_index(const uint32_t* off, const uint32_t* starts, uint32_t* idx,
                              uint64_t n, hipStream_t s);

// 2-way LCP merge (pipeline.hip).  Outputs per merge tile (one workgroup
// each); merge_splits(total) = ceil(total / kMergeTile) + 1 is the number of
// merge-path split entries a merge of `total` strings uses; merge_launch
// queues the partition and merge kernels on `st` (no sync), with `split` a
// device array of at least merge_splits(na + nb) entries, each written
// before it is read.  merge_check validates the arguments of a merge entry
// point (GSS_OK or the GSS_E_* code, with the message set).
constexpr uint64_t kMergeTile = 2048;
uint64_t merge_splits(uint64_t total);
hipError_t merge_launch(const uint64_t* t64, uint64_t tlen, const uint32_t* a,
                        const uint64_t* alcp, uint64_t na, const uint32_t* b,
                        const uint64_t* blcp, uint64_t nb, uint32_t* out, uint64_t* olcp,
                        uint64_t* split, hipStream_t st);
int merge_check(const unsigned char* d_text, uint64_t text_bytes, const uint32_t* d_a,
                const uint64_t* d_a_lcp, uint64_t na, const uint32_t* d_b,
                const uint64_t* d_b_lcp, uint64_t nb, const uint32_t* d_out);

// One piece of a multi-array copy: `words` 32-bit words from src to dst.
// The served loop gathers its requests' offsets into one batch array and
// scatters the results back with one launch each.
constexpr uint32_t kChunkWords = 16384;
struct CopyChunk {
  const uint32_t* src;
  uint32_t* dst;
  uint64_t words;  // <= kChunkWords
};
// Copies chunks[0, nchunks) (device array), one workgroup per chunk.
hipError_t launch_copy_chunks(const CopyChunk* chunks, uint32_t nchunks, hipStream_t s);

}  // namespace gss
