/*
 * gss.h -- C ABI of the MI355X (gfx950) MSD radix string sort.
 *
 * Drop-in boundary for the reference's routine plugin ABI: every entry point
 * takes plain pointers and sizes, has C linkage and needs no torch types.
 * Reference paths are relative to the rantala/string-sorting source tree.
 *
 * Ordering is the reference's: unsigned bytewise lexicographic order with NUL
 * as the terminator (strcmp).  The sort is stable -- equal strings keep their
 * input order -- so the output permutation equals msd_CE2's pointer for
 * pointer (src/msd_ce.cpp:143-178 is stable) and its bytes equal every other
 * reference routine's.
 */
#ifndef GSS_H
#define GSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Drop-in routines, signature of `struct routine::f` (src/routine.h:32-37).
 *
 * gss_msd replaces the msd_CE / msd_ci / msd_A entry points
 * (src/msd_ce.cpp:176-178, src/msd_ci.cpp:156-177, src/msd_a.cpp:111-113):
 * it permutes strings[0..n) in place into sorted order.  The strings are
 * NUL-terminated; they may point into one text buffer (sortstring,
 * src/sortstring.c:310-335), into separate allocations (unit-test/main.cpp:223)
 * or overlap (suffix mode, src/sortstring.c:349-358).  String bytes are never
 * written.  n <= 1 is a no-op.  The ABI has no error channel, so any device
 * failure prints a message on stderr and calls abort(), as the reference's
 * routines do on fatal errors (src/msd_ci.cpp:158-164).
 * The device is initialised lazily on the first call, never at load time.
 */
void gss_msd(unsigned char **strings, size_t n);

/* As gss_msd, and also writes the LCP array: lcp[i] = number of equal bytes
 * of strings[i] and strings[i+1] before the first mismatch or NUL, for
 * i in [0, n-2] (n-1 entries, none written when n <= 1).  Follows lcp() and
 * the array convention of src/mergesort_lcp.cpp:75,88-100,222-226; the
 * reference computes this array inside mergesort_lcp but never returns it. */
void gss_msd_lcp(unsigned char **strings, size_t n, size_t *lcp);

/* How the pointer entry points move the text (no reference counterpart:
 * the reference sorts in place in host memory).  When the strings lie in
 * one readable address range of at most max_text_bytes that is not much
 * larger than their bytes, that range is copied to the device as it is and
 * the pointers become 32-bit distances into it.  Otherwise the strings are
 * packed one after another; a set whose packed bytes exceed max_text_bytes
 * or whose count exceeds max_strings is first split on the host by leading
 * bytes, MSD fashion (src/msd_ce.cpp:151-165), into pieces that each fit
 * one device sort, so gss_msd takes any n and any total size (the
 * reference's size_t n), e.g. a sortstring input of more than 4 GiB.
 * Strings that overlap (suffix sets) must lie within max_text_bytes.
 * Defaults (also set by passing 0): max_text_bytes = 2^32-64 (32-bit
 * offsets), max_strings = 2^31-1; smaller values bound the device memory
 * one call takes.  Returns GSS_E_INVALID outside [64, 2^32-64] x
 * [2, 2^31-1]. */
int gss_set_pointer_limits(uint64_t max_text_bytes, uint64_t max_strings);

/* What the last gss_msd / gss_msd_lcp call of this process did. */
typedef struct gss_pointer_info {
	uint64_t uploaded_text_bytes; /* text bytes moved to the device      */
	uint64_t pieces;              /* device sorts the call was split into */
	uint32_t packed;              /* 1: packed; 0: the address range copied */
	uint32_t pad;
} gss_pointer_info;
int gss_pointer_last_info(gss_pointer_info *out);

/* ------------------------------------------------------------------------
 * Device-resident entry (no reference counterpart: the reference has no
 * device).  The text and the 32-bit string offsets already sit in device
 * memory; this is the path the throughput figures are quoted on.
 *
 *   d_text     device pointer, 8-byte aligned, text_bytes >= 1 bytes (the
 *              last, partial 8-byte word may be read, never written, in
 *              full); every string starts at an offset < text_bytes.  A
 *              string running off the end is cut there, as if a NUL
 *              followed: bytes at or beyond text_bytes are never part of
 *              a string, also when they share the text's last word.
 *   d_off_in   n string offsets into d_text (any order, may repeat).
 *   d_off_out  n offsets, receives the sorted order (may equal d_off_in).
 *   d_lcp      NULL, or n-1 uint64 LCP values as for gss_msd_lcp.
 *   stream     a hipStream_t (NULL = the default stream).
 *
 * Returns 0 on success, a negative GSS_E* code otherwise (gss_last_error()
 * explains).  The call returns once the result is complete on the stream.
 * The global passes are queued back to back (each later pass's tile table
 * is built on the device); the calling thread then polls a pinned word the
 * device writes with the counters, before the local phase and at the end:
 * it busy-waits rather than sleeps.
 */
int gss_sort_device32(const unsigned char *d_text, uint64_t text_bytes,
		      const uint32_t *d_off_in, uint32_t *d_off_out,
		      uint64_t *d_lcp, uint64_t n, void *stream);

/* ------------------------------------------------------------------------
 * Batched sort (no reference counterpart: the reference sorts one input per
 * call).  nreq independent requests share one text and one offset array:
 * request r owns offsets [req_begin[r], req_begin[r+1]) (host array of
 * nreq+1 non-decreasing entries, req_begin[0] = 0, req_begin[nreq] = total
 * strings) and is sorted on its own, exactly as gss_sort_device32 would sort
 * it alone; all requests go through the same kernel launches.  d_lcp, when
 * given, receives each request's LCPs at its positions; the entry between
 * two requests (index req_begin[r+1]-1) is left untouched. */
int gss_sort_batch_device32(const unsigned char *d_text, uint64_t text_bytes,
			    const uint32_t *d_off_in, uint32_t *d_off_out,
			    uint64_t *d_lcp, const uint64_t *req_begin,
			    uint32_t nreq, void *stream);

/* ------------------------------------------------------------------------
 * Sort contexts (no reference counterpart).  The entry points above use one
 * default context per device, so their calls are serialised.  A context owns
 * a stream and a device workspace; calls on different contexts run
 * concurrently (for example from two host threads), which lets a server keep
 * several batches on the GPU at once.  A context belongs to the device that
 * was current when it was created. */
typedef struct gss_ctx gss_ctx;

gss_ctx *gss_ctx_create(void);
void gss_ctx_destroy(gss_ctx *ctx);
void *gss_ctx_stream(gss_ctx *ctx); /* the context's own hipStream_t */
int gss_ctx_sort_batch_device32(gss_ctx *ctx, const unsigned char *d_text,
				uint64_t text_bytes, const uint32_t *d_off_in,
				uint32_t *d_off_out, uint64_t *d_lcp,
				const uint64_t *req_begin, uint32_t nreq,
				void *stream);
/* Sizes the context's device workspace for sorts of up to n strings now, so
 * that later calls allocate nothing until they exceed n. */
int gss_ctx_reserve(gss_ctx *ctx, uint64_t n);

/* ------------------------------------------------------------------------
 * Served request loop (no reference counterpart).  A server is bound to one
 * device-resident text; requests name strings of it by device offset arrays.
 * Requests queue as they arrive; each of `slots` worker slots (a sort context
 * and its batch buffers) repeatedly takes as many queued requests as fit its
 * share of the HBM budget (at most max_batch_requests), gathers their
 * offsets with one kernel, sorts them as one batch, scatters the results
 * back and completes them.  Slots run concurrently, so requests arriving
 * while others are on the GPU start as soon as a slot frees up.  Results are
 * identical to sorting each request alone (gss_sort_device32). */
typedef struct gss_server gss_server;

typedef struct gss_server_config {
	uint64_t hbm_budget;         /* device bytes for all slots' batches */
	uint32_t slots;              /* batches on the GPU at once, 1..8 */
	uint32_t max_batch_requests; /* requests folded into one batch, >= 1 */
} gss_server_config;

typedef struct gss_server_stats {
	uint64_t submitted;
	uint64_t completed;
	uint64_t batches;
	uint64_t strings;           /* strings of completed requests */
	uint64_t batch_capacity;    /* strings one slot's batch may hold */
	uint32_t max_batch_requests_seen;
	double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion */
} gss_server_stats;

/* Creates a server on the current device for the text d_text (device, 8-byte
 * aligned, text_bytes bytes, as for gss_sort_device32), which must stay valid
 * until gss_server_destroy.  The tuning current at creation is kept: every
 * batch of this server uses it, whatever gss_set_tuning does later.  One
 * slot's batch capacity is the largest string count whose buffers and
 * workspace under that tuning, times `slots`, fit hbm_budget.  Returns NULL
 * on error (gss_last_error()). */
gss_server *gss_server_create(const unsigned char *d_text, uint64_t text_bytes,
			      const gss_server_config *cfg);
/* Queues a sort of the n strings at device offsets d_off (into the server's
 * text).  The sorted offsets go to d_out (device, n entries, may equal d_off)
 * and, when d_lcp is not NULL, the n-1 LCPs to d_lcp.  The arrays must stay
 * valid until the request completes.
 *
 * Stream order: ready_event, when not NULL, is a hipEvent_t recorded after
 * the device work that writes d_off (or still reads d_out / d_lcp); the
 * request's device work waits for that work.  The event is captured as
 * recorded at the time of the call, so the caller may record it again
 * right away.  With ready_event NULL the server waits for no other stream,
 * and such work must be complete before the call.  Either way the results
 * are complete on the device when gss_server_wait returns (or
 * gss_server_done reports 1).
 *
 * Returns the request id (>= 0), or a negative GSS_E* code: GSS_E_TOO_LARGE
 * when n exceeds one slot's batch capacity (that request is not admitted;
 * sort it with gss_sort_device32). */
int64_t gss_server_submit(gss_server *s, const uint32_t *d_off, uint32_t *d_out,
			  uint64_t *d_lcp, uint64_t n, void *ready_event);
/* Waits for request `id` and releases its record; *latency_ms (optional) =
 * completion - submission.  Returns the request's status (0 = sorted). */
int gss_server_wait(gss_server *s, int64_t id, double *latency_ms);
/* 1 once request `id` has completed, 0 before, GSS_E_INVALID for an id that
 * is unknown or already waited for. */
int gss_server_done(gss_server *s, int64_t id);
int gss_server_get_stats(gss_server *s, gss_server_stats *out);
/* Finishes every queued request, stops the slots and frees the server. */
void gss_server_destroy(gss_server *s);

/* ------------------------------------------------------------------------
 * The steps either side of the sort, on the device.  All take device
 * pointers, return 0 or a negative GSS_E* code and return once the result is
 * complete on `stream` (NULL = the null stream).  A text is 8-byte aligned;
 * its last, partial 8-byte word may be read (never written) in full.
 *
 * Ingest replaces create_strings_delim (src/sortstring.c:310-335): every
 * byte equal to `delim` ends a string, which starts after the previous
 * delimiter (or at 0); unless delim is 0 each delimiter is overwritten with
 * a NUL.  Bytes after the last delimiter belong to no string.  Writes the
 * string count to *n_strings and the offsets to d_off (max_strings entries;
 * d_off == NULL only counts).  No delimiter at all is an error, as in the
 * reference ("unable to read any lines").  sortstring's plain input uses
 * delim '\n', its --raw input delim 0 (src/sortstring.c:337-347). */
int gss_ingest_device32(unsigned char *d_text, uint64_t text_bytes, int delim,
			uint32_t *d_off, uint64_t max_strings,
			uint64_t *n_strings, void *stream);

/* Suffix mode, create_suffixes (src/sortstring.c:349-358): d_off[i] = i for
 * every text position; the suffixes are then sorted like any strings (each
 * runs to the text's NUL or end). */
int gss_suffixes_device32(uint64_t text_bytes, uint32_t *d_off, void *stream);

/* --write formatting, write_result (src/sortstring.c:360-391): the strings
 * d_off[0..n) in array order, each followed by '\n', into d_out (16-byte
 * aligned, out_cap bytes).  *out_bytes receives the output size; d_out ==
 * NULL only computes it. */
int gss_write_device32(const unsigned char *d_text, uint64_t text_bytes,
		       const uint32_t *d_off, uint64_t n, unsigned char *d_out,
		       uint64_t out_cap, uint64_t *out_bytes, void *stream);

/* LCP-aware merge of two sorted runs of strings of one text, with their LCP
 * arrays (na-1 and nb-1 entries, as written by the sort), replacing
 * merge_lcp_2way (src/mergesort_lcp.cpp:133-211).  Ties take A first, so
 * merging the stable sorts of two inputs gives the stable sort of A followed
 * by B.  d_out receives na+nb offsets, d_out_lcp (optional) their LCP array.
 * The LCPs of the runs decide most comparisons; only strings whose LCP with
 * the last output is equal are compared, from that depth on. */
int gss_merge_lcp_device32(const unsigned char *d_text, uint64_t text_bytes,
			   const uint32_t *d_a, const uint64_t *d_a_lcp, uint64_t na,
			   const uint32_t *d_b, const uint64_t *d_b_lcp, uint64_t nb,
			   uint32_t *d_out, uint64_t *d_out_lcp, void *stream);

/* Allocation-free variant of gss_merge_lcp_device32 (no counterpart in the
 * reference: it is for callers that cannot allocate or block, such as code
 * under HIP graph capture or a tight serving loop).  Same arguments, same
 * output, bit for bit; but it allocates nothing (its one scratch array, the
 * merge-path splits, is a fixed 2 MiB device array of the library, one per
 * GPU, held until the process ends) and does not synchronize: it queues the
 * merge on `stream` and returns, and the result is complete when the
 * stream's work is.  It takes at most 536,868,864 strings (na + nb), beyond
 * that it returns GSS_E_TOO_LARGE.  Calls on one GPU share that GPU's scratch
 * array: calls queued in order on one stream are safe, calls whose work
 * overlaps (other streams, or host threads racing on one stream) race. */
int gss_merge_lcp_device32_noalloc(const unsigned char *d_text, uint64_t text_bytes,
				   const uint32_t *d_a, const uint64_t *d_a_lcp,
				   uint64_t na, const uint32_t *d_b,
				   const uint64_t *d_b_lcp, uint64_t nb,
				   uint32_t *d_out, uint64_t *d_out_lcp, void *stream);

/* K-way LCP merge of sorted runs of one text, replacing merge_lcp_3way
 * (src/mergesort_lcp.cpp:399-709) for any number of runs: run r is
 * d_runs[r] (run_len[r] device offsets) with its LCP array d_run_lcps[r]
 * (run_len[r]-1 entries); the three arrays are host arrays of k entries.
 * Equal strings come out in run order (run 0 first), as merge_lcp_3way's
 * ties do, so merging the stable sorts of k parts gives the stable sort of
 * their concatenation.  Three to 64 runs (under 2^31 strings in all) merge
 * by key groups: each run's strings sharing their first 8 bytes (found from
 * its LCP array) form a group keyed by those bytes, the groups of all runs
 * merge stably as integers, and only the output spans where groups of
 * different runs share a key are sorted, in place, from byte 8 (the batch
 * sort's kernels).  When a sample of the LCP arrays shows more than one
 * group per 8 strings (short or diverse strings), the runs are sorted side
 * by side instead (the same order: the sort is stable).  Two runs, or more
 * than 64, merge pairwise in rounds (a
 * balanced tree of the 2-way merge above, all queued on the stream).  Both
 * use the library's per-device scratch (about 44 and 24 bytes per string),
 * which grows as needed and is kept; the key-group merge also sizes the
 * default context's workspace for the total.  d_out receives the merged
 * offsets, d_out_lcp (optional) their LCP array. */
int gss_merge_lcp_kway_device32(const unsigned char *d_text, uint64_t text_bytes,
				const uint32_t *const *d_runs,
				const uint64_t *const *d_run_lcps,
				const uint64_t *run_len, uint32_t k,
				uint32_t *d_out, uint64_t *d_out_lcp,
				void *stream);

/* ------------------------------------------------------------------------
 * Tuning knobs (no reference counterpart: the reference fixes its
 * thresholds at compile time, e.g. src/msd_ce.cpp:79 n < 32 and :216
 * n >= 0x10000).  Values outside the supported range are rejected with
 * GSS_E_INVALID. */
typedef struct gss_tuning {
	uint32_t local_threshold; /* segments of at most this many strings finish
				     inside one workgroup (64..6144)          */
	uint32_t radix_bits;      /* digit width of a global MSD pass (8 or 11) */
	uint32_t tile;            /* strings per workgroup in a global pass
				     (2048 or 4096)                            */
	uint32_t key_bytes;       /* bytes of every string the largest local
				     class reads and compares at once in a
				     segment's first round (32 or 64)         */
} gss_tuning;

int gss_get_tuning(gss_tuning *out);
int gss_set_tuning(const gss_tuning *in);

/* Autotuner (no reference counterpart).  An input's signature is its mean
 * string length, alphabet size and string count, each as a power-of-two
 * bucket, read from 512 evenly spaced strings.  gss_autotune_device32 sorts
 * a strided sample of the input (at most 2M strings) with the candidate
 * tunings -- local_threshold 1024/2048/4096/6144 x radix_bits 11/8 with the
 * current tile and key_bytes, then the best of those with the other tile
 * size (2048/4096), then with the other key_bytes (32/64): 10 trials when
 * the current tuning is one of them -- and returns the fastest (best_ms;
 * default_ms is the current tuning's time, 0 when it is not a candidate),
 * caching it for the signature.  It does not
 * change the current tuning.  With gss_autotune_enable(1), gss_sort_device32
 * (and the pointer entry points) sort inputs of 65536 strings or more with
 * the cached tuning of their signature, tuning a new signature first. */
typedef struct gss_autotune_result {
	gss_tuning best;
	double best_ms;
	double default_ms;
	uint64_t sample;       /* strings sorted per trial */
	uint32_t signature;
	uint32_t candidates;
} gss_autotune_result;

int gss_autotune_device32(const unsigned char *d_text, uint64_t text_bytes,
			  const uint32_t *d_off, uint64_t n,
			  gss_autotune_result *out, void *stream);
int gss_autotune_enable(int on);
int gss_autotune_reset(void);   /* forgets every cached signature */
int gss_autotune_cached(void);  /* number of cached signatures */

/* ------------------------------------------------------------------------
 * Per-call statistics of the last sort on this thread's library instance
 * (no reference counterpart).  active[r] is the number of strings still
 * unresolved when global MSD pass r began, so the algorithmic bytes of a
 * pass can be recomputed from it. */
#define GSS_MAX_PASSES 64
typedef struct gss_stats {
	uint64_t n;
	uint64_t text_bytes;
	uint32_t global_passes;
	uint32_t local_segments;
	uint64_t active[GSS_MAX_PASSES];
	uint64_t local_strings;  /* strings finished inside one workgroup */
	uint32_t local_rounds_max;  /* most 8-byte key rounds one such segment took */
	uint64_t local_round_strings; /* sum over those rounds of active strings */
	uint64_t local_level_strings; /* sum over strings of the 8-byte key levels
					 the local rounds took them through (a
					 round reads 32 bytes: up to 4 levels);
					 counted only under GSS_PROFILE_LEVELS,
					 else 0 */
	uint32_t first_pass_fused; /* 1: the first global pass's digit counts
				      came from its key gather (its digit was
				      formed as in the last sort on the same
				      context), 0: a separate count */
} gss_stats;

int gss_get_stats(gss_stats *out);            /* last sort on this thread */
int gss_ctx_get_stats(gss_ctx *ctx, gss_stats *out); /* last sort of ctx */

/* Kernel timing: with GSS_PROFILE_EVENTS set, every kernel launch is
 * bracketed by HIP events recorded on the stream the kernel runs on; totals
 * per kernel accumulate until reset.  `name` selects the kernel ("gather",
 * "histogram", "scatter", "local", ...) or a pass ("pass0", ..., "pass9+",
 * "local_phase").  An event between two kernels holds the second one back
 * (about 10 us on the MI355X, while queued kernels follow each other with no
 * gap), so pass times are read with GSS_PROFILE_PASSES alone: only each pass
 * and the local phase are bracketed.  GSS_PROFILE_LEVELS runs the local
 * kernels' counting variant, which fills gss_stats.local_level_strings (and
 * the items of the local kernels' totals) at some cost in kernel time: time
 * with EVENTS or PASSES, count with LEVELS.  0 turns all off. */
#define GSS_PROFILE_EVENTS 1
#define GSS_PROFILE_LEVELS 2
#define GSS_PROFILE_PASSES 4
int gss_profile_enable(int mode);
int gss_profile_reset(void);
int gss_profile_get(const char *name, uint64_t *launches, double *total_ms,
		    uint64_t *items);

const char *gss_last_error(void);

#define GSS_OK 0
#define GSS_E_INVALID (-1)
#define GSS_E_DEVICE (-2)
#define GSS_E_NOMEM (-3)
#define GSS_E_TOO_LARGE (-4)
#define GSS_E_INTERNAL (-5)

#ifdef __cplusplus
}
#endif

#endif /* GSS_H */
