/*
 * CPU oracle for the MSD radix string sort -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's string-sort semantics, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to compare
 * results with.  Nothing in the product library links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the rantala/string-sorting source tree).
 */
#ifndef GSS_ORACLE_H
#define GSS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-digit MSD radix sort with an O(n) byte oracle and a stable external
 * scatter; insertion sort below 32; bucket 0 (ended strings) is final.
 * Restates msd_CE2, src/msd_ce.cpp:143-178. */
void or_msd_ce2(unsigned char **strings, size_t n);

/* Multikey quicksort over an 8-byte big-endian NUL-padded cache.
 * Restates multikey_cache<8>, src/multikey_cache.cpp:48-235. */
void or_multikey_cache8(unsigned char **strings, size_t n);

/* Stable insertion sort comparing from `depth`, ties keep input order.
 * Restates src/util/insertion_sort.h:29-48. */
void or_insertion_sort(unsigned char **strings, int n, size_t depth);

/* 8 bytes from p[depth..], packed big-endian, stopping after the first NUL
 * and zero-padding the rest.  Restates fill_cache,
 * src/multikey_cache.cpp:107-125 (get_char<uint64_t>, src/util/get_char.h:68-89,
 * without its sign-extension of bytes 4..7). */
uint64_t or_key8(const unsigned char *p, size_t depth);

/* Number of equal bytes before the first mismatch or NUL.
 * Restates lcp(), src/mergesort_lcp.cpp:88-100. */
size_t or_lcp(const unsigned char *a, const unsigned char *b);

/* lcp[i] = or_lcp(s[i], s[i+1]) for i in [0, n-2] (n-1 entries).
 * Restates the array convention of src/mergesort_lcp.cpp:222-226. */
void or_lcp_array(unsigned char *const *strings, size_t n, size_t *lcp);

/* Adjacent-pair result check: counts identical pointers, NULL pointers and
 * strcmp inversions; returns nonzero when any is found.
 * Restates check_result, src/util/debug.h:41-70 (n == 0 and 1 are valid
 * here; the reference underflows n-1 at n == 0). */
int or_check_result(unsigned char *const *strings, size_t n,
		    size_t *identical, size_t *invalid, size_t *wrong);

/* Offset-array front end used by the tests: builds pointers text+off[i],
 * runs algorithm `alg` (0 = or_msd_ce2, 1 = or_multikey_cache8), writes the
 * sorted offsets to off_out and, when lcp_out is not NULL, the n-1 LCPs. */
int or_sort_offsets32(int alg, const unsigned char *text, const uint32_t *off,
		      size_t n, uint32_t *off_out, uint64_t *lcp_out);

/* Splits text[0..len) at every byte equal to `delim`: string j starts after
 * delimiter j-1 (or at 0) and ends at delimiter j, which becomes a NUL
 * unless delim is 0; bytes after the last delimiter belong to no string.
 * Returns the string count and writes the first `max` offsets.
 * Restates create_strings_delim, src/sortstring.c:311-335. */
size_t or_ingest32(unsigned char *text, size_t len, int delim, uint32_t *off,
		   size_t max);

/* The strings text+off[i], in array order, each followed by '\n' (a string
 * ends at its NUL or at byte len).  Returns the output size; writes it to
 * `out` when it fits `cap`.  Restates write_result, src/sortstring.c:361-391
 * (fputs then fputc('\n') per string). */
size_t or_write32(const unsigned char *text, size_t len, const uint32_t *off,
		  size_t n, unsigned char *out, size_t cap);

/* strcmp order and LCP of a and b, known equal on their first `depth` bytes:
 * returns A-B of the first differing or ending byte and its index in *lcp.
 * Restates compare(), src/mergesort_lcp.cpp:102-114. */
int or_compare(const unsigned char *a, const unsigned char *b, size_t depth,
	       size_t *lcp);

/* Merges the sorted runs a (na offsets, alcp: na-1 LCPs) and b into out
 * (na+nb offsets) and olcp (na+nb-1 LCPs, or NULL); ties take a first.
 * Restates merge_lcp_2way, src/mergesort_lcp.cpp:133-211, with LCP arrays
 * of n-1 entries (lcp[i] = lcp(s[i], s[i+1])). */
void or_merge_lcp32(const unsigned char *text, const uint32_t *a,
		    const uint64_t *alcp, size_t na, const uint32_t *b,
		    const uint64_t *blcp, size_t nb, uint32_t *out,
		    uint64_t *olcp);

#ifdef __cplusplus
}
#endif

#endif
