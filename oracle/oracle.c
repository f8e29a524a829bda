/*
 * CPU oracle for the MSD radix string sort -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's string-sort semantics.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it, and
 * only to compare results with; the product library never links or calls it.
 * Reference paths are relative to the rantala/string-sorting source tree.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* src/util/insertion_sort.h:29-48 -- stable: the scan stops at the first
 * predecessor that is <= the element, so equal strings keep input order. */
void
or_insertion_sort(unsigned char **strings, int n, size_t depth)
{
	for (int i = 1; i < n; ++i) {
		unsigned char *cur = strings[i];
		int j = i;
		while (j > 0) {
			const unsigned char *a = strings[j - 1] + depth;
			const unsigned char *b = cur + depth;
			while (*a == *b && *a != 0) {
				++a;
				++b;
			}
			if (*a <= *b)
				break;
			strings[j] = strings[j - 1];
			--j;
		}
		strings[j] = cur;
	}
}

/* src/msd_ce.cpp:143-178 (msd_CE2).  The byte oracle is filled in its own
 * loop, counted, prefix-summed and scattered stably; bucket 0 holds strings
 * that end at `depth` and is final.  Scratch is allocated once for the whole
 * recursion (each call uses the sub-range of its own strings), which changes
 * neither the order nor the result. */
static void
msd_ce2_rec(unsigned char **strings, unsigned char *oracle,
	    unsigned char **sorted, size_t n, size_t depth)
{
	if (n < 32) {
		or_insertion_sort(strings, (int)n, depth);
		return;
	}
	size_t count[256];
	size_t start[256];
	memset(count, 0, sizeof(count));
	for (size_t i = 0; i < n; ++i)
		oracle[i] = strings[i][depth];
	for (size_t i = 0; i < n; ++i)
		++count[oracle[i]];
	size_t run = 0;
	for (int c = 0; c < 256; ++c) {
		start[c] = run;
		run += count[c];
	}
	for (size_t i = 0; i < n; ++i)
		sorted[start[oracle[i]]++] = strings[i];
	memcpy(strings, sorted, n * sizeof(*strings));
	size_t pos = count[0];
	for (int c = 1; c < 256; ++c) {
		if (count[c] == 0)
			continue;
		msd_ce2_rec(strings + pos, oracle + pos, sorted + pos, count[c],
			    depth + 1);
		pos += count[c];
	}
}

void
or_msd_ce2(unsigned char **strings, size_t n)
{
	if (n < 2)
		return;
	unsigned char *oracle = malloc(n);
	unsigned char **sorted = malloc(n * sizeof(*sorted));
	if (!oracle || !sorted)
		abort();
	msd_ce2_rec(strings, oracle, sorted, n, 0);
	free(oracle);
	free(sorted);
}

/* src/multikey_cache.cpp:107-125: big-endian packing that stops after the
 * NUL byte (which packs as zero) and leaves the remaining low bytes zero. */
uint64_t
or_key8(const unsigned char *p, size_t depth)
{
	uint64_t key = 0;
	p += depth;
	for (int i = 0; i < 8; ++i) {
		const uint64_t c = p[i];
		key |= c << (8 * (7 - i));
		if (c == 0)
			break;
	}
	return key;
}

/* src/mergesort_lcp.cpp:88-100 */
size_t
or_lcp(const unsigned char *a, const unsigned char *b)
{
	size_t i = 0;
	while (a[i] != 0 && a[i] == b[i])
		++i;
	return i;
}

/* src/mergesort_lcp.cpp:222-226: lcp[i] relates s[i] and s[i+1]. */
void
or_lcp_array(unsigned char *const *strings, size_t n, size_t *lcp)
{
	for (size_t i = 0; i + 1 < n; ++i)
		lcp[i] = or_lcp(strings[i], strings[i + 1]);
}

/* src/util/debug.h:41-70 */
int
or_check_result(unsigned char *const *strings, size_t n, size_t *identical,
		size_t *invalid, size_t *wrong)
{
	size_t same = 0, bad = 0, inv = 0;
	for (size_t i = 0; i + 1 < n; ++i) {
		if (strings[i] == strings[i + 1])
			++same;
		if (strings[i] == NULL || strings[i + 1] == NULL)
			++bad;
		else if (strcmp((const char *)strings[i],
				(const char *)strings[i + 1]) > 0)
			++inv;
	}
	if (identical)
		*identical = same;
	if (invalid)
		*invalid = bad;
	if (wrong)
		*wrong = inv;
	return (same || bad || inv) ? 1 : 0;
}

/* ---- multikey_cache<8>, src/multikey_cache.cpp:48-235 ------------------- */

struct mk_block {
	uint64_t key;
	unsigned char *ptr;
};

static int
mk_cmp(const struct mk_block *a, const struct mk_block *b)
{
	return (a->key > b->key) - (a->key < b->key);
}

static void
mk_swap(struct mk_block *a, struct mk_block *b)
{
	struct mk_block t = *a;
	*a = *b;
	*b = t;
}

/* Median of three by cached key (src/util/median.h:28-58 semantics). */
static struct mk_block *
mk_med3(struct mk_block *a, struct mk_block *b, struct mk_block *c)
{
	if (mk_cmp(a, b) == 0)
		return a;
	if (mk_cmp(a, c) == 0 || mk_cmp(b, c) == 0)
		return c;
	if (mk_cmp(a, b) < 0) {
		if (mk_cmp(b, c) < 0)
			return b;
		return mk_cmp(a, c) < 0 ? c : a;
	}
	if (mk_cmp(b, c) > 0)
		return b;
	return mk_cmp(a, c) < 0 ? a : c;
}

/* :71-89 -- full-string insertion sort from `depth`, ignoring the cache. */
static void
mk_inssort_strings(struct mk_block *c, size_t n, size_t depth)
{
	for (size_t i = 1; i < n; ++i) {
		unsigned char *cur = c[i].ptr;
		size_t j = i;
		while (j > 0) {
			const unsigned char *a = c[j - 1].ptr + depth;
			const unsigned char *b = cur + depth;
			while (*a == *b && *a != 0) {
				++a;
				++b;
			}
			if (*a <= *b)
				break;
			c[j].ptr = c[j - 1].ptr;
			--j;
		}
		c[j].ptr = cur;
	}
}

/* :91-105 -- insertion sort by cached key only. */
static void
mk_inssort_keys(struct mk_block *c, size_t n)
{
	for (size_t i = 1; i < n; ++i) {
		struct mk_block cur = c[i];
		size_t j = i;
		while (j > 0 && mk_cmp(&c[j - 1], &cur) > 0) {
			c[j] = c[j - 1];
			--j;
		}
		c[j] = cur;
	}
}

static void
mk_rec(struct mk_block *c, size_t n, size_t depth, int dirty)
{
	if (n < 32) {
		/* :130-153 */
		if (n == 0)
			return;
		if (dirty) {
			mk_inssort_strings(c, n, depth);
			return;
		}
		mk_inssort_keys(c, n);
		size_t s = 0;
		for (size_t i = 1; i <= n; ++i) {
			if (i < n && c[i].key == c[s].key)
				continue;
			if (i - s > 1 && (c[s].key & 0xFF))
				mk_inssort_strings(c + s, i - s, depth + 8);
			s = i;
		}
		return;
	}
	if (dirty)
		for (size_t i = 0; i < n; ++i)
			c[i].key = or_key8(c[i].ptr, depth);
	/* :157-161 pseudo-median of nine as the pivot */
	struct mk_block *m = mk_med3(
		mk_med3(&c[0], &c[n / 8], &c[n / 4]),
		mk_med3(&c[n / 2 - n / 8], &c[n / 2], &c[n / 2 + n / 8]),
		mk_med3(&c[n - 1 - n / 4], &c[n - 1 - n / 8], &c[n - 3]));
	const uint64_t pivot = m->key;
	/* Three-way partition (Dijkstra): [0,lt) < pivot, [lt,gt) == pivot,
	 * [gt,n) > pivot.  The reference uses Bentley-McIlroy swapping
	 * (:162-199); the partition sets are the same. */
	size_t lt = 0, i = 0, gt = n;
	while (i < gt) {
		if (c[i].key < pivot)
			mk_swap(&c[lt++], &c[i++]);
		else if (c[i].key > pivot)
			mk_swap(&c[i], &c[--gt]);
		else
			++i;
	}
	/* :201-211 -- the equal part continues at depth+8 only when its key
	 * has no NUL, otherwise its strings are identical. */
	mk_rec(c, lt, depth, 0);
	mk_rec(c + gt, n - gt, depth, 0);
	if (pivot & 0xFF)
		mk_rec(c + lt, gt - lt, depth + 8, 1);
}

void
or_multikey_cache8(unsigned char **strings, size_t n)
{
	if (n < 2)
		return;
	struct mk_block *c = malloc(n * sizeof(*c));
	if (!c)
		abort();
	for (size_t i = 0; i < n; ++i)
		c[i].ptr = strings[i];
	mk_rec(c, n, 0, 1);
	for (size_t i = 0; i < n; ++i)
		strings[i] = c[i].ptr;
	free(c);
}

int
or_sort_offsets32(int alg, const unsigned char *text, const uint32_t *off,
		  size_t n, uint32_t *off_out, uint64_t *lcp_out)
{
	unsigned char **s = malloc((n ? n : 1) * sizeof(*s));
	if (!s)
		return -1;
	for (size_t i = 0; i < n; ++i)
		s[i] = (unsigned char *)text + off[i];
	if (alg == 0)
		or_msd_ce2(s, n);
	else if (alg == 1)
		or_multikey_cache8(s, n);
	else {
		free(s);
		return -1;
	}
	for (size_t i = 0; i < n; ++i)
		off_out[i] = (uint32_t)(s[i] - text);
	if (lcp_out)
		for (size_t i = 0; i + 1 < n; ++i)
			lcp_out[i] = or_lcp(s[i], s[i + 1]);
	free(s);
	return 0;
}

/* src/sortstring.c:311-335: count the delimiters, then record the start of
 * the line each one ends and turn it into a NUL (unless it already is). */
size_t
or_ingest32(unsigned char *text, size_t len, int delim, uint32_t *off,
	    size_t max)
{
	size_t cnt = 0;
	for (size_t i = 0; i < len; ++i)
		if (text[i] == (unsigned char)delim)
			++cnt;
	size_t line_start = 0, j = 0;
	for (size_t i = 0; i < len; ++i)
		if (text[i] == (unsigned char)delim) {
			if (j < max)
				off[j] = (uint32_t)line_start;
			++j;
			line_start = i + 1;
			if (delim != '\0')
				text[i] = '\0';
		}
	return cnt;
}

/* src/sortstring.c:361-391: fputs(strings[i]); fputc('\n'). */
size_t
or_write32(const unsigned char *text, size_t len, const uint32_t *off,
	   size_t n, unsigned char *out, size_t cap)
{
	size_t total = 0;
	for (size_t i = 0; i < n; ++i) {
		size_t e = off[i];
		while (e < len && text[e] != 0)
			++e;
		total += e - off[i] + 1;
	}
	if (!out || total > cap)
		return total;
	size_t p = 0;
	for (size_t i = 0; i < n; ++i) {
		for (size_t e = off[i]; e < len && text[e] != 0; ++e)
			out[p++] = text[e];
		out[p++] = '\n';
	}
	return total;
}

/* src/mergesort_lcp.cpp:102-114. */
int
or_compare(const unsigned char *a, const unsigned char *b, size_t depth,
	   size_t *lcp)
{
	for (size_t i = depth;; ++i) {
		const unsigned char A = a[i], B = b[i];
		if (A == 0 || A != B) {
			*lcp = i;
			return (int)A - (int)B;
		}
	}
}

/* src/mergesort_lcp.cpp:133-211.  lcp0/lcp1 hold the LCP of the last output
 * with the heads of run 0 / run 1; the larger one wins without comparing,
 * equal values compare from that depth. */
void
or_merge_lcp32(const unsigned char *text, const uint32_t *a,
	       const uint64_t *alcp, size_t na, const uint32_t *b,
	       const uint64_t *blcp, size_t nb, uint32_t *out, uint64_t *olcp)
{
	size_t i = 0, j = 0, k = 0;
	size_t lcp0 = 0, lcp1 = 0;
	if (na == 0 || nb == 0) {
		for (; i < na; ++i) {
			out[k] = a[i];
			if (olcp && i + 1 < na)
				olcp[k] = alcp[i];
			++k;
		}
		for (; j < nb; ++j) {
			out[k] = b[j];
			if (olcp && j + 1 < nb)
				olcp[k] = blcp[j];
			++k;
		}
		return;
	}
	{
		size_t l01;
		const int c = or_compare(text + a[0], text + b[0], 0, &l01);
		if (c <= 0) {
			out[k++] = a[i];
			lcp0 = na > 1 ? alcp[0] : 0;
			lcp1 = l01;
			++i;
		} else {
			out[k++] = b[j];
			lcp1 = nb > 1 ? blcp[0] : 0;
			lcp0 = l01;
			++j;
		}
	}
	while (i < na && j < nb) {
		if (lcp0 > lcp1) {
			if (olcp)
				olcp[k - 1] = lcp0;
			out[k++] = a[i];
			lcp0 = i + 1 < na ? alcp[i] : 0;
			++i;
		} else if (lcp0 < lcp1) {
			if (olcp)
				olcp[k - 1] = lcp1;
			out[k++] = b[j];
			lcp1 = j + 1 < nb ? blcp[j] : 0;
			++j;
		} else {
			size_t l01;
			const int c = or_compare(text + a[i], text + b[j], lcp0, &l01);
			if (olcp)
				olcp[k - 1] = lcp0;
			if (c <= 0) {
				out[k++] = a[i];
				lcp1 = l01;
				lcp0 = i + 1 < na ? alcp[i] : 0;
				++i;
			} else {
				out[k++] = b[j];
				lcp0 = l01;
				lcp1 = j + 1 < nb ? blcp[j] : 0;
				++j;
			}
		}
	}
	/* One run is exhausted: the other's head follows the last output with
	 * the tracked LCP, then its own LCPs. */
	for (int first = 1; i < na; ++i, first = 0) {
		if (olcp)
			olcp[k - 1] = first ? lcp0 : alcp[i - 1];
		out[k++] = a[i];
	}
	for (int first = 1; j < nb; ++j, first = 0) {
		if (olcp)
			olcp[k - 1] = first ? lcp1 : blcp[j - 1];
		out[k++] = b[j];
	}
}
