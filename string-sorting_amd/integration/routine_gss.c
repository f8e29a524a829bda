/*
 * Registration of the MI355X string sort in the reference's routine registry.
 *
 * This is the translation unit a maintainer adds to rantala/string-sorting
 * (next to src/msd_ce.cpp, compiled into both sortstring and unit-test,
 * CMakeLists.txt:61-64) and links against libgss.so.  It uses the
 * reference's own macro, ROUTINE_REGISTER_MULTICORE (src/routine.h:41-59), so
 * `sortstring -L` lists the routines, `sortstring --check gss_msd FILE` runs
 * them (src/sortstring.c:424-451) and unit-test's test_routines() exercises
 * them (unit-test/main.cpp:207-272) with no change to either driver.
 *
 * The device is initialised on the first call, not in the constructor that
 * ROUTINE_REGISTER creates.
 */
#include "routine.h"
#include "gss.h"

ROUTINE_REGISTER_MULTICORE(gss_msd, "MI355X MSD radix sort, 8-byte keys (HIP, gfx950)")

/* The same sort with the LCP array computed and discarded, so the harness
 * times the LCP-emitting kernels through the unchanged routine signature. */
void gss_msd_with_lcp(unsigned char **strings, size_t n)
{
	size_t *lcp = n > 1 ? (size_t *)malloc((n - 1) * sizeof(size_t)) : NULL;
	if (n > 1 && !lcp)
		abort();
	gss_msd_lcp(strings, n, lcp);
	free(lcp);
}

ROUTINE_REGISTER_MULTICORE(gss_msd_with_lcp, "MI355X MSD radix sort + LCP array (HIP, gfx950)")
