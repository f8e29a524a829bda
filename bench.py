#!/usr/bin/env python3
"""Benchmark of the MI355X MSD radix string sort.

One step = one device-resident sort (gss_sort_device32 through the C ABI) of
10,000,000 URL-like strings (2,000 shared prefixes of 20-60 bytes plus 5-40
byte suffixes over [a-z0-9./:-], 30% exact duplicates), with the text and
offsets already in HBM.  `value` is Mstrings/s over the timed steps.

Besides the contract fields the JSON line carries:
  roofline      the dominant kernel's algorithmic bytes per launch / its
                average launch time (HIP events on the launch stream),
                against the 8 TB/s HBM peak;
  cpu_baseline  the reference's own single-core routines (msd_CE2, msd_CE6,
                msd_ci_adaptive, msd_A2_adaptive, multikey_cache8, burstsortA,
                compiled from its sources into oracle/_ref) on one core of
                this host, on the same 10M strings: `value` is the fastest,
                7 runs, trimmed mean of the middle 5; msd_CE2 beside it, and
                parallel_msd_radix_sort on the host's cores (CPU model and
                thread count stated), plus the 1M uniform a-z set;
  hbm_gbps      algorithmic bytes of the whole sort / sort time;
  e2e           the routine.h pointer path (gss_msd), PCIe transfers included;
  pipeline      the steps either side of the sort on the same set: device
                ingest of the newline file, --write formatting, LCP-aware
                merge of two sorted halves, suffix sorting of a 10 MB ACGT
                text, and file-in / file-out with PCIe;
  batched       the request stream (20,000 sorts of 1e3..4e6 strings from the
                four generators' seeded pools): the served loop (continuous
                batching, gss_server_*) against one plain gss_sort_device32
                call per request -- requests/s, p50/p99 latency at swept
                Poisson rates, and whether both outputs are identical.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "string-sorting_amd"), str(ROOT / "oracle")]

METRIC = "Mstrings/s + HBM GB/s (10M URL-like, 1 GPU); batched p50/p99 vs plain"
HBM_PEAK_GBPS = 8000.0
# Algorithmic bytes per active string per MSD round: read its 32-bit offset
# and 8 text bytes, write the partitioned (8-byte key, offset) pair.
BYTES_PER_ROUND_STRING = 2 * (4 + 8)  # per string and 8-byte key level
KERNELS = ["plan", "gather", "segreduce", "layout", "histogram", "flat_scan", "segscan", "scatter",
           "boundary", "retire", "local128", "local1024", "local6144", "local128_early",
           "local1024_early", "local6144_early"]
PASSES = [f"pass{r}" for r in range(9)] + ["pass9+", "local_phase"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Library profile names -> device symbol fragments in the rocprofv3 tables.
# (The local kernels' last template argument is the first round's key
# words, so the fragments stop before it.)
SYMBOLS = {"gather": "k_gather_minmax", "histogram": "k_histogram", "scatter": "k_scatter",
           "segscan": "k_segscan", "local128": "k_local<64, 128, false,",
           "local1024": "k_local<256, 1024, false,", "local6144": "k_local<1024, 6144, false,"}
PROFILE = ROOT / "profiles" / "url_like_10M_kernels.md"


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (calibrated FETCH_SIZE + WRITE_SIZE)
    from the committed rocprofv3 summary, or None when it has no row."""
    sym = SYMBOLS.get(kernel)
    if not sym or not PROFILE.exists():
        return None
    text = PROFILE.read_text()
    at = text.find("## HBM traffic")
    if at < 0:
        return None
    for line in text[at:].splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and sym in cells[0]:
            try:
                return (float(cells[2]) + float(cells[3])) * 2**20
            except ValueError:
                return None
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--lcp", action="store_true", help="also emit the LCP array in the timed steps")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-e2e", action="store_true", help="skip the pointer-path timing")
    ap.add_argument("--cpu-runs", type=int, default=CPU_RUNS)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, copy 1 GiB device to device twice: the "
                         "reference dispatch that scales FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--stream-requests", type=int, default=20000,
                    help="requests of the served-stream comparison (0 = skip it)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="skip the ingest / write / merge / suffix rows")
    ap.add_argument("--all-configs", action="store_true",
                    help="also time every workload at its full size (adds about a minute)")
    ap.add_argument("--stream-only", action="store_true",
                    help="run only the served-stream comparison")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--serve-max-batch", type=int, default=1024)
    ap.add_argument("--rates", type=str, default="1,2,4",
                    help="swept Poisson rates, as multiples of the plain path's capacity")
    return ap.parse_args()


def _pctl(x, q):
    """Nearest-rank percentile (the smallest sample with >= q% at or below)."""
    return float(np.percentile(np.asarray(x), q, method="inverted_cdf")) if len(x) else 0.0


def _best_time(fn, runs, prepare=None):
    """Best wall time of `fn` (which returns once its work is complete on the
    device) over `runs` calls; `prepare` runs untimed before each call."""
    import torch

    best = float("inf")
    for _ in range(runs):
        if prepare is not None:
            prepare()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def pipeline_rows(ss, dev, runs=5):
    """The steps either side of the sort on the same 10M URL-like set:
    device ingest of the newline file, --write formatting of the sorted
    output, LCP-aware merge of two sorted halves and of eight sorted runs,
    suffix sorting of a 10 MB ACGT text, and the whole file-in / file-out
    pipeline with PCIe."""
    import torch

    import gss

    res = {}
    n, tb = ss.n, ss.text_bytes
    nl = ss.text.copy()
    nl[:tb][nl[:tb] == 0] = ord("\n")
    pristine = torch.from_numpy(nl).to(dev)
    work = torch.empty_like(pristine)
    box = {}
    off_buf = torch.empty(n, dtype=torch.int32, device=dev)

    def ingest():
        box["off"] = gss.ingest(work, b"\n", text_bytes=tb, out=off_buf)

    t = _best_time(ingest, runs, prepare=lambda: work.copy_(pristine))
    off = box["off"]
    ok = bool(torch.equal(off.cpu(), torch.from_numpy(ss.offsets.view(np.int32))))
    alg = tb + 4 * n
    res["ingest"] = {"seconds": t, "strings": n, "gb_per_s": alg / t / 1e9,
                     "frac_hbm": alg / t / 1e9 / HBM_PEAK_GBPS, "matches_generator": ok,
                     "alg_bytes": alg, "note": "text read once + 32-bit offsets written"}

    out = torch.empty_like(off)
    lcp = torch.empty(n - 1, dtype=torch.int64, device=dev)
    gss.sort_device(work, off, out, lcp, text_bytes=tb)
    box.clear()

    def write():
        box["w"] = gss.write(work, out, text_bytes=tb)

    t = _best_time(write, runs)
    wbytes = int(box["w"].numel())
    alg = 4 * n + 2 * wbytes - n
    res["write"] = {"seconds": t, "out_bytes": wbytes, "gb_per_s": alg / t / 1e9,
                    "frac_hbm": alg / t / 1e9 / HBM_PEAK_GBPS, "alg_bytes": alg,
                    "note": "offsets + string bytes read, formatted output written"}
    del box["w"]

    # Merge: sort each half of the input order, then merge the two runs.
    h = n // 2
    a, b = torch.empty(h, dtype=torch.int32, device=dev), torch.empty(n - h, dtype=torch.int32,
                                                                      device=dev)
    al = torch.empty(h - 1, dtype=torch.int64, device=dev)
    bl = torch.empty(n - h - 1, dtype=torch.int64, device=dev)
    gss.sort_device(work, off[:h].contiguous(), a, al, text_bytes=tb)
    gss.sort_device(work, off[h:].contiguous(), b, bl, text_bytes=tb)

    def merge():
        box["m"] = gss.merge_lcp(work, a, al, b, bl, text_bytes=tb)

    t = _best_time(merge, runs)
    m, ml = box.pop("m")
    same = bool(torch.equal(m, out)) and bool(torch.equal(ml, lcp))
    alg = 24 * n
    res["merge_lcp"] = {"seconds": t, "strings": n, "mstrings_per_s": n / t / 1e6,
                        "gb_per_s": alg / t / 1e9, "frac_hbm": alg / t / 1e9 / HBM_PEAK_GBPS,
                        "equals_full_sort": same,
                        "note": "two sorted 5M halves; 24 B/string = offset+LCP in, offset+LCP out"}
    del m, ml, a, b, al, bl

    # K-way merge: eight sorted eighths of the input order.
    k = 8
    bounds = [n * r // k for r in range(k + 1)]
    kruns, klcps = [], []
    for r in range(k):
        lo, hi = bounds[r], bounds[r + 1]
        ro = torch.empty(hi - lo, dtype=torch.int32, device=dev)
        rl = torch.empty(max(1, hi - lo - 1), dtype=torch.int64, device=dev)
        gss.sort_device(work, off[lo:hi].contiguous(), ro, rl, text_bytes=tb)
        kruns.append(ro)
        klcps.append(rl[: max(0, hi - lo - 1)])

    def merge_kway():
        box["k"] = gss.merge_lcp_kway(work, kruns, klcps, text_bytes=tb)

    t = _best_time(merge_kway, runs)
    m, ml = box.pop("k")
    same = bool(torch.equal(m, out)) and bool(torch.equal(ml, lcp))
    res["merge_lcp_kway"] = {"seconds": t, "strings": n, "runs": k,
                             "mstrings_per_s": n / t / 1e6, "equals_full_sort": same,
                             "note": "eight sorted 1.25M runs merged in three rounds of 2-way merges"}
    del m, ml, kruns, klcps

    # Suffix sorting of a 10 MB i.i.d. ACGT text.
    rng = np.random.default_rng(20242)
    dna = np.concatenate([np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, size=10_000_000)],
                          np.zeros(16, np.uint8)])
    dt_ = torch.from_numpy(dna).to(dev)
    soff = gss.suffixes(10_000_000, dev)
    sout = torch.empty_like(soff)

    def suffix_sort():
        gss.sort_device(dt_, soff, sout, text_bytes=10_000_000)

    t = _best_time(suffix_sort, runs)
    s_sorted = sout.cpu().numpy().view(np.uint32)
    import oracle

    inv = oracle.check_result(dna, s_sorted)[2]
    res["suffix_sort"] = {"seconds": t, "suffixes": 10_000_000, "mstrings_per_s": 1e7 / t / 1e6,
                          "inversions": int(inv),
                          "permutation": bool(np.array_equal(np.sort(s_sorted),
                                                             np.arange(10_000_000, dtype=np.uint32)))}
    del dt_, soff, sout

    # File in, file out: H2D of the newline file, ingest, sort, write, D2H.
    host_out = torch.empty(wbytes, dtype=torch.uint8, pin_memory=True)
    host_in = torch.from_numpy(nl).pin_memory()

    def file_pipeline():
        w = pristine.copy_(host_in, non_blocking=True)
        o = gss.ingest(w, b"\n", text_bytes=tb, out=off_buf)
        srt = torch.empty_like(o)
        gss.sort_device(w, o, srt, text_bytes=tb)
        host_out.copy_(gss.write(w, srt, text_bytes=tb))

    t = _best_time(file_pipeline, 3)
    res["file_pipeline"] = {"seconds": t, "mstrings_per_s": n / t / 1e6,
                            "note": "pinned H2D of the 635 MB newline file, ingest, sort, write, "
                                    "D2H of the sorted file"}
    return res


def all_configs(dev, runs=5):
    """Each workload at its full size (1M uniform a-z, 10M URL-like, 32M DNA,
    100M Zipf words): median device-resident sort time over `runs` sorts,
    without and with the LCP array, and the pass statistics."""
    import torch

    import gss
    import gss.gen as gen

    out = {}
    for cfg in range(4):
        t0 = time.time()
        ss = gen.make(cfg)
        log(f"config {cfg}: {ss.n} strings generated in {time.time() - t0:.1f}s")
        text = torch.from_numpy(ss.text).to(dev)
        off = torch.from_numpy(ss.offsets.view(np.int32)).to(dev)
        o = torch.empty_like(off)
        lcp = torch.empty(ss.n - 1, dtype=torch.int64, device=dev)
        row = {"workload": gen.NAMES[cfg], "strings": ss.n, "text_bytes": ss.text_bytes}
        for name, lc in (("sort", None), ("sort_lcp", lcp)):
            for _ in range(2):
                gss.sort_device(text, off, o, lc, text_bytes=ss.text_bytes)
            ts = []
            for _ in range(runs):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                gss.sort_device(text, off, o, lc, text_bytes=ss.text_bytes)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t1)
            t = sorted(ts)[len(ts) // 2]
            row[name] = {"ms": t * 1e3, "mstrings_per_s": ss.n / t / 1e6}
        st = gss.stats()
        row["global_passes"] = int(st.global_passes)
        row["local_round_strings"] = int(st.local_round_strings)
        row["local_level_strings"] = int(st.local_level_strings)
        out[gen.NAMES[cfg]] = row
        del text, off, o, lcp, ss
        torch.cuda.empty_cache()
    return out


def served_stream(args, dev):
    """Request-stream comparison: the served loop (continuous batching over
    gss_server_*) against one plain gss_sort_device32 call per request, on
    the same requests and arrival times.  Inputs are device-resident pools;
    each request's sorted offsets land in its own range of an output arena,
    one arena per mode, and the two arenas must be identical."""
    import torch

    import gss
    import gss.gen as gen

    t0 = time.time()
    rs = gen.request_stream(nreq=args.stream_requests)
    nreq = rs.nreq
    sizes = rs.size.astype(np.int64)
    firsts = np.array([rs.first(i) for i in range(nreq)], dtype=np.int64)
    pos = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    total = int(sizes.sum())
    log(f"request stream: {nreq} requests, {total} strings, pools {rs.pools.n} strings / "
        f"{rs.pools.text_bytes} bytes, built in {time.time() - t0:.1f}s")
    tb = rs.pools.text_bytes
    text = torch.from_numpy(rs.pools.text).to(dev)
    off = torch.from_numpy(rs.pools.offsets.view(np.int32)).to(dev)
    arena = {m: torch.empty(total, dtype=torch.int32, device=dev) for m in ("plain", "served")}
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()
    L = gss.lib()
    tp, op = text.data_ptr(), off.data_ptr()
    in_ptr = [op + 4 * int(a) for a in firsts]
    ns = [int(n) for n in sizes]
    out_ptr = {m: [arena[m].data_ptr() + 4 * int(p) for p in pos] for m in arena}
    st = stream.cuda_stream

    # Warm-up: size the default context's workspace for the largest request.
    big = int(np.argmax(sizes))
    for _ in range(2):
        rc = L.gss_sort_device32(tp, tb, in_ptr[big], out_ptr["plain"][big], None, ns[big], st)
        if rc:
            raise RuntimeError(f"plain sort failed: {gss.last_error()}")

    def run_plain(due):
        lat = np.empty(nreq)
        start = time.perf_counter()
        for i in range(nreq):
            t_due = start + due[i]
            while time.perf_counter() < t_due:
                pass
            rc = L.gss_sort_device32(tp, tb, in_ptr[i], out_ptr["plain"][i], None, ns[i], st)
            if rc:
                raise RuntimeError(f"plain sort of request {i} failed: {gss.last_error()}")
            lat[i] = time.perf_counter() - t_due
        return time.perf_counter() - start, lat * 1e3

    budget = int(args.serve_budget_gib * 2**30)
    srv = gss.Server(text, tb, hbm_budget=budget, slots=args.serve_slots,
                     max_batch_requests=args.serve_max_batch)
    w = srv.submit(off[firsts[big]:firsts[big] + ns[big]],
                   arena["served"][pos[big]:pos[big] + ns[big]])
    srv.wait(w)

    def run_served(due):
        ids = [0] * nreq
        start = time.perf_counter()
        for i in range(nreq):
            t_due = start + due[i]
            while time.perf_counter() < t_due:
                pass
            rid = L.gss_server_submit(srv.handle, in_ptr[i], out_ptr["served"][i], None, ns[i],
                                      None)
            if rid < 0:
                raise RuntimeError(f"submit of request {i} failed: {gss.last_error()}")
            ids[i] = rid
        lat = np.array([srv.wait(r) for r in ids])
        return time.perf_counter() - start, lat

    def identical():
        torch.cuda.synchronize()
        step = 1 << 28
        return all(torch.equal(arena["plain"][a:a + step], arena["served"][a:a + step])
                   for a in range(0, total, step))

    def summary(wall, lat):
        return {"wall_s": wall, "req_per_s": nreq / wall, "mstrings_per_s": total / wall / 1e6,
                "p50_ms": _pctl(lat, 50), "p99_ms": _pctl(lat, 99), "max_ms": float(lat.max())}

    zero = np.zeros(nreq)
    log("plain, all requests at once ...")
    plain = summary(*run_plain(zero))
    log("served, all requests at once ...")
    served = summary(*run_served(zero))
    same = identical()
    sst = srv.stats()
    served.update({"batches": int(sst.batches) - 1, "batch_capacity": int(sst.batch_capacity),
                   "max_batch_requests": int(sst.max_batch_requests_seen)})
    sweep = []
    for mult in [float(x) for x in args.rates.split(",") if x]:
        rate = mult * plain["req_per_s"]
        due = gen.arrivals(nreq, rate)
        log(f"rate {rate:.0f}/s: plain ...")
        p = summary(*run_plain(due))
        log(f"rate {rate:.0f}/s: served ...")
        s = summary(*run_served(due))
        same = same and identical()
        sweep.append({"rate_req_per_s": rate, "x_plain_capacity": mult,
                      "plain": {k: p[k] for k in ("p50_ms", "p99_ms", "req_per_s")},
                      "served": {k: s[k] for k in ("p50_ms", "p99_ms", "req_per_s")}})
    srv.close()
    return {
        "workload": (f"first {nreq} of the {gen.STREAM_REQUESTS}-request stream (seed "
                     f"{gen.STREAM_SEED}): sizes floor(exp(U[ln 1e3, ln 4e6])), generator "
                     f"U{{0..3}}, payload = a contiguous slice of that generator's seeded "
                     f"{gen.POOL_N}-string pool, pools device-resident"),
        "requests": nreq, "strings": total,
        "slots": args.serve_slots, "hbm_budget_bytes": budget,
        "max_batch_requests": args.serve_max_batch,
        "plain": plain, "served": served,
        "throughput_ratio": served["req_per_s"] / plain["req_per_s"],
        "sweep": sweep, "identical_output": bool(same),
    }


# The reference's best single-core routines for this path and its multicore
# baseline (BASELINE.md's plan): msd_CE2 is the parity oracle, the others
# are the fastest of their families on these inputs.
CPU_ROUTINES = ["msd_CE2", "msd_CE6", "msd_ci_adaptive", "msd_A2_adaptive", "multikey_cache8",
                "burstsortA"]
CPU_RUNS = 7  # the reference's report: 7 runs, trimmed mean of the middle 5


def _trimmed_mean(times):
    """Mean of the middle 5 of 7 runs (report/htmlreport.xsl:265-277); with
    fewer runs, the mean of all but the fastest and slowest when possible."""
    t = sorted(times)
    if len(t) >= 3:
        t = t[1:-1]
    return sum(t) / len(t)


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_routine(name, ptrs0, runs):
    """Wall times (perf_counter_ns, seconds) of reference routine `name` on
    fresh copies of the pointer array; the copy is outside the timing, as
    sortstring's pointer build is (src/sortstring.c:424-451)."""
    import oracle

    fn = oracle.ref_routine(name)
    times = []
    for _ in range(runs):
        p = ptrs0.copy()
        t0 = time.perf_counter_ns()
        fn(p.ctypes.data, p.shape[0])
        times.append((time.perf_counter_ns() - t0) * 1e-9)
    return times, p


def cpu_baseline(ss, runs=CPU_RUNS):
    """The reference's own routines (oracle/_ref/libref.so, compiled from its
    sources) on this host, in this process:

    * single core (the process pinned to one CPU with sched_setaffinity):
      every routine of CPU_ROUTINES once on the full URL-like set to screen,
      then `runs` runs of the fastest and of msd_CE2, trimmed mean;
    * parallel_msd_radix_sort on all CPUs this process may use, `runs` runs;
    * the same on the 1,000,000-string uniform a-z set (config 0), every
      routine `runs` times, on a nanosecond clock.

    `value` is the fastest single-core routine on the URL-like set.  Falls
    back to the oracle's C restatement of msd_CE2 when the reference build
    is absent."""
    import oracle

    import gss.gen as gen

    base = np.uint64(ss.text.ctypes.data)
    ptrs0 = (base + ss.offsets.astype(np.uint64)).astype(np.uint64)
    if not oracle.ref_available():
        lib = oracle.lib()
        times = []
        for _ in range(runs):
            p = ptrs0.copy()
            t0 = time.perf_counter_ns()
            lib.or_msd_ce2(p.ctypes.data, p.shape[0])
            times.append((time.perf_counter_ns() - t0) * 1e-9)
        t = _trimmed_mean(times)
        return {"value": ss.n / t / 1e6, "unit": "Mstrings/s", "cores": 1, "kind": "port",
                "sample": f"msd_CE2 restatement (oracle.c) on the full {ss.n}-string URL-like "
                          f"set, trimmed mean of {runs} runs"}
    allowed = sorted(os.sched_getaffinity(0))
    cpu_model = _cpu_model()
    want = np.sort(ss.offsets)
    out = {"unit": "Mstrings/s", "cores": 1, "kind": "reference", "cpu_model": cpu_model,
           "nproc": len(allowed), "os_cpu_count": os.cpu_count()}
    try:
        os.sched_setaffinity(0, {allowed[0]})
        screen = {}
        for name in CPU_ROUTINES:
            t, p = _time_routine(name, ptrs0, 1)
            got = (p - base).astype(np.uint32)
            ok = bool(np.array_equal(np.sort(got), want)) and oracle.check_result(ss.text, got)[2] == 0
            screen[name] = {"seconds": t[0], "mstrings_per_s": ss.n / t[0] / 1e6, "check": ok}
            log(f"  {name}: {t[0]:.3f} s (sorted: {ok})")
        best = min((k for k in screen if screen[k]["check"]), key=lambda k: screen[k]["seconds"])
        timed = {}
        for name in dict.fromkeys([best, "msd_CE2"]):
            t, _ = _time_routine(name, ptrs0, runs)
            timed[name] = {"trimmed_mean_s": _trimmed_mean(t), "runs_s": t,
                           "mstrings_per_s": ss.n / _trimmed_mean(t) / 1e6}
        small = gen.make(0)
        sbase = np.uint64(small.text.ctypes.data)
        sptrs = (sbase + small.offsets.astype(np.uint64)).astype(np.uint64)
        config0 = {}
        for name in CPU_ROUTINES:
            t, _ = _time_routine(name, sptrs, runs)
            config0[name] = {"trimmed_mean_ms": _trimmed_mean(t) * 1e3,
                             "mstrings_per_s": small.n / _trimmed_mean(t) / 1e6}
    finally:
        os.sched_setaffinity(0, allowed)
    out.update({
        "value": timed[best]["mstrings_per_s"],
        "routine": best,
        "sample": (f"{best} (reference sources, gcc -O3), one core of {cpu_model}, on the full "
                   f"{ss.n}-string URL-like set: trimmed mean of the middle 5 of {runs} runs; "
                   f"screened against {', '.join(CPU_ROUTINES)} (one run each)"),
        "screen": screen, "timed": timed,
        "msd_CE2": timed["msd_CE2"]["mstrings_per_s"],
        "config0_uniform_1M": config0,
    })
    # The multicore baseline on every CPU this process may use (bounded by
    # OMP_NUM_THREADS when the environment sets it: a host shared between
    # jobs gives each its share that way).  OpenMP reads its thread count
    # when libref.so loads, so the count is set through the OpenMP runtime
    # the library uses.
    try:
        import ctypes as C

        env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = min(len(allowed), env_threads) if env_threads > 0 else len(allowed)
        gomp = C.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(threads)
        t, p = _time_routine("parallel_msd_radix_sort", ptrs0, runs)
        got = (p - base).astype(np.uint32)
        t0, _ = _time_routine("parallel_msd_radix_sort", sptrs, runs)
        out["multicore"] = {"routine": "parallel_msd_radix_sort", "threads": threads,
                            "cpu_model": cpu_model, "nproc": len(allowed),
                            "omp_num_threads_env": env_threads or None,
                            "value": ss.n / _trimmed_mean(t) / 1e6,
                            "trimmed_mean_s": _trimmed_mean(t), "runs_s": t,
                            "config0_mstrings_per_s": small.n / _trimmed_mean(t0) / 1e6,
                            "sorted": bool(np.array_equal(np.sort(got), want))}
    except Exception as e:  # the baseline is a report, not the target
        out["multicore"] = {"error": str(e)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); sharding the sort "
              "across GPUs is designed but not built.", file=sys.stderr)
        sys.exit(2)
    import torch

    import gss
    import gss.gen as gen

    dev = torch.device("cuda:0")
    if args.stream_only:
        print(json.dumps({"batched": served_stream(args, dev)}), flush=True)
        return
    t0 = time.time()
    ss = gen.url_like(args.n)
    log(f"generated {ss.n} strings, {ss.text_bytes} text bytes in {time.time() - t0:.1f}s")
    text = torch.from_numpy(ss.text).to(dev)
    off = torch.from_numpy(ss.offsets.view(np.int32)).to(dev)
    out = torch.empty_like(off)
    lcp = torch.empty(max(1, ss.n - 1), dtype=torch.int64, device=dev) if args.lcp else None
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()

    def step():
        gss.sort_device(text, off, out, lcp, text_bytes=ss.text_bytes, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # Kernel and pass breakdown: the same sort again, outside the timed steps
    # above.  Pass times come from sorts in which only each pass and the
    # local phase are bracketed by HIP events (an event between two kernels
    # holds the second back about 10 us; queued kernels otherwise follow each
    # other with no gap), kernel times from sorts with every launch
    # bracketed.  The key levels the local kernels take strings through
    # (their work units) come from one more sort with their counting variant
    # on; durations come from the production kernels.
    prof_steps = max(1, min(args.steps, 10))
    gss.profile_reset()
    gss.profile_enable(True, passes_only=True)
    for _ in range(prof_steps):
        step()
    torch.cuda.synchronize()
    pass_timing = {k: gss.profile_get(k) for k in PASSES}
    gss.profile_reset()
    gss.profile_enable(True)
    for _ in range(prof_steps):
        step()
    torch.cuda.synchronize()
    timing = {k: gss.profile_get(k) for k in KERNELS}
    timing.update(pass_timing)
    gss.profile_reset()
    gss.profile_enable(True, levels=True)
    step()
    torch.cuda.synchronize()
    gss.profile_enable(False)
    counting = {k: gss.profile_get(k) for k in KERNELS + PASSES}
    st = gss.stats()
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ms_per_step = dt / args.steps * 1e3
    value = ss.n * args.steps / dt / 1e6

    def per_sort(k):
        """(launches, ms, units) of kernel or pass `k` per sort: launches
        and time from the timing steps, units from the counting sort."""
        launches, ms, _ = timing[k]
        c_launches, _, c_items = counting[k]
        return launches / prof_steps, ms / prof_steps, c_items if c_launches else 0

    prof = {}
    for k in KERNELS:
        launches, ms, items = per_sort(k)
        if launches:
            prof[k] = {"launches": launches, "ms_total": ms, "ms_avg": ms / launches,
                       "items": items}
    # Work units of one sort: active strings summed over all 8-byte MSD key
    # levels -- one level per global pass, and in the local kernel the levels
    # each string's 32-byte rounds took it through.
    global_rounds = [int(st.active[i]) for i in range(min(st.global_passes, 64))]
    round_strings = sum(global_rounds) + int(st.local_level_strings)
    alg_bytes = BYTES_PER_ROUND_STRING * round_strings + 4 * ss.n + (8 * (ss.n - 1) if args.lcp else 0)
    hbm_gbps = alg_bytes / (ms_per_step * 1e-3) / 1e9

    # Per pass: every global MSD pass and the local phase, bracketed as a
    # whole (HIP events on the launch stream), against 24 B per active
    # string per 8-byte key level (a local round covers up to four levels).
    passes = []
    for name in PASSES:
        launches, ms, items = per_sort(name)
        if not launches:
            continue
        units = items / launches
        avg = ms / launches
        gbps = BYTES_PER_ROUND_STRING * units / (avg * 1e-3) / 1e9
        passes.append({"pass": name, "strings": units, "ms": avg,
                       "alg_bytes": BYTES_PER_ROUND_STRING * units, "gb_per_s": gbps,
                       "frac": gbps / HBM_PEAK_GBPS})

    # A local class launched early (overlapping the next global pass, on a
    # second stream) and finally is one kernel: combine both for the roofline.
    combined = {}
    for k, v in prof.items():
        c = combined.setdefault(k.replace("_early", ""),
                                {"launches": 0, "ms_total": 0.0, "items": 0})
        c["launches"] += v["launches"]
        c["ms_total"] += v["ms_total"]
        c["items"] += v["items"]
    for v in combined.values():
        v["ms_avg"] = v["ms_total"] / v["launches"]
    dominant = max(combined, key=lambda k: combined[k]["ms_total"])
    d = combined[dominant]
    per_launch_bytes = BYTES_PER_ROUND_STRING * d["items"] / d["launches"]
    achieved = per_launch_bytes / (d["ms_avg"] * 1e-3) / 1e9
    traffic = measured_traffic(dominant)
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                "traffic_source": str(PROFILE.relative_to(ROOT)) if traffic else None,
                "traffic_gbps": traffic / (d["ms_avg"] * 1e-3) / 1e9 if traffic else None,
                "alg_bytes_per_launch": per_launch_bytes, "avg_launch_ms": d["ms_avg"]}

    res = {
        "metric": METRIC, "value": value, "unit": "Mstrings/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "profiled_steps": prof_steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded URL-like generator, string-sorting_amd/gss/gen.py)",
        "config": {"workload": "url_like_10M", "strings": ss.n, "text_bytes": ss.text_bytes,
                   "prefixes": 2000, "prefix_len": [20, 60], "suffix_len": [5, 40],
                   "duplicates": 0.3, "lcp": bool(args.lcp), "parallelism": "single GPU"},
        "hbm_gbps": hbm_gbps,
        "alg_bytes_per_sort": alg_bytes,
        "rounds": {"global_passes": int(st.global_passes), "global_active": global_rounds,
                   "local_segments": int(st.local_segments),
                   "local_strings": int(st.local_strings),
                   "local_round_strings": int(st.local_round_strings),
                   "local_level_strings": int(st.local_level_strings),
                   "local_rounds_max": int(st.local_rounds_max)},
        "kernels": prof,
        "passes": passes,
        "roofline": roofline,
    }

    # Parity of the benchmarked output (full size, size-independent checks).
    import oracle

    got = out.cpu().numpy().view(np.uint32)
    ident, inval, wrong = oracle.check_result(ss.text, got)
    perm_ok = bool(np.array_equal(np.sort(got), np.sort(ss.offsets)))
    res["check"] = {"inversions": int(wrong), "identical_ptrs": int(ident), "permutation": perm_ok}

    if not args.no_e2e:
        base = ss.text.ctypes.data
        ptrs = (np.uint64(base) + ss.offsets.astype(np.uint64)).astype(np.uint64)
        arr = ptrs.ctypes.data_as(C.POINTER(C.c_void_p))
        gss.lib().gss_msd(arr, ss.n)  # warm: device buffers for the pointer path
        ptrs[:] = (np.uint64(base) + ss.offsets.astype(np.uint64))
        t0 = time.perf_counter()
        gss.lib().gss_msd(arr, ss.n)
        te = time.perf_counter() - t0
        e2e_ok = bool(np.array_equal((ptrs - np.uint64(base)).astype(np.uint32), got))
        res["e2e"] = {"entry": "gss_msd (host pointers: H2D text + 32-bit offsets, D2H offsets widened to pointers on the host)",
                      "seconds": te, "mstrings_per_s": ss.n / te / 1e6, "matches_device": e2e_ok}

    if not args.no_pipeline:
        log("ingest / write / merge / suffix rows ...")
        res["pipeline"] = pipeline_rows(ss, dev)
        torch.cuda.empty_cache()

    if args.stream_requests > 0:
        del text, off, out, lcp
        torch.cuda.empty_cache()
        res["batched"] = served_stream(args, dev)

    if args.all_configs:
        torch.cuda.empty_cache()
        res["configs"] = all_configs(dev)

    if not args.no_cpu:
        log("timing the CPU baseline ...")
        res["cpu_baseline"] = cpu_baseline(ss, args.cpu_runs)
        res["vs_cpu"] = value / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
