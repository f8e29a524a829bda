"""GPU: batched sorting of independent requests (gss_sort_batch_device32) and
sort contexts.  Every request of a batch must come out exactly as it does
alone and as the oracle's msd_CE2 sorts it."""

import threading

import numpy as np
import pytest

from gpu_util import device, gpu_sort, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402


def _mixed_requests(seed, sizes):
    rng = np.random.default_rng(seed)
    sets = []
    for i, n in enumerate(sizes):
        cfg = int(rng.integers(0, 4))
        sets.append(gen.make(cfg, n=n, seed=seed * 1000 + i))
    return sets


def _batch_sort(joined, req, lcp=True, ctx=None):
    import torch

    dev = device()
    text = torch.from_numpy(joined.text).to(dev)
    off = torch.from_numpy(joined.offsets.view(np.int32)).to(dev)
    out = torch.empty_like(off)
    lc = None
    if lcp:
        lc = torch.full((max(1, joined.n - 1),), -7, dtype=torch.int64, device=dev)
    gss.sort_batch(text, off, req, out=out, lcp=lc, text_bytes=joined.text_bytes, ctx=ctx)
    torch.cuda.synchronize()
    o = out.cpu().numpy().view(np.uint32)
    l = lc.cpu().numpy() if lcp else None
    return o, l


def _check_requests(joined, req, sets, o, l, oracle):
    base = 0
    for r, s in enumerate(sets):
        a, b = req[r], req[r + 1]
        got = o[a:b].astype(np.int64) - base
        want, want_lcp = oracle.sort_offsets(s.text, s.offsets, "msd_ce2", lcp=True)
        assert np.array_equal(got, want.astype(np.int64)), f"request {r} (n={s.n}) differs"
        if l is not None and b - a > 1:
            assert np.array_equal(l[a:b - 1].astype(np.uint64), want_lcp), f"request {r} LCP"
        if l is not None and b < joined.n and b > 0:
            assert l[b - 1] == -7, f"LCP entry between requests {r} and {r + 1} was written"
        base += s.text_bytes


@pytest.mark.parametrize("sizes", [
    [0, 1, 2, 3, 64, 65, 128, 129, 1000, 1024, 1025, 5000, 6144, 6145, 8192, 8193, 30000],
    [40000, 1, 0, 70000, 12, 9000],
    [1] * 50 + [2] * 50,
])
def test_batch_matches_oracle(sizes, oracle_mod):
    sets = _mixed_requests(len(sizes), sizes)
    joined, req = gen.concat(sets)
    log(f"batch of {len(sets)} requests, {joined.n} strings")
    o, l = _batch_sort(joined, req)
    _check_requests(joined, req, sets, o, l, oracle_mod)


def test_batch_equals_alone(oracle_mod):
    """Each request's result in the mix equals its result sorted alone."""
    sets = _mixed_requests(77, [3000, 20000, 150, 9000, 1, 45000])
    joined, req = gen.concat(sets)
    o, _ = _batch_sort(joined, req, lcp=False)
    base = 0
    for r, s in enumerate(sets):
        alone, _ = gpu_sort(s, lcp=False)
        assert np.array_equal(o[req[r]:req[r + 1]].astype(np.int64) - base,
                              alone.astype(np.int64))
        base += s.text_bytes


def test_contexts_run_concurrently(oracle_mod):
    """Two contexts sort different batches from two host threads."""
    import torch

    batches = [gen.concat(_mixed_requests(5 + k, [5000, 30000, 700, 12000])) for k in range(2)]
    ctxs = [gss.Context(), gss.Context()]
    results = [None, None]
    errors = []

    def work(k):
        try:
            joined, req = batches[k]
            results[k] = _batch_sort(joined, req, lcp=True, ctx=ctxs[k])
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(2):
        joined, req = batches[k]
        sets = _mixed_requests(5 + k, [5000, 30000, 700, 12000])
        _check_requests(joined, req, sets, results[k][0], results[k][1], oracle_mod)
        st = ctxs[k].stats()
        assert st.n == joined.n
    for c in ctxs:
        c.close()
    torch.cuda.synchronize()


def test_interleaved_shapes_on_one_context(oracle_mod):
    """Single sorts and batches of different shapes back to back on the
    default context: every call re-fills the staged tables and hands its
    counters back through the same host block, so a stale table or a
    hand-over read too early would show up as a wrong order here."""
    plan = [("one", 0, 70000), ("batch", [1, 2, 65, 7000, 1]), ("one", 1, 5000),
            ("batch", [30000, 0, 12, 6145]), ("one", 2, 2), ("one", 3, 90000),
            ("batch", [1] * 9 + [40000]), ("one", 1, 1025), ("batch", [64, 64, 9000, 3]),
            ("one", 0, 6144), ("one", 2, 50000), ("batch", [20000, 20000])]
    for step, item in enumerate(plan):
        if item[0] == "one":
            _, cfg, n = item
            ss = gen.make(cfg, n=n, seed=100 + step)
            log(f"interleaved step {step}: {gen.NAMES[cfg]} n={n}")
            off, lcp = gpu_sort(ss)
            want, want_lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, "msd_ce2", lcp=True)
            assert np.array_equal(off, want), f"step {step}"
            assert np.array_equal(lcp, want_lcp), f"step {step} LCP"
        else:
            sets = _mixed_requests(200 + step, item[1])
            joined, req = gen.concat(sets)
            log(f"interleaved step {step}: batch of {len(sets)}, {joined.n} strings")
            o, l = _batch_sort(joined, req)
            _check_requests(joined, req, sets, o, l, oracle_mod)
