"""GPU: the served request loop (gss_server_*).  Requests of uneven size
arrive while others are on the GPU, are folded into batches and must each
come out exactly as when sorted alone (gss_sort_device32) and as the oracle's
msd_CE2 sorts them, LCP arrays included."""

import time

import numpy as np
import pytest

from gpu_util import device, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402

POOL_N = 150_000


@pytest.fixture(scope="module")
def stream():
    return gen.request_stream(nreq=48, seed=4242, pool_n=POOL_N, sizes=(1_000, POOL_N))


@pytest.fixture(scope="module")
def pools_dev(stream):
    import torch

    dev = device()
    text = torch.from_numpy(stream.pools.text).to(dev)
    off = torch.from_numpy(stream.pools.offsets.view(np.int32)).to(dev)
    return text, off


def _budget_for(slots, cap_strings):
    """A budget whose per-slot batch capacity is roughly cap_strings."""
    return slots * cap_strings * 110


def _alone(text, off, a, n, text_bytes):
    import torch

    out = torch.empty(n, dtype=torch.int32, device=text.device)
    lc = torch.empty(max(1, n - 1), dtype=torch.int64, device=text.device)
    gss.sort_device(text, off[a:a + n], out, lc, text_bytes=text_bytes)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32), lc.cpu().numpy().view(np.uint64)[: max(0, n - 1)]


def test_served_stream_matches_alone_and_oracle(stream, pools_dev, oracle_mod):
    import torch

    text, off = pools_dev
    tb = stream.pools.text_bytes
    rng = np.random.default_rng(7)
    reqs = [(stream.first(i), int(stream.size[i])) for i in range(stream.nreq)]
    # Edge sizes: empty, single string, a pair, one past the local threshold.
    reqs += [(0, 0), (5, 1), (9, 2), (stream.pool_begin[1] + 11, 8193)]
    with gss.Server(text, tb, hbm_budget=_budget_for(3, 400_000), slots=3,
                    max_batch_requests=16) as srv:
        cap = srv.stats().batch_capacity
        log(f"served stream: {len(reqs)} requests, batch capacity {cap}")
        assert cap < sum(n for _, n in reqs), "capacity must force several batches"
        outs, lcps, ids = [], [], []
        for k, (a, n) in enumerate(reqs):
            outs.append(torch.full((max(1, n),), -1, dtype=torch.int32, device=text.device))
            lcps.append(torch.full((max(1, n - 1),), -7, dtype=torch.int64, device=text.device)
                        if k % 2 == 1 else None)
        # The server's streams do not wait for torch's: the fills above must
        # be done before a request is handed over.
        torch.cuda.synchronize()
        for k, (a, n) in enumerate(reqs):
            ids.append(srv.submit(off[a:a + n], outs[k][:n], lcps[k]))
            if k >= 20:  # the first 20 arrive at once, the rest spread out
                time.sleep(float(rng.uniform(0, 0.002)))
        lat = [srv.wait(i) for i in ids]
        st = srv.stats()
        log(f"batches {st.batches}, max requests per batch {st.max_batch_requests_seen}, "
            f"p50 {st.p50_ms:.3f} ms, p99 {st.p99_ms:.3f} ms")
        assert st.completed == len(reqs) and st.submitted == len(reqs)
        assert st.max_batch_requests_seen > 1, "queued requests must be folded into one batch"
        assert st.batches < len(reqs)
        assert all(x >= 0 for x in lat)
    for k, (a, n) in enumerate(reqs):
        if n == 0:
            continue
        got = outs[k].cpu().numpy().view(np.uint32)[:n]
        alone, alone_lcp = _alone(text, off, a, n, tb)
        assert np.array_equal(got, alone), f"request {k} (n={n}) differs from its solo sort"
        want, want_lcp = oracle_mod.sort_offsets(stream.pools.text, stream.pools.offsets[a:a + n],
                                                 "msd_ce2", lcp=True)
        assert np.array_equal(got, want), f"request {k} (n={n}) differs from msd_CE2"
        if lcps[k] is not None and n > 1:
            gl = lcps[k].cpu().numpy().view(np.uint64)[: n - 1]
            assert np.array_equal(gl, want_lcp), f"request {k} LCP"
            assert np.array_equal(gl, alone_lcp)


def test_served_admission_by_budget(stream, pools_dev):
    import torch

    text, off = pools_dev
    with gss.Server(text, stream.pools.text_bytes, hbm_budget=_budget_for(1, 20_000), slots=1,
                    max_batch_requests=4) as srv:
        cap = srv.stats().batch_capacity
        assert 1_000 < cap < 60_000
        out = torch.empty(cap + 1, dtype=torch.int32, device=text.device)
        with pytest.raises(gss.GssError, match="exceeds the batch capacity"):
            srv.submit(off[: cap + 1], out)
        rid = srv.submit(off[:cap], out[:cap])
        srv.wait(rid)
        alone, _ = _alone(text, off, 0, cap, stream.pools.text_bytes)
        assert np.array_equal(out[:cap].cpu().numpy().view(np.uint32), alone)
        with pytest.raises(gss.GssError):
            srv.wait(rid)  # already released


def test_served_close_finishes_queued_requests(stream, pools_dev):
    import torch

    text, off = pools_dev
    tb = stream.pools.text_bytes
    srv = gss.Server(text, tb, hbm_budget=_budget_for(2, 200_000), slots=2, max_batch_requests=8)
    reqs = [(stream.first(i), int(stream.size[i])) for i in range(12)]
    outs = []
    for a, n in reqs:
        o = torch.empty(n, dtype=torch.int32, device=text.device)
        outs.append(o)
        srv.submit(off[a:a + n], o)
    srv.close()  # no wait: closing drains the queue
    for (a, n), o in zip(reqs, outs):
        alone, _ = _alone(text, off, a, n, tb)
        assert np.array_equal(o.cpu().numpy().view(np.uint32), alone)


def _delay(stream_, ms):
    """Queues about `ms` milliseconds of device work on `stream_` (a bounded
    spin kernel where torch has one, else elementwise passes)."""
    import torch

    with torch.cuda.stream(stream_):
        if hasattr(torch.cuda, "_sleep"):
            torch.cuda._sleep(int(ms * 2.0e6))  # cycles at about 2 GHz
        else:
            x = torch.ones(1 << 26, device=device())
            for _ in range(max(1, int(ms))):
                x.mul_(1.0000001)


def test_served_waits_for_ready_event(stream, pools_dev, oracle_mod):
    """Stream order through the ready event: each request's offsets are
    written on another torch stream behind a delay, the event is recorded
    there and the request is submitted at once, with no host
    synchronization.  The results must equal the oracle's.  The offset
    buffers first hold valid offsets (all 0), so a server that ignored the
    event would sort the wrong strings, never read out of bounds."""
    import torch

    text, off = pools_dev
    tb = stream.pools.text_bytes
    reqs = [(stream.first(i), int(stream.size[i])) for i in range(10)]
    side = torch.cuda.Stream(device())
    bufs = [torch.zeros(n, dtype=torch.int32, device=text.device) for _, n in reqs]
    outs = [torch.full((n,), -1, dtype=torch.int32, device=text.device) for _, n in reqs]
    torch.cuda.synchronize()  # the zero fills above are done; nothing below waits
    with gss.Server(text, tb, hbm_budget=_budget_for(2, 300_000), slots=2,
                    max_batch_requests=4) as srv:
        ids, events = [], []
        for k, (a, n) in enumerate(reqs):
            _delay(side, 20 if k < 3 else 2)
            with torch.cuda.stream(side):
                bufs[k].copy_(off[a:a + n])
            ev = torch.cuda.Event()
            ev.record(side)
            events.append(ev)
            ids.append(srv.submit(bufs[k], outs[k], ready=ev))
        log(f"ready events: {len(ids)} requests submitted without a host sync")
        for i in ids:
            srv.wait(i)
    torch.cuda.synchronize()
    for k, (a, n) in enumerate(reqs):
        want = oracle_mod.sort_offsets(stream.pools.text, stream.pools.offsets[a:a + n], "msd_ce2")
        got = outs[k].cpu().numpy().view(np.uint32)
        assert np.array_equal(got, want), f"request {k} (n={n}) ignored its ready event"


def test_request_stream_sample_at_scale(oracle_mod):
    """The served request stream at its stated scale: pools of 4,000,000
    strings per generator and requests of 1e3..4e6 strings.  A seeded sample
    of the first 20,000 requests, always including the largest, is served
    (several slots, batches folded) and every request is checked against the
    oracle's msd_CE2, LCP arrays included."""
    import torch

    rs = gen.request_stream()
    big = int(np.argmax(rs.size))
    rng = np.random.default_rng(99)
    pick = [big] + [int(i) for i in rng.choice(rs.nreq, size=23, replace=False) if i != big]
    sizes = [int(rs.size[i]) for i in pick]
    log(f"stream sample: {len(pick)} requests, largest {sizes[0]} strings, total {sum(sizes)}")
    dev = device()
    text = torch.from_numpy(rs.pools.text).to(dev)
    off = torch.from_numpy(rs.pools.offsets.view(np.int32)).to(dev)
    tb = rs.pools.text_bytes
    outs = [torch.full((n,), -1, dtype=torch.int32, device=dev) for n in sizes]
    lcps = [torch.full((max(1, n - 1),), -1, dtype=torch.int64, device=dev) for n in sizes]
    torch.cuda.synchronize()
    with gss.Server(text, tb, hbm_budget=16 << 30, slots=4, max_batch_requests=1024) as srv:
        ids = [srv.submit(off[rs.first(i):rs.first(i) + n], outs[k], lcps[k])
               for k, (i, n) in enumerate(zip(pick, sizes))]
        for r in ids:
            srv.wait(r)
        st = srv.stats()
        log(f"served {st.completed} requests in {st.batches} batches")
    for k, (i, n) in enumerate(zip(pick, sizes)):
        a = rs.first(i)
        want, want_lcp = oracle_mod.sort_offsets(rs.pools.text, rs.pools.offsets[a:a + n],
                                                 "msd_ce2", lcp=True)
        assert np.array_equal(outs[k].cpu().numpy().view(np.uint32), want), f"request {i}"
        assert np.array_equal(lcps[k].cpu().numpy().view(np.uint64)[: max(0, n - 1)],
                              want_lcp), f"request {i} LCP"
        log(f"request {i} (generator {int(rs.gen_id[i])}, {n} strings) equals msd_CE2")
