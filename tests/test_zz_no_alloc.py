"""GPU: the allocation-free LCP merge (gss_merge_lcp_device32_noalloc, through
the C ABI) against the allocating merge it varies (gss_merge_lcp_device32)
and the oracle, offset for offset and LCP for LCP, on two calls in a row --
the second with a smaller input, so it runs over the first call's leftover
split table in the shared scratch array."""

import ctypes as C

import numpy as np
import pytest

from gpu_util import device, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device())


def _sorted_run(text, off, text_bytes):
    import torch

    o = _dev(off.view(np.int32))
    out = torch.empty_like(o)
    lc = torch.empty(max(1, o.numel() - 1), dtype=torch.int64, device=device())
    gss.sort_device(text, o, out, lc, text_bytes=text_bytes)
    return out, lc


def _noalloc_merge(text, tb, a, al, b, bl):
    """One call of the allocation-free entry on torch's current stream; the
    result is read after the stream's work is complete."""
    import torch

    na, nb = a.numel(), b.numel()
    out = torch.full((max(1, na + nb),), -1, dtype=torch.int32, device=device())
    ol = torch.full((max(1, na + nb - 1),), -1, dtype=torch.int64, device=device())
    st = torch.cuda.current_stream(device()).cuda_stream
    rc = gss.lib().gss_merge_lcp_device32_noalloc(
        text.data_ptr(), tb, a.data_ptr() if na else None, al.data_ptr() if na > 1 else None, na,
        b.data_ptr() if nb else None, bl.data_ptr() if nb > 1 else None, nb, out.data_ptr(),
        ol.data_ptr(), C.c_void_p(st))
    assert rc == 0, gss.last_error()
    torch.cuda.synchronize()
    return out[: na + nb], ol[: max(0, na + nb - 1)]


def test_noalloc_merge_equals_allocating_merge(oracle_mod):
    for call, (na, nb, seed) in enumerate([(700_000, 500_000, 11), (30_001, 9_000, 12)]):
        parts = [gen.make(c, n=(na + nb) // 4 + 1, seed=seed + c) for c in range(4)]
        joined, _ = gen.concat(parts)
        rng = np.random.default_rng(seed)
        pick = rng.integers(0, joined.n, size=na + nb)
        pick[na: na + nb // 4] = pick[: nb // 4]  # duplicates across the runs
        a_in, b_in = joined.offsets[pick[:na]], joined.offsets[pick[na:]]
        log(f"no-alloc merge call {call}: na={na} nb={nb} (text {joined.text_bytes} B)")
        text = _dev(joined.text)
        a, al = _sorted_run(text, a_in, joined.text_bytes)
        b, bl = _sorted_run(text, b_in, joined.text_bytes)
        got, got_lcp = _noalloc_merge(text, joined.text_bytes, a[:na], al, b[:nb], bl)
        ref, ref_lcp = gss.merge_lcp(text, a[:na], al[: na - 1], b[:nb], bl[: nb - 1],
                                     text_bytes=joined.text_bytes)
        g = got.cpu().numpy().view(np.uint32)
        gl = got_lcp.cpu().numpy().view(np.uint64)
        assert np.array_equal(g, ref.cpu().numpy().view(np.uint32))
        assert np.array_equal(gl, ref_lcp.cpu().numpy().view(np.uint64))
        want, want_lcp = oracle_mod.sort_offsets(joined.text, np.concatenate([a_in, b_in]),
                                                 lcp=True)
        assert np.array_equal(g, want)
        assert np.array_equal(gl, want_lcp)
