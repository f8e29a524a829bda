"""Helpers shared by the GPU tests: run the device sort through the C ABI and
compare with the oracle."""

from __future__ import annotations

import sys

import numpy as np


def log(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def device():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch.device("cuda:0")


def gpu_sort(ss, lcp: bool = True):
    """Sorts StringSet `ss` on the GPU (gss_sort_device32).  Returns the sorted
    offsets (uint32) and the LCP array (uint64, n-1 entries) or None."""
    import torch

    import gss

    dev = device()
    n = ss.n
    text = torch.from_numpy(ss.text).to(dev)
    off = torch.from_numpy(ss.offsets.view(np.int32)).to(dev)
    out = torch.empty_like(off)
    lc = torch.empty(max(1, n - 1), dtype=torch.int64, device=dev) if lcp else None
    gss.sort_device(text, off, out, lc, text_bytes=ss.text_bytes)
    torch.cuda.synchronize()
    o = out.cpu().numpy().view(np.uint32)
    l = lc.cpu().numpy().view(np.uint64)[: max(0, n - 1)] if lcp else None
    return o, l


def assert_matches_oracle(ss, off, lcp, oracle):
    """Bit-exact parity with the oracle's msd_CE2 (stable, so the offsets
    match one for one) and the closed-form LCP array."""
    want, want_lcp = oracle.sort_offsets(ss.text, ss.offsets, "msd_ce2", lcp=True)
    if not np.array_equal(off, want):
        bad = np.nonzero(off != want)[0]
        raise AssertionError(f"{ss.name}: {bad.size} offsets differ from msd_CE2, first at "
                             f"{bad[0]}: got {off[bad[0]]}, want {want[bad[0]]}")
    if lcp is not None and not np.array_equal(lcp, want_lcp):
        bad = np.nonzero(lcp != want_lcp)[0]
        raise AssertionError(f"{ss.name}: {bad.size} LCPs differ, first at {bad[0]}: "
                             f"got {lcp[bad[0]]}, want {want_lcp[bad[0]]}")
