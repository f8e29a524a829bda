"""GPU: the four workloads at their full sizes (1M uniform a-z, 10M URL-like,
32M DNA substrings = 3.2 GB of text, 100M Zipf words), checked through
size-independent properties instead of a CPU sort:

* the output is a permutation of the input offsets;
* no adjacent pair is out of order (the oracle's check_result, a restatement
  of the reference's src/util/debug.h:41-70);
* identical strings keep their input order (input offsets increase).

A stable sort's output is unique, so these three make it equal to msd_CE2's
(which is stable) offset for offset.  The LCP array is compared with the
oracle's lcp() of every adjacent pair."""

import numpy as np
import pytest

from gpu_util import device, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402


@pytest.mark.parametrize("config", [0, 1, 2, 3], ids=["uniform_1M", "url_10M", "dna_32M",
                                                      "zipf_100M"])
def test_full_size_workload(config, oracle_mod):
    import torch

    log(f"generating config {config} at full size ...")
    ss = gen.make(config)
    n = ss.n
    assert np.all(np.diff(ss.offsets.astype(np.int64)) > 0), "generator offsets increase"
    log(f"{ss.name}: {n} strings, {ss.text_bytes} bytes; sorting on the GPU")
    dev = device()
    text = torch.from_numpy(ss.text).to(dev)
    off = torch.from_numpy(ss.offsets.view(np.int32)).to(dev)
    out = torch.empty_like(off)
    lcp = torch.empty(n - 1, dtype=torch.int64, device=dev)
    gss.sort_device(text, off, out, lcp, text_bytes=ss.text_bytes)
    torch.cuda.synchronize()
    del text

    # Permutation: the sorted output, sorted numerically, is the input.
    o64 = out.to(torch.int64) & 0xFFFFFFFF
    assert torch.equal(torch.sort(o64).values, off.to(torch.int64) & 0xFFFFFFFF)
    # Stability: identical neighbours (LCP = both lengths) keep input order.
    starts = off.to(torch.int64) & 0xFFFFFFFF
    lens_in = torch.empty_like(starts)
    lens_in[:-1] = starts[1:] - starts[:-1] - 1
    lens_in[-1] = ss.text_bytes - int(starts[-1]) - 1
    lens = lens_in[torch.searchsorted(starts, o64)]
    same = (lcp == lens[:-1]) & (lens[:-1] == lens[1:])
    assert bool(torch.all(o64[1:][same] > o64[:-1][same])), "identical strings reordered"
    log(f"{int(same.sum())} identical neighbour pairs keep their input order")
    del starts, lens_in, lens, same, o64

    got = out.cpu().numpy().view(np.uint32)
    got_lcp = lcp.cpu().numpy().view(np.uint64)
    del out, lcp, off
    torch.cuda.empty_cache()
    ident, inval, wrong = oracle_mod.check_result(ss.text, got)
    assert (ident, inval, wrong) == (0, 0, 0)
    assert np.array_equal(got_lcp, oracle_mod.lcp_array(ss.text, got))
    log(f"{ss.name}: sorted, stable, exact LCPs")
