"""GPU: seeded random string sets of varied structure -- alphabets of 1 to
255 byte values, shared prefixes of up to 70 bytes, a quarter exact
duplicates, empty strings, 0 to 20,000 strings -- through the sort with
LCPs, the k-way merge of 2 to 70 sorted runs (empty ones included, so every
merge path: the 2-way rounds, the key groups with ties and without, the sort
of the runs side by side) and --write of the merged order, each bit-exact
against the oracle (the stable sort of the runs' concatenation for the
merge)."""

import numpy as np
import pytest

from gpu_util import device, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402


def _case(seed):
    rng = np.random.default_rng(7000 + seed)
    n = int(rng.choice([0, 1, 2, 3, 10, 100, 1000, 5000, 20000]))
    asz = int(rng.choice([1, 2, 4, 26, 200, 255]))
    alpha = rng.choice(np.arange(1, 256), size=asz, replace=False).astype(np.uint8)
    prefixes = [bytes(alpha[rng.integers(0, asz, int(rng.integers(0, 70)))])
                for _ in range(int(rng.integers(1, 50)))]
    out = []
    for _ in range(n):
        if out and rng.random() < 0.25:
            out.append(out[int(rng.integers(0, len(out)))])
            continue
        p = prefixes[int(rng.integers(0, len(prefixes)))] if rng.random() < 0.7 else b""
        out.append(p + bytes(alpha[rng.integers(0, asz, int(rng.integers(0, 30)))]))
    k = int(rng.choice([2, 3, 4, 8, 17, 64, 70]))
    return rng, gen.from_strings(out), k


@pytest.mark.parametrize("seed", range(48))
def test_random_sort_merge_write(seed, oracle_mod):
    import torch

    rng, ss, k = _case(seed)
    n = ss.n
    log(f"random case {seed}: n={n} k={k}")
    if n == 0:
        return
    dev = device()
    text = torch.from_numpy(ss.text).to(dev)
    o = torch.from_numpy(ss.offsets.view(np.int32).copy()).to(dev)
    out = torch.empty_like(o)
    lc = torch.empty(max(1, n), dtype=torch.int64, device=dev)
    gss.sort_device(text, o, out, lc, text_bytes=ss.text_bytes)
    want, want_lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, lcp=True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
    assert np.array_equal(lc.cpu().numpy()[: n - 1].view(np.uint64), want_lcp)

    offs = ss.offsets[rng.permutation(n)]
    cuts = np.sort(rng.integers(0, n + 1, size=k - 1))
    bounds = [0] + [int(c) for c in cuts] + [n]
    runs, lcps, ins = [], [], []
    for r in range(k):
        part = offs[bounds[r]:bounds[r + 1]]
        ins.append(part)
        m = part.size
        if m:
            po = torch.from_numpy(part.view(np.int32).copy()).to(dev)
            pout = torch.empty_like(po)
            pl = torch.empty(max(1, m), dtype=torch.int64, device=dev)
            gss.sort_device(text, po, pout, pl, text_bytes=ss.text_bytes)
            runs.append(pout)
            lcps.append(pl[: max(0, m - 1)])
        else:
            runs.append(torch.empty(0, dtype=torch.int32, device=dev))
            lcps.append(torch.empty(0, dtype=torch.int64, device=dev))
    mo, ml = gss.merge_lcp_kway(text, runs, lcps, text_bytes=ss.text_bytes)
    want, want_lcp = oracle_mod.sort_offsets(ss.text, np.concatenate(ins), lcp=True)
    assert np.array_equal(mo.cpu().numpy().view(np.uint32), want)
    assert np.array_equal(ml.cpu().numpy()[: n - 1].view(np.uint64), want_lcp)
    got = gss.write(text, mo, text_bytes=ss.text_bytes).cpu().numpy().tobytes()
    assert got == oracle_mod.write(ss.text, want, length=ss.text_bytes)
