"""GPU: the autotuner.  Tuning returns one of the candidates and never picks
something slower than the current tuning on its own sample; with the
autotuner on, gss_sort_device32 sorts with the cached tuning of each input's
signature and the results stay bit-exact against the oracle."""

import numpy as np
import pytest

from gpu_util import device, gpu_sort, log, assert_matches_oracle

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402


def _dev(ss):
    import torch

    return (torch.from_numpy(ss.text).to(device()),
            torch.from_numpy(ss.offsets.view(np.int32)).to(device()))


def test_autotune_picks_a_candidate():
    ss = gen.make(1, n=300_000, seed=3)
    text, off = _dev(ss)
    before = gss.get_tuning()
    r = gss.autotune(text, off, text_bytes=ss.text_bytes)
    log(f"best threshold={r.best.local_threshold} radix={r.best.radix_bits} "
        f"tile={r.best.tile} key_bytes={r.best.key_bytes} "
        f"{r.best_ms:.3f} ms (current tuning {r.default_ms:.3f} ms), sample {r.sample}")
    # 4 thresholds x 2 digit widths, then the other tile size, then the other key window
    assert r.candidates == 10 and r.sample == ss.n
    assert r.best.local_threshold in (1024, 2048, 4096, 6144) and r.best.radix_bits in (8, 11)
    assert r.best.tile in (2048, 4096) and r.best.key_bytes in (32, 64)
    assert 0 < r.best_ms <= r.default_ms
    after = gss.get_tuning()
    key = lambda t: (t.local_threshold, t.radix_bits, t.tile, t.key_bytes)  # noqa: E731
    assert key(after) == key(before)
    assert gss.autotune_cached() >= 1


@pytest.mark.parametrize("config", [0, 1, 2, 3])
def test_sorts_with_autotuner_on_match_oracle(config, oracle_mod):
    gss.autotune_reset()
    gss.autotune_enable(True)
    try:
        ss = gen.make(config, n=120_000, seed=60 + config)
        off, lcp = gpu_sort(ss)  # first sight of this signature: tunes, then sorts
        assert gss.autotune_cached() == 1
        assert_matches_oracle(ss, off, lcp, oracle_mod)
        off2, _ = gpu_sort(ss, lcp=False)  # cached
        assert gss.autotune_cached() == 1
        assert np.array_equal(off, off2)
    finally:
        gss.autotune_enable(False)
        gss.autotune_reset()
