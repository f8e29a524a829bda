"""GPU: the first global pass counts its digits inside its key gather when the
digit is formed as in the last sort on the same context (the same alphabet
and first differing key byte, checked on the device); otherwise a separate
count runs.  Every sort of the sequence below is checked against the oracle
offset for offset and LCP for LCP, and the statistics show which way its
first pass took: a repeated set (guess held), a set with another alphabet, a
set with the same alphabet whose keys first differ at another byte, a set
whose first keys are all equal (no guess left for the next sort), and the
8-bit digit width."""

import numpy as np
import pytest

from gpu_util import assert_matches_oracle, gpu_sort, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402

N = 40_000  # above the largest local class: the sort takes a global pass


def _ab(prefix: int, seed: int):
    """N strings: `prefix` bytes 'a', then 12 random bytes of {a, b}."""
    rng = np.random.default_rng(seed)
    tails = rng.choice(np.frombuffer(b"ab", np.uint8), size=(N, 12))
    return gen.from_strings([b"a" * prefix + t.tobytes() for t in tails])


def _sort_checked(label, ss, oracle_mod, fused):
    log(f"fused count: {label} (n={ss.n})")
    off, lcp = gpu_sort(ss, lcp=True)
    assert_matches_oracle(ss, off, lcp, oracle_mod)
    st = gss.stats()
    assert st.global_passes >= 1, label
    if fused is not None:
        assert st.first_pass_fused == fused, f"{label}: first_pass_fused {st.first_pass_fused}"


def test_fused_first_pass_count(oracle_mod):
    url = gen.url_like(N, seed=5)
    uni = gen.uniform_az(N, seed=6)
    ab0, ab2, ab10 = _ab(0, 7), _ab(2, 8), _ab(10, 9)
    _sort_checked("URL", url, oracle_mod, None)
    _sort_checked("URL again", url, oracle_mod, 1)
    _sort_checked("uniform a-z (other alphabet)", uni, oracle_mod, 0)
    _sort_checked("uniform a-z again", uni, oracle_mod, 1)
    _sort_checked("{a,b} from byte 0 (other alphabet)", ab0, oracle_mod, 0)
    _sort_checked("{a,b} from byte 0 again", ab0, oracle_mod, 1)
    _sort_checked("{a,b} from byte 2 (same alphabet, other byte)", ab2, oracle_mod, 0)
    _sort_checked("{a,b} from byte 2 again", ab2, oracle_mod, 1)
    _sort_checked("first 8 bytes all equal", ab10, oracle_mod, 0)
    _sort_checked("{a,b} from byte 2 after it (no guess)", ab2, oracle_mod, 0)
    _sort_checked("{a,b} from byte 2 again", ab2, oracle_mod, 1)


def test_fused_first_pass_count_8bit(oracle_mod):
    saved = gss.get_tuning()
    try:
        gss.set_tuning(radix_bits=8)
        uni = gen.uniform_az(N, seed=11)
        _sort_checked("8-bit digits, uniform a-z", uni, oracle_mod, None)
        _sort_checked("8-bit digits, uniform a-z again", uni, oracle_mod, 1)
        url = gen.url_like(N, seed=12)
        _sort_checked("8-bit digits, URL", url, oracle_mod, 0)
        _sort_checked("8-bit digits, URL again", url, oracle_mod, 1)
    finally:
        gss.set_tuning(local_threshold=saved.local_threshold, radix_bits=saved.radix_bits,
                       tile=saved.tile, key_bytes=saved.key_bytes)
