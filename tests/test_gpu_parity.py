"""GPU parity: the device sort (gss_sort_device32, through the C ABI) against
the oracle's msd_CE2 restatement, offset for offset and LCP for LCP."""

import numpy as np
import pytest

from gpu_util import assert_matches_oracle, gpu_sort, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402


@pytest.fixture(autouse=True)
def default_tuning():
    gss.lib()
    t = gss.get_tuning()
    yield
    gss.set_tuning(t.local_threshold, t.radix_bits, t.tile, t.key_bytes)


@pytest.mark.parametrize("config", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [3000, 60000])
def test_workload_parity(config, n, oracle_mod):
    ss = gen.make(config, n=n)
    log(f"workload {gen.NAMES[config]} n={n}")
    off, lcp = gpu_sort(ss)
    assert_matches_oracle(ss, off, lcp, oracle_mod)


def _set(strs, name):
    return gen.from_strings(strs, name)


EDGE_CASES = {
    "two": [b"b", b"a"],
    "three_equal": [b"aaa"] * 3,
    "all_equal_201": [b"aaa"] * 201,
    "all_equal_9000": [b"aaa"] * 9000,
    "a_bb_bbb": [b"bb", b"a", b"bbb"] * 801,
    "empty_strings": [b"", b"a", b"", b"b", b""] * 50,
    "insertion_known": [b"c", b"", b"bbaaa", b"aaaaa", b"bbbbb"],
    "high_bytes": [bytes([0x41, 0x61, 0x7F, 0x80, 0xC3, 0xFF][i % 6] for i in range(j % 13))
                   + bytes([0xFF - (j % 7)]) for j in range(5000)],
    "key_boundaries": [b"x" * L for L in range(0, 40)] + [b"x" * L + b"y" for L in range(0, 40)],
    "modular_9": [bytes([ord("a") + (i % 30)]) * 9 for i in range(8001)],
    # One local segment whose first sorting key varies in one byte only.
    "one_varying_byte": [b"ppp" + bytes([ord("a") + (i * 7) % 26]) + b"qqqq" + str(i % 997).encode()
                         for i in range(5000)],
    # Two varying bytes, four combinations: sub-buckets far above the
    # counting bound, so the round falls back to the LSD passes.
    "two_bytes_few": [b"pp" + b"ab"[i % 2:i % 2 + 1] + b"qq" + b"xy"[(i // 2) % 2:(i // 2) % 2 + 1]
                      + b"rr" + str((i * 31) % 1009).encode() for i in range(5000)],
    # Two varying bytes, 676 combinations and exact duplicates: counting
    # inside small sub-buckets, ties kept in input order.
    "two_bytes_many": [b"pp" + bytes([ord("a") + (i * 11) % 26]) + b"qq"
                       + bytes([ord("a") + (i * 5 // 26) % 26]) + b"rr" + str(i % 3).encode()
                       for i in range(5000)],
}


def _descending(n, seed, dup_every=0):
    """n strings in non-increasing order: the input on which msd_CE7 and
    msd_CE8 take a descending bucket of 65,536 or more strings for sorted and
    skip its distribution (src/msd_ce.cpp:444-463, 520-542)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(3, 13, size=n)
    strs = [bytes(rng.integers(97, 123, size=int(k)).astype(np.uint8)) for k in lens]
    if dup_every:
        strs = [strs[i - i % dup_every] for i in range(n)]
    return sorted(strs, reverse=True)


EDGE_CASES["reverse_sorted_70000"] = _descending(70_000, 1)
EDGE_CASES["non_increasing_dups_66000"] = _descending(66_000, 2, dup_every=3)
# Every string of the trap sharing one long prefix: a single descending
# bucket of 70,000 strings below the first byte.
EDGE_CASES["reverse_sorted_shared_prefix_70000"] = [b"prefix/" + s
                                                     for s in _descending(70_000, 3)]


@pytest.mark.parametrize("case", sorted(EDGE_CASES))
def test_edge_cases(case, oracle_mod):
    ss = _set(EDGE_CASES[case], case)
    log(f"edge case {case} n={ss.n}")
    off, lcp = gpu_sort(ss)
    assert_matches_oracle(ss, off, lcp, oracle_mod)


@pytest.mark.parametrize("n", [5, 3000, 60000])
def test_unterminated_last_string(n, oracle_mod):
    """The text ends inside the last string (no NUL) at a length that is not
    a multiple of 8, and the bytes after text_bytes are 0xFF: the last
    string is cut at text_bytes, as if a NUL followed, so it orders before
    every string it is a proper prefix of.  Offsets and LCPs against the
    oracle on a NUL-terminated copy."""
    import torch

    rng = np.random.default_rng(n)
    strs = [b"abc" + bytes(rng.integers(97, 100, size=int(rng.integers(0, 9))).astype(np.uint8))
            for _ in range(n - 1)]
    last = b"abc" + b"a" * int(rng.integers(0, 3))
    head = b"".join(s + b"\0" for s in strs)
    if (len(head) + len(last)) % 8 == 0:  # the text must end inside a word
        last += b"a"
    body = head + last
    starts = np.zeros(n, dtype=np.uint32)
    pos = 0
    for i, s in enumerate(strs + [last]):
        starts[i] = pos
        pos += len(s) + 1
    tb = len(body)
    assert tb % 8 != 0 and body[starts[-1]:] == last
    dev = torch.device("cuda:0")
    garbage = np.frombuffer(body + b"\xff" * 24, np.uint8).copy()
    text = torch.from_numpy(garbage).to(dev)
    off = torch.from_numpy(starts.view(np.int32)).to(dev)
    out = torch.empty_like(off)
    lc = torch.empty(max(1, n - 1), dtype=torch.int64, device=dev)
    log(f"unterminated last string, n={n}, text_bytes={tb}")
    gss.sort_device(text, off, out, lc, text_bytes=tb)
    torch.cuda.synchronize()
    clean = np.frombuffer(body + b"\0" * 24, np.uint8).copy()
    want, want_lcp = oracle_mod.sort_offsets(clean, starts, "msd_ce2", lcp=True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
    assert np.array_equal(lc.cpu().numpy().view(np.uint64)[: n - 1], want_lcp)


def test_empty_and_single():
    for strs in ([], [b"only"]):
        ss = _set(strs, "tiny")
        off, lcp = gpu_sort(ss)
        assert list(off) == list(ss.offsets)


def _deep_prefix(n, plen, seed):
    rng = np.random.default_rng(seed)
    pre = b"p" * plen
    return [pre + bytes(rng.integers(97, 100, size=int(rng.integers(0, 6))).astype(np.uint8))
            for _ in range(n)]


@pytest.mark.parametrize("plen", [7, 8, 64, 301])
def test_deep_shared_prefix(plen, oracle_mod):
    ss = _set(_deep_prefix(20000, plen, plen), f"deep{plen}")
    log(f"deep prefix plen={plen}")
    off, lcp = gpu_sort(ss)
    assert_matches_oracle(ss, off, lcp, oracle_mod)


def _first_round_set(n, plen, seed):
    """n strings sharing a prefix of plen bytes, then 0-12 bytes over a
    4-letter alphabet: some end at plen (their NUL is the first differing
    byte), many are duplicates, and with n <= 6144 the whole set is one
    segment of the largest local class at depth 0."""
    rng = np.random.default_rng(seed)
    pre = bytes(rng.integers(97, 123, size=plen).astype(np.uint8))
    out = []
    for _ in range(n):
        k = int(rng.integers(0, 13))
        out.append(pre + bytes(rng.integers(97, 101, size=k).astype(np.uint8)))
    return out


@pytest.mark.parametrize("plen,key_bytes", [(0, 64), (5, 64), (50, 64), (56, 64), (57, 64),
                                            (60, 64), (63, 64), (64, 64), (65, 64), (121, 64),
                                            (20, 32), (24, 32), (25, 32), (31, 32), (33, 32)])
def test_first_round_key_position(plen, key_bytes, oracle_mod):
    """The largest class's first round sorts by the 8 bytes at the group's
    first differing byte G, capped where the window still holds all 8
    (56 for a 64-byte window, 24 for 32); past the window it probes again.
    Shared prefixes around those bounds, strings ending at G and duplicates
    sort offset for offset as msd_CE2, LCPs included."""
    gss.set_tuning(key_bytes=key_bytes)
    ss = _set(_first_round_set(5000, plen, 1000 + plen), f"first{plen}_{key_bytes}")
    log(f"first-round key position plen={plen} key_bytes={key_bytes}")
    off, lcp = gpu_sort(ss)
    assert_matches_oracle(ss, off, lcp, oracle_mod)


@pytest.mark.parametrize("threshold,radix_bits,tile,key_bytes",
                         [(64, 8, 4096, 64), (64, 11, 2048, 32), (1024, 8, 2048, 64),
                          (1024, 11, 4096, 32), (6144, 8, 4096, 32), (6144, 11, 4096, 64),
                          (6144, 11, 2048, 64), (6144, 8, 2048, 32)])
def test_tuning_knobs(threshold, radix_bits, tile, key_bytes, oracle_mod):
    """Every knob of the tuning (local threshold, digit width, tile size,
    the largest class's first-round key window) keeps the output exact."""
    gss.set_tuning(local_threshold=threshold, radix_bits=radix_bits, tile=tile,
                   key_bytes=key_bytes)
    for config in (0, 1, 3):
        ss = gen.make(config, n=40000, seed=7 + config)
        log(f"tuning threshold={threshold} radix_bits={radix_bits} tile={tile} "
            f"key_bytes={key_bytes} workload {ss.name}")
        off, lcp = gpu_sort(ss)
        assert_matches_oracle(ss, off, lcp, oracle_mod)
    ss = _set(_deep_prefix(30000, 40, 3), "deep40")
    off, lcp = gpu_sort(ss)
    assert_matches_oracle(ss, off, lcp, oracle_mod)


def test_suffixes(oracle_mod):
    """Overlapping strings: every suffix of one text (sortstring's
    --suffix-sorting builds these, src/sortstring.c:349-358)."""
    rng = np.random.default_rng(11)
    body = bytes(rng.choice(np.frombuffer(b"ab", np.uint8), size=3000))
    text = np.zeros(len(body) + 1 + gen.PAD, dtype=np.uint8)
    text[: len(body)] = np.frombuffer(body, np.uint8)
    ss = gen.StringSet(text, np.arange(len(body) + 1, dtype=np.uint32), "suffixes")
    off, lcp = gpu_sort(ss)
    assert_matches_oracle(ss, off, lcp, oracle_mod)


def test_repeat_calls_reuse_workspace(oracle_mod):
    for n in (50000, 1000, 70000, 10):
        ss = gen.make(1, n=n, seed=n)
        off, lcp = gpu_sort(ss)
        assert_matches_oracle(ss, off, lcp, oracle_mod)


def _grouped(groups, sub, per, seed):
    """groups x sub groups of `per` strings: every group has its own two
    first bytes, every subgroup its own next two, then 6 random bytes."""
    rng = np.random.default_rng(seed)
    out = []
    for g in range(groups):
        for h in range(sub):
            head = bytes([97 + g % 26, 97 + g // 26, 65 + h % 26, 65 + h // 26])
            tails = rng.integers(97, 123, size=(per, 6)).astype(np.uint8)
            out += [head + bytes(t) for t in tails]
    return [out[i] for i in rng.permutation(len(out))]


def test_largest_class_launched_before_readback(oracle_mod):
    """The largest local class goes out before the counters are read back
    when the last similar sort on the context had work for it: repeated
    sorts (its segments all known after the hinted pass), a sort that needs
    one more pass than the hint (the early launch finds no segment; the
    later pass's segments are launched after it), and a sort with one pass
    fewer than the hint (an empty pass is queued before the early launch).
    Every result equals msd_CE2's with exact LCPs."""
    one_pass = _set(_grouped(12, 1, 3000, 1), "one_pass")   # 12 segments of 3,000 after pass 0
    two_pass = _set(_grouped(4, 3, 3000, 2), "two_pass")    # 4 x 9,000, then 12 x 3,000
    for ss in (one_pass, one_pass, one_pass, two_pass, two_pass, one_pass, one_pass):
        log(f"largest-class early launch: {ss.name}")
        off, lcp = gpu_sort(ss)
        assert_matches_oracle(ss, off, lcp, oracle_mod)
        st = gss.stats()
        assert st.global_passes == (1 if ss.name == "one_pass" else 2)


def test_without_lcp(oracle_mod):
    ss = gen.make(1, n=30000, seed=5)
    off, _ = gpu_sort(ss, lcp=False)
    assert np.array_equal(off, oracle_mod.sort_offsets(ss.text, ss.offsets))


def test_in_place_output(oracle_mod):
    import torch

    ss = gen.make(0, n=20000, seed=3)
    dev = torch.device("cuda:0")
    text = torch.from_numpy(ss.text).to(dev)
    off = torch.from_numpy(ss.offsets.view(np.int32)).to(dev)
    gss.sort_device(text, off, off, None, text_bytes=ss.text_bytes)
    got = off.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, oracle_mod.sort_offsets(ss.text, ss.offsets))
