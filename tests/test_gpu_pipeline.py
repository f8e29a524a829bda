"""GPU: the steps either side of the sort -- device ingest of a delimited
text, suffix offsets, --write formatting and the LCP-aware merge of sorted
runs -- each bit-exact against the oracle (whose ingest and write are pinned
against the reference's own sortstring, and whose merge against brute
force), and the whole file -> ingest -> sort -> write pipeline on the device
against the oracle's."""

import hashlib
import os

import numpy as np
import pytest

from gpu_util import device, log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device())


def _text_tensor(data: bytes, pad=16):
    return _dev(np.frombuffer(data + b"\0" * pad, np.uint8).copy())


def _edge_text(delim):
    d = bytes([delim])
    lines = [b"bb", b"", b"ab", b"", b"", b"a" * 40, b"\x01\xff", b"zz", b"ab"]
    return d.join(lines) + d + b"tail-without-delimiter"


@pytest.mark.parametrize("delim", [ord("\n"), 0])
@pytest.mark.parametrize("kind", ["edge", "url_like", "dna", "long_lines"])
def test_ingest_matches_oracle(delim, kind, oracle_mod):
    if kind == "edge":
        data = _edge_text(delim)
    elif kind == "long_lines":  # lines spanning many 16 KiB ingest chunks
        rng = np.random.default_rng(3)
        parts = [bytes(rng.integers(97, 123, size=int(rng.integers(0, 70000)), dtype=np.uint8))
                 for _ in range(12)]
        data = bytes([delim]).join(parts) + bytes([delim])
    else:
        ss = gen.make({"url_like": 1, "dna": 2}[kind], n=200_000, seed=7)
        t = ss.text[: ss.text_bytes].copy()
        t[t == 0] = delim
        data = t.tobytes()
    log(f"ingest {kind} delim={delim}: {len(data)} bytes")
    text = _text_tensor(data)
    off = gss.ingest(text, delim, text_bytes=len(data))
    want_off, want_text = oracle_mod.ingest(np.frombuffer(data, np.uint8), delim)
    assert off.numel() == want_off.shape[0]
    assert np.array_equal(off.cpu().numpy().view(np.uint32), want_off)
    assert np.array_equal(text.cpu().numpy()[: len(data)], want_text)
    assert not text.cpu().numpy()[len(data):].any(), "bytes past the text were written"


def test_ingest_capacity(oracle_mod):
    """gss_ingest_device32 through the C ABI: max_strings one short gives
    GSS_E_TOO_LARGE with the count, and leaves the text and the offsets as
    they were; the exact capacity splits the text as the oracle does."""
    import ctypes as C
    import torch

    rng = np.random.default_rng(9)
    lines = [bytes(rng.integers(97, 123, int(k)).astype(np.uint8)) for k in rng.integers(0, 90, 3000)]
    data = b"\n".join(lines) + b"\n"
    want_off, want_text = oracle_mod.ingest(np.frombuffer(data, np.uint8), ord("\n"))
    n = len(want_off)
    lib = gss.lib()
    cnt = C.c_uint64()
    text = _text_tensor(data)
    before = text.clone()
    off = torch.full((n,), -1, dtype=torch.int32, device=device())
    rc = lib.gss_ingest_device32(text.data_ptr(), len(data), ord("\n"), off.data_ptr(), n - 1,
                                 C.byref(cnt), None)
    assert rc == -4 and cnt.value == n  # GSS_E_TOO_LARGE
    assert torch.equal(text, before) and bool((off == -1).all())
    rc = lib.gss_ingest_device32(text.data_ptr(), len(data), ord("\n"), off.data_ptr(), n,
                                 C.byref(cnt), None)
    assert rc == 0 and cnt.value == n
    assert np.array_equal(off.cpu().numpy().view(np.uint32), want_off)
    assert text.cpu().numpy().tobytes()[: len(data)] == bytes(want_text[: len(data)])


def test_ingest_without_delimiter_is_an_error():
    text = _text_tensor(b"no newline here")
    with pytest.raises(gss.GssError, match="unable to read any lines"):
        gss.ingest(text, b"\n", text_bytes=15)


@pytest.mark.parametrize("config", [0, 1, 2, 3])
def test_write_matches_oracle(config, oracle_mod):
    import torch

    ss = gen.make(config, n=50_000, seed=21)
    text = _dev(ss.text)
    off = _dev(ss.offsets.view(np.int32))
    out = torch.empty_like(off)
    gss.sort_device(text, off, out, text_bytes=ss.text_bytes)
    got = gss.write(text, out, text_bytes=ss.text_bytes).cpu().numpy().tobytes()
    srt = oracle_mod.sort_offsets(ss.text, ss.offsets)
    assert got == oracle_mod.write(ss.text, srt, length=ss.text_bytes)


def test_write_edge_cases(oracle_mod):
    # strings of every length 0..40 around the 16-byte output blocks, and a
    # string cut by the end of the text
    strs = [b"x" * k for k in range(41)] + [b"", b"", b"q"]
    ss = gen.from_strings(strs)
    text = _dev(ss.text)
    rng = np.random.default_rng(1)
    order = ss.offsets[rng.permutation(ss.n)]
    got = gss.write(text, _dev(order.view(np.int32)), text_bytes=ss.text_bytes)
    assert got.cpu().numpy().tobytes() == oracle_mod.write(ss.text, order, length=ss.text_bytes)
    raw = np.frombuffer(b"abcdefghij", np.uint8)
    got = gss.write(_text_tensor(raw.tobytes()), _dev(np.array([7, 2], np.int32)), text_bytes=10)
    assert got.cpu().numpy().tobytes() == b"hij\ncdefghij\n"


def test_write_capacity_long_strings(oracle_mod):
    """gss_write_device32 through the C ABI: the size query (no output), an
    output one byte too small (GSS_E_TOO_LARGE with the exact size, nothing
    written), and the exact size; strings of up to 3,000 bytes, so that some
    blocks of 256 strings overflow the write kernel's 32 KiB stage and take
    its direct path, among short ones that do not."""
    import ctypes as C
    import torch

    rng = np.random.default_rng(5)
    lens = np.where(rng.random(1500) < 0.3, rng.integers(500, 3000, 1500), rng.integers(0, 40, 1500))
    strs = [bytes(rng.integers(97, 123, int(k)).astype(np.uint8)) for k in lens]
    ss = gen.from_strings(strs)
    text = _dev(ss.text)
    order = ss.offsets[rng.permutation(ss.n)]
    off = _dev(order.view(np.int32))
    want = oracle_mod.write(ss.text, order, length=ss.text_bytes)
    lib = gss.lib()
    size = C.c_uint64()
    assert lib.gss_write_device32(text.data_ptr(), ss.text_bytes, off.data_ptr(), ss.n, None, 0,
                                  C.byref(size), None) == 0
    assert size.value == len(want)
    out = torch.full((len(want) + 32,), 0xAB, dtype=torch.uint8, device=device())
    rc = lib.gss_write_device32(text.data_ptr(), ss.text_bytes, off.data_ptr(), ss.n,
                                out.data_ptr(), len(want) - 1, C.byref(size), None)
    assert rc == -4 and size.value == len(want)  # GSS_E_TOO_LARGE
    assert bool((out == 0xAB).all())
    rc = lib.gss_write_device32(text.data_ptr(), ss.text_bytes, off.data_ptr(), ss.n,
                                out.data_ptr(), len(want), C.byref(size), None)
    assert rc == 0 and size.value == len(want)
    got = out.cpu().numpy().tobytes()
    assert got[: len(want)] == want and got[len(want):] == b"\xab" * 32


@pytest.mark.parametrize("kind", ["plain", "raw"])
def test_file_pipeline_on_device(kind, oracle_mod):
    """sortstring's path on the device: the file moves once, strings are cut
    there, sorted and written; equals the oracle's ingest + msd_CE2 + write."""
    import torch

    delim = 0 if kind == "raw" else ord("\n")
    ss = gen.make(1, n=300_000, seed=31)
    t = ss.text[: ss.text_bytes].copy()
    t[t == 0] = delim
    data = t.tobytes() + b"trailing"
    text = _text_tensor(data)
    off = gss.ingest(text, delim, text_bytes=len(data))
    out = torch.empty_like(off)
    gss.sort_device(text, off, out, text_bytes=len(data))
    got = gss.write(text, out, text_bytes=len(data)).cpu().numpy().tobytes()
    want_off, want_text = oracle_mod.ingest(np.frombuffer(data, np.uint8), delim)
    want_text = np.concatenate([want_text, np.zeros(16, np.uint8)])
    want = oracle_mod.write(want_text, oracle_mod.sort_offsets(want_text, want_off),
                            length=len(data))
    assert hashlib.md5(got).hexdigest() == hashlib.md5(want).hexdigest()


@pytest.mark.parametrize("text", [b"banana\n", b"mississippi" * 3, b"ab" * 700 + b"b",
                                  b"a" * 2000], ids=["banana", "mississippi", "abab", "a2000"])
def test_suffix_sort_equals_naive(text, oracle_mod):
    """Suffix mode: every suffix of the text (each runs to the text's end),
    sorted on the device, equals a naive suffix sort; LCPs equal lcp()."""
    import torch

    t = _text_tensor(text)
    off = gss.suffixes(len(text), device())
    out = torch.empty_like(off)
    lc = torch.empty(max(1, len(text) - 1), dtype=torch.int64, device=device())
    gss.sort_device(t, off, out, lc, text_bytes=len(text))
    got = out.cpu().numpy().tolist()
    assert got == sorted(range(len(text)), key=lambda i: text[i:])
    suf = [text[i:] for i in got]
    want_lcp = [len(os.path.commonprefix([a, b])) for a, b in zip(suf, suf[1:])]
    assert lc.cpu().numpy()[: len(text) - 1].tolist() == want_lcp


def test_suffix_sort_dna_text(oracle_mod):
    """Suffixes of a 200 kB random ACGT text, against the oracle's msd_CE2."""
    import torch

    rng = np.random.default_rng(9)
    data = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=200_000).astype(np.uint8))
    text = np.frombuffer(data + b"\0" * 16, np.uint8)
    t = _dev(text)
    off = gss.suffixes(len(data), device())
    out = torch.empty_like(off)
    lc = torch.empty(len(data) - 1, dtype=torch.int64, device=device())
    gss.sort_device(t, off, out, lc, text_bytes=len(data))
    want, want_lcp = oracle_mod.sort_offsets(text, np.arange(len(data), dtype=np.uint32),
                                             lcp=True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
    assert np.array_equal(lc.cpu().numpy().view(np.uint64), want_lcp)


def _sorted_run(text, off, text_bytes):
    import torch

    o = _dev(off.view(np.int32))
    out = torch.empty_like(o)
    lc = torch.empty(max(1, o.numel() - 1), dtype=torch.int64, device=device())
    gss.sort_device(text, o, out, lc, text_bytes=text_bytes)
    return out, lc


@pytest.mark.parametrize("sizes", [(60_000, 45_000), (0, 5_000), (5_000, 0), (1, 1), (1, 9_000),
                                   (100_003, 2_048)])
def test_merge_lcp_matches_oracle(sizes, oracle_mod):
    """Two sorted runs of one text (drawn from all four generators, with
    duplicates across the runs) merge into the stable sort of A then B with
    exact LCPs."""
    na, nb = sizes
    parts = [gen.make(c, n=30_000, seed=40 + c) for c in range(4)]
    joined, _ = gen.concat(parts)
    rng = np.random.default_rng(na * 7 + nb)
    pick = rng.integers(0, joined.n, size=na + nb)
    pick[na: na + min(na, nb) // 3] = pick[: min(na, nb) // 3]  # cross-run duplicates
    a_in = joined.offsets[pick[:na]]
    b_in = joined.offsets[pick[na:]]
    text = _dev(joined.text)
    a, al = _sorted_run(text, a_in, joined.text_bytes)
    b, bl = _sorted_run(text, b_in, joined.text_bytes)
    out, ol = gss.merge_lcp(text, a[:na], al[: max(0, na - 1)], b[:nb], bl[: max(0, nb - 1)],
                            text_bytes=joined.text_bytes)
    got = out.cpu().numpy().view(np.uint32)
    want, want_lcp = oracle_mod.sort_offsets(joined.text, np.concatenate([a_in, b_in]), lcp=True)
    assert np.array_equal(got, want)
    assert np.array_equal(ol.cpu().numpy().view(np.uint64), want_lcp)
    ao, alo = oracle_mod.sort_offsets(joined.text, a_in, lcp=True)
    bo, blo = oracle_mod.sort_offsets(joined.text, b_in, lcp=True)
    mo, mlo = oracle_mod.merge_lcp(joined.text, ao, alo, bo, blo)
    assert np.array_equal(got, mo) and np.array_equal(ol.cpu().numpy().view(np.uint64), mlo)


def _prefix_family(n, seed):
    """Strings that are prefixes of one another, share 0-300 bytes, repeat
    exactly, or are empty: the cases where a merge must tell "shorter" from
    "smaller" and "equal" from "longer"."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.choice([0, 1, 7, 8, 9, 31, 32, 33, 150, 300]))
        tail = bytes(rng.integers(97, 99, size=int(rng.integers(0, 4))).astype(np.uint8))
        out.append(b"p" * k + tail)
    out += out[: n // 3]
    return gen.from_strings([out[i] for i in rng.permutation(len(out))], "prefix_family")


@pytest.mark.parametrize("k", [2, 3, 8])
def test_merge_lcp_prefix_family(k, oracle_mod):
    """Runs of strings that are prefixes of each other and exact duplicates,
    spread over merge blocks of their own and of mixed runs: the k-way (and,
    for k = 2, the 2-way) merge equals the stable sort of the concatenation
    with exact LCPs."""
    ss = _prefix_family(12_000, 90 + k)
    text = _dev(ss.text)
    bounds = np.linspace(0, ss.n, k + 1).astype(np.int64)
    runs, lcps = [], []
    for r in range(k):
        o, lc = _sorted_run(text, ss.offsets[bounds[r]:bounds[r + 1]], ss.text_bytes)
        runs.append(o)
        lcps.append(lc[: max(0, o.numel() - 1)])
    log(f"prefix-family merge k={k} n={ss.n}")
    want, want_lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, lcp=True)
    out, ol = gss.merge_lcp_kway(text, runs, lcps, text_bytes=ss.text_bytes)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
    assert np.array_equal(ol.cpu().numpy().view(np.uint64), want_lcp)
    if k == 2:
        o2, l2 = gss.merge_lcp(text, runs[0], lcps[0], runs[1], lcps[1], text_bytes=ss.text_bytes)
        assert np.array_equal(o2.cpu().numpy().view(np.uint32), want)
        assert np.array_equal(l2.cpu().numpy().view(np.uint64), want_lcp)


@pytest.mark.parametrize("sizes", [(7_000,), (9_000, 0, 4_000), (3_000, 3_000, 3_000),
                                   (1, 20_000, 0, 5, 2_500), (800,) * 8, (0, 0)])
def test_merge_lcp_kway_matches_oracle(sizes, oracle_mod):
    """k sorted runs (k = 1..8, uneven and empty ones, duplicates across
    runs) merge into the stable sort of their concatenation with exact LCPs:
    equal strings keep run order, as merge_lcp_3way's ties do
    (src/mergesort_lcp.cpp:399-709).  The k = 2 case equals the 2-way merge."""
    import torch

    parts = [gen.make(c, n=20_000, seed=70 + c) for c in range(4)]
    joined, _ = gen.concat(parts)
    rng = np.random.default_rng(sum(sizes) + len(sizes))
    total = sum(sizes)
    pick = rng.integers(0, joined.n, size=max(1, total))[:total]
    if total > 20:
        pick[total // 2: total // 2 + total // 10] = pick[: total // 10]  # cross-run duplicates
    text = _dev(joined.text)
    runs, lcps, ins = [], [], []
    start = 0
    for n in sizes:
        part = joined.offsets[pick[start:start + n]]
        ins.append(part)
        start += n
        if n:
            o, lc = _sorted_run(text, part, joined.text_bytes)
            runs.append(o[:n])
            lcps.append(lc[: max(0, n - 1)])
        else:
            runs.append(torch.empty(0, dtype=torch.int32, device=device()))
            lcps.append(torch.empty(0, dtype=torch.int64, device=device()))
    log(f"k-way merge sizes={sizes}")
    out, ol = gss.merge_lcp_kway(text, runs, lcps, text_bytes=joined.text_bytes)
    got = out.cpu().numpy().view(np.uint32)
    assert got.size == total
    if total == 0:
        return
    want, want_lcp = oracle_mod.sort_offsets(joined.text, np.concatenate(ins), lcp=True)
    assert np.array_equal(got, want)
    assert np.array_equal(ol.cpu().numpy().view(np.uint64), want_lcp)
    if len(sizes) == 2:
        o2, l2 = gss.merge_lcp(text, runs[0], lcps[0], runs[1], lcps[1],
                               text_bytes=joined.text_bytes)
        assert np.array_equal(o2.cpu().numpy(), out.cpu().numpy())


def _big_tie(n, seed):
    """Strings that share their first 12 bytes, with short suffixes over a
    three-letter alphabet and repeats: one tie of every run, far over the
    local capacity, sorted by global passes from byte 8."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        tail = bytes(rng.integers(97, 100, size=int(rng.integers(0, 9))).astype(np.uint8))
        out.append(b"abcdefghijkl" + tail)
    out += out[: n // 4]
    out += [b"abcdefg", b"abcdefgh", b"abc"] * 50  # keys with a NUL next to the tie
    return gen.from_strings([out[i] for i in rng.permutation(len(out))], "big_tie")


def _coded(n, ncodes, seed):
    """n strings of an 8-letter code (ncodes of them) and two letters of
    a-c: about n / ncodes strings per code, so a run of n / k strings has
    some ncodes groups of 8-byte keys, and every code ties across runs."""
    from gss.gen import PAD, StringSet

    rng = np.random.default_rng(seed)
    code = rng.integers(0, ncodes, size=n)
    rec = np.zeros((n, 11), dtype=np.uint8)
    for q in range(8):
        rec[:, 7 - q] = 97 + (code // 26 ** q) % 26
    rec[:, 8:10] = rng.integers(97, 100, size=(n, 2))
    text = np.concatenate([rec.ravel(), np.zeros(PAD, np.uint8)])
    return StringSet(text, (np.arange(n, dtype=np.uint64) * 11).astype(np.uint32), "coded")


@pytest.mark.parametrize("case", ["url8", "uniform4", "zipf3", "mixed17", "prefix5", "bigtie3",
                                  "tree3", "runs70", "tiny64"])
def test_merge_kway_key_groups(case, oracle_mod):
    """Three to 64 runs merge by 8-byte key groups: groups from the runs' LCP
    arrays, a stable merge of their keys, and a sort from byte 8 of only the
    spans where groups of different runs share a key.  The output equals the
    stable sort of the runs' concatenation, LCPs included -- with ties
    everywhere (URL prefix groups) and almost nowhere (uniform strings),
    short words whose keys hold a NUL (Zipf), empty and uneven runs with
    duplicates across runs, strings that are prefixes of each other, and a
    tie far over the local capacity.  About 300,000 groups merge in 2-way
    rounds of their keys, fewer by ranks in one launch; more than n / 8
    groups (uniform, Zipf, mixed) take a sort of the runs side by side; 70
    runs take the 2-way string merge rounds."""
    import torch

    cuts = None
    if case == "url8":
        ss, k = gen.url_like(400_000, seed=81), 8
    elif case == "uniform4":
        ss, k = gen.uniform_az(200_000, seed=85), 4
    elif case == "zipf3":
        ss, k = gen.zipf_words(300_000, seed=82), 3
    elif case == "mixed17":
        parts = [gen.make(c, n=80_000, seed=83 + c) for c in range(4)]
        ss, _ = gen.concat(parts)
        rng = np.random.default_rng(83)
        cuts = np.sort(rng.integers(0, ss.n, size=16))
        cuts[3] = cuts[4]  # an empty run
        k = 17
    elif case == "prefix5":
        ss, k = _prefix_family(220_000, 84), 5
    elif case == "bigtie3":
        ss, k = _big_tie(40_000, 86), 3
    elif case == "tree3":  # ~300,000 groups (over the one-launch limit), ~n/10
        ss, k = _coded(3_000_000, 100_000, 88), 3
    elif case == "tiny64":  # 64 runs of 1-4 strings (runs of one have no LCPs), ties and NULs
        words = [b"abcdefghij", b"abcdefghik", b"abcdefgh", b"abc", b"abcdefghij", b""]
        rng = np.random.default_rng(89)
        ss = gen.from_strings([words[int(i)] for i in rng.integers(0, len(words), 160)])
        k = 64
        cuts = np.sort(rng.choice(np.arange(1, ss.n), size=63, replace=False))
    else:
        parts = [gen.make(c, n=5_000, seed=87 + c) for c in range(4)]
        ss, _ = gen.concat(parts)
        k = 70
    if cuts is None:
        cuts = np.linspace(0, ss.n, k + 1).astype(np.int64)[1:-1]
    bounds = [0] + [int(c) for c in cuts] + [ss.n]
    text = _dev(ss.text)
    offs = ss.offsets.copy()
    if case in ("mixed17", "runs70"):  # the same strings again, in another run
        offs[ss.n // 2: ss.n // 2 + ss.n // 10] = offs[: ss.n // 10]
    runs, lcps, ins = [], [], []
    for r in range(k):
        part = offs[bounds[r]:bounds[r + 1]]
        ins.append(part)
        n = part.size
        if n:
            o, lc = _sorted_run(text, part, ss.text_bytes)
            runs.append(o[:n])
            lcps.append(lc[: max(0, n - 1)])
        else:
            runs.append(torch.empty(0, dtype=torch.int32, device=device()))
            lcps.append(torch.empty(0, dtype=torch.int64, device=device()))
    log(f"k-way merge by key groups {case}: k={k} n={ss.n}")
    out, ol = gss.merge_lcp_kway(text, runs, lcps, text_bytes=ss.text_bytes)
    want, want_lcp = oracle_mod.sort_offsets(ss.text, np.concatenate(ins), lcp=True)
    got = out.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, want)
    assert np.array_equal(ol.cpu().numpy().view(np.uint64), want_lcp)
