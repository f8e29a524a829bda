"""Pins the CPU oracle (oracle/oracle.c) before results are compared with it:
against the golden vectors produced by the reference's own compiled code,
against a textbook comparison sort, against the reference unit test's known
answers, and against closed forms.  CPU only."""

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import gss.gen as gen
import oracle

GOLDEN = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())


def strings_at(ss, off):
    t = ss.text.tobytes()
    return [t[o:t.index(b"\0", o)] for o in off.tolist()]


def brute_lcp(strs):
    out = []
    for a, b in zip(strs[:-1], strs[1:]):
        i = 0
        while i < len(a) and i < len(b) and a[i] == b[i]:
            i += 1
        out.append(i)
    return out


@pytest.mark.parametrize("w", GOLDEN["workloads"], ids=lambda w: f"{w['name']}-{w['n']}")
def test_oracle_matches_reference_golden(w, oracle_mod):
    ss = gen.make(w["config"], n=w["n"], seed=w["seed"])
    assert ss.md5() == w["input_md5"], "generator output changed"
    off, lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, "msd_ce2", lcp=True)
    assert hashlib.md5(off.astype("<u4").tobytes()).hexdigest() == w["msd_CE2_offsets_md5"]
    strs = strings_at(ss, off)
    assert hashlib.md5(b"\n".join(strs) + b"\n").hexdigest() == w["msd_CE2_bytes_md5"]
    assert hashlib.md5(lcp.astype("<u8").tobytes()).hexdigest() == w["lcp_md5"]
    mk = oracle_mod.sort_offsets(ss.text, ss.offsets, "multikey_cache8")
    mk_strs = strings_at(ss, mk)
    assert hashlib.md5(b"\n".join(mk_strs) + b"\n").hexdigest() == w["multikey_cache8_bytes_md5"]


@pytest.mark.parametrize("t", GOLDEN["tiny"], ids=lambda t: f"tiny{t['config']}")
def test_oracle_tiny_fixtures(t, oracle_mod):
    strs = [s.encode("latin-1") for s in t["input"]]
    ss = gen.from_strings(strs)
    off, lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, "msd_ce2", lcp=True)
    index = {int(o): i for i, o in enumerate(ss.offsets.tolist())}
    assert [index[int(o)] for o in off.tolist()] == t["msd_CE2_order"]
    assert lcp.tolist() == t["lcp"]


def _random_strings(rng, n, alphabet, maxlen):
    return [bytes(rng.choice(alphabet, size=int(rng.integers(0, maxlen + 1))).astype(np.uint8))
            for _ in range(n)]


@pytest.mark.parametrize("alg", ["msd_ce2", "multikey_cache8"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_equals_textbook_sort(alg, seed, oracle_mod):
    """Unsigned bytewise order with NUL as terminator is exactly Python's bytes
    order; msd_CE2 is stable, so its order equals a stable sort's."""
    rng = np.random.default_rng(seed)
    alphabets = [np.array([0x41, 0x61, 0x7F, 0x80, 0xC3, 0xFF], np.uint8),
                 np.arange(1, 256, dtype=np.uint8), np.array([97, 98], np.uint8)]
    for alpha in alphabets:
        strs = _random_strings(rng, 3000, alpha, 12)
        ss = gen.from_strings(strs)
        off, lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, alg, lcp=True)
        got = strings_at(ss, off)
        assert got == sorted(strs)
        if alg == "msd_ce2":
            idx = {int(o): i for i, o in enumerate(ss.offsets.tolist())}
            order = [idx[int(o)] for o in off.tolist()]
            assert order == sorted(range(len(strs)), key=lambda i: (strs[i], i))
        assert lcp.tolist() == brute_lcp(got)


def _ptr_array(strs):
    import ctypes as C

    bufs = [C.create_string_buffer(s) for s in strs]
    arr = (C.c_void_p * len(bufs))(*[C.addressof(b) for b in bufs])
    return bufs, arr


def test_insertion_sort_known_answers(oracle_mod):
    """unit-test/main.cpp:160-205"""
    import ctypes as C

    lib = oracle_mod.lib()
    bufs, arr = _ptr_array([b"c", b"", b"bbaaa", b"aaaaa", b"bbbbb"])
    for _ in range(2):
        lib.or_insertion_sort(arr, 5, 0)
        assert [C.string_at(p) for p in arr] == [b"", b"aaaaa", b"bbaaa", b"bbbbb", b"c"]
    bufs, arr = _ptr_array([b"bbbb1", b"aaaa3", b"aaaa2"])
    lib.or_insertion_sort(arr, 3, 4)
    assert [C.string_at(p) for p in arr] == [b"bbbb1", b"aaaa2", b"aaaa3"]
    bufs, arr = _ptr_array([b""])
    lib.or_insertion_sort(arr, 1, 0)
    lib.or_insertion_sort(arr, 0, 0)


def test_unit_test_families(oracle_mod):
    """unit-test/main.cpp:220-270 known answers for the routine families."""
    for k in (1, 201, 801):
        ss = gen.from_strings([b"bb", b"a", b"bbb"] * k)
        got = strings_at(ss, oracle_mod.sort_offsets(ss.text, ss.offsets))
        assert got == [b"a"] * k + [b"bb"] * k + [b"bbb"] * k
    for k in (1, 2001, 8001):
        ss = gen.from_strings([bytes([ord("a") + (i % 30)]) * 9 for i in range(k)])
        off = oracle_mod.sort_offsets(ss.text, ss.offsets)
        assert oracle_mod.check_result(ss.text, off) == (0, 0, 0)


def test_key8_packing(oracle_mod):
    """multikey_cache.cpp:107-125: big-endian, stop after NUL, zero pad; no
    sign extension of bytes >= 0x80 (get_char.h:82 would)."""
    assert oracle_mod.key8(b"abc") == 0x6162630000000000
    assert oracle_mod.key8(b"") == 0
    assert oracle_mod.key8(b"\xff" * 8) == 0xFFFFFFFFFFFFFFFF
    assert oracle_mod.key8(b"\x80" * 9) == 0x8080808080808080
    assert oracle_mod.key8(b"abcdefghij", 3) == int.from_bytes(b"defghij", "big") << 8
    assert oracle_mod.key8(b"abcdefgh") & 0xFF == ord("h")
    assert oracle_mod.key8(b"abcdefg") & 0xFF == 0


def test_lcp_closed_form(oracle_mod):
    """mergesort_lcp.cpp:88-100: equal bytes before the first mismatch or NUL;
    identical strings give their length."""
    lib = oracle_mod.lib()
    assert lib.or_lcp(b"abc", b"abd") == 2
    assert lib.or_lcp(b"abc", b"abc") == 3
    assert lib.or_lcp(b"", b"") == 0
    assert lib.or_lcp(b"ab", b"abc") == 2
    assert lib.or_lcp(b"\xff\x01", b"\xff\x02") == 1


def test_check_result_detects_planted_errors(oracle_mod):
    """debug.h:41-70 counts inversions, identical pointers and NULLs; a
    mistake planted in the oracle's own output must be caught."""
    ss = gen.make(0, n=5000, seed=1)
    off = oracle_mod.sort_offsets(ss.text, ss.offsets)
    assert oracle_mod.check_result(ss.text, off) == (0, 0, 0)
    bad = off.copy()
    bad[[10, 11]] = bad[[11, 10]]
    assert oracle_mod.check_result(ss.text, bad)[2] >= 1
    dup = off.copy()
    dup[20] = dup[21]
    assert oracle_mod.check_result(ss.text, dup)[0] == 1
    for n in (0, 1):
        assert oracle_mod.check_result(ss.text, off[:n]) == (0, 0, 0)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("config", [0, 1, 2, 3])
def test_restatement_equals_reference_build(config, oracle_mod):
    """Where the reference compiles (this container), its msd_CE2 and the
    restatement agree offset for offset on each workload."""
    ss = gen.make(config, n=100000, seed=99 + config)
    want = oracle_mod.ref_sort_offsets(ss.text, ss.offsets, "msd_CE2")
    assert np.array_equal(oracle_mod.sort_offsets(ss.text, ss.offsets), want)
    for name in ("msd_CE6", "msd_ci_adaptive", "msd_A2_adaptive", "multikey_cache8"):
        other = oracle_mod.ref_sort_offsets(ss.text, ss.offsets, name)
        assert strings_at(ss, other) == strings_at(ss, want), name


# ---------------------------------------------------------------------------
# The steps either side of the sort: ingest, --write, LCP merge.

def _delim_text(rng, n, delim, trailing=b"tail"):
    lines = _random_strings(rng, n, np.frombuffer(b"abcxyz\x01\xff", np.uint8), 12)
    lines = [l.replace(bytes([delim]), b"") for l in lines] + [b"", b""]
    return np.frombuffer(bytes([delim]).join(lines) + bytes([delim]) + trailing, np.uint8)


@pytest.mark.parametrize("delim", [ord("\n"), 0])
def test_ingest_closed_form(delim, oracle_mod):
    """Offsets are 0 and every delimiter position + 1 except the last; the
    count is the number of delimiters; delimiters become NUL, other bytes
    are untouched; bytes after the last delimiter form no string."""
    rng = np.random.default_rng(5)
    text = _delim_text(rng, 300, delim)
    off, t2 = oracle_mod.ingest(text, delim)
    pos = np.nonzero(text == delim)[0]
    assert off.shape[0] == pos.shape[0]
    assert np.array_equal(off, np.concatenate([[0], pos[:-1] + 1]).astype(np.uint32))
    want = text.copy()
    want[pos] = 0
    assert np.array_equal(t2, want)
    assert oracle_mod.ingest(np.frombuffer(b"no delimiter", np.uint8), delim)[0].shape[0] == 0


def test_write_format(oracle_mod):
    """Each string then '\\n', in array order; a string also ends at the end
    of the text."""
    ss = gen.from_strings([b"bb", b"", b"a\xffz", b"bb"])
    order = ss.offsets[[3, 1, 2, 0]]
    assert oracle_mod.write(ss.text, order) == b"bb\n\na\xffz\nbb\n"
    t = np.frombuffer(b"abcdef", np.uint8)
    assert oracle_mod.write(t, np.array([4, 0], np.uint32)) == b"ef\nabcdef\n"
    assert oracle_mod.write(t, np.array([4, 0], np.uint32), length=5) == b"e\nabcde\n"


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_merge_lcp_equals_brute_force(seed, oracle_mod):
    """Merging two sorted runs equals the stable sort of A then B, and the
    LCPs equal lcp() of every adjacent output pair; the LCP shortcuts are
    exercised by long shared prefixes and cross-run duplicates."""
    rng = np.random.default_rng(seed)
    base = _random_strings(rng, 200, np.frombuffer(b"ab", np.uint8), 10)
    strs = [b"prefix/" * int(rng.integers(0, 3)) + s for s in base]
    ss = gen.from_strings(strs + strs[:50])  # duplicates across the runs
    cut = int(rng.integers(0, ss.n))
    a_in, b_in = ss.offsets[:cut], ss.offsets[cut:]
    a, al = oracle_mod.sort_offsets(ss.text, a_in, lcp=True)
    b, bl = oracle_mod.sort_offsets(ss.text, b_in, lcp=True)
    out, ol = oracle_mod.merge_lcp(ss.text, a, al, b, bl)
    want = oracle_mod.sort_offsets(ss.text, np.concatenate([a, b]))
    assert np.array_equal(out, want)
    got = strings_at(ss, out)
    assert got == sorted(got)
    assert ol.tolist() == brute_lcp(got)
    for x, y in ((a[:0], b), (a, b[:0]), (a[:1], b[:1])):
        o, l = oracle_mod.merge_lcp(ss.text, x, al[: max(0, len(x) - 1)], y,
                                    bl[: max(0, len(y) - 1)])
        assert np.array_equal(o, oracle_mod.sort_offsets(ss.text, np.concatenate([x, y])))
        assert l.tolist() == brute_lcp(strings_at(ss, o))


SORTSTRING = Path(oracle.__file__).resolve().parent / "_ref" / "sortstring"


@pytest.mark.skipif(not SORTSTRING.exists(), reason="oracle/_ref not built")
@pytest.mark.parametrize("mode", ["plain", "raw", "suffix"])
def test_reference_sortstring_equals_oracle_pipeline(mode, tmp_path, oracle_mod):
    """The reference's own sortstring (compiled from its sources) reads the
    file, sorts with msd_CE2 and writes it (--write); the oracle's ingest,
    msd_CE2 and write give the same bytes.  Pins ingest and write against
    the reference, for newline input, --raw input and --suffix-sorting."""
    import subprocess

    rng = np.random.default_rng(11)
    if mode == "suffix":
        data = bytes(rng.choice(np.frombuffer(b"abn\n", np.uint8), size=300).astype(np.uint8))
    else:
        delim = 0 if mode == "raw" else ord("\n")
        data = _delim_text(rng, 400, delim).tobytes()
    f = tmp_path / "in.txt"
    f.write_bytes(data)
    out = tmp_path / "out.txt"
    args = [str(SORTSTRING), f"--write={out}"]
    args += {"plain": [], "raw": ["--raw"], "suffix": ["--suffix-sorting"]}[mode]
    # sortstring appends to ./sortstring_log: run it in the test's directory
    subprocess.run(args + ["msd_CE2", str(f)], check=True, capture_output=True, timeout=120,
                   cwd=tmp_path)
    text = np.frombuffer(data + b"\0" * 16, np.uint8).copy()
    if mode == "suffix":
        off = np.arange(len(data), dtype=np.uint32)
        t2 = text
    else:
        off, t2 = oracle_mod.ingest(text[: len(data)], 0 if mode == "raw" else ord("\n"))
        t2 = np.concatenate([t2, np.zeros(16, np.uint8)])
    srt = oracle_mod.sort_offsets(t2, off)
    assert oracle_mod.write(t2, srt, length=len(data)) == out.read_bytes()
