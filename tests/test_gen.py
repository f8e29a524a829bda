"""Workload generators: deterministic and shaped as specified.  CPU only."""

import numpy as np

import gss.gen as gen


def lengths(ss):
    ends = np.nonzero(ss.text[: ss.text_bytes] == 0)[0]
    return ends - ss.offsets.astype(np.int64)


def test_deterministic():
    for c in range(4):
        assert gen.make(c, n=3000).md5() == gen.make(c, n=3000).md5()
        assert gen.make(c, n=3000).md5() != gen.make(c, n=3000, seed=1).md5()


def test_uniform_az_shape():
    ss = gen.make(0, n=20000)
    l = lengths(ss)
    assert l.min() >= 5 and l.max() <= 20
    body = ss.text[: ss.text_bytes]
    body = body[body != 0]
    assert body.min() >= ord("a") and body.max() <= ord("z")


def test_url_like_shape():
    ss = gen.make(1, n=50000)
    l = lengths(ss)
    assert l.min() >= 25 and l.max() <= 100
    strs = ss.strings()
    dup_frac = 1 - len(set(strs)) / len(strs)
    assert 0.25 < dup_frac < 0.35
    allowed = set(gen.A40.tolist())
    assert set(ss.text[: ss.text_bytes].tolist()) <= allowed | {0}


def test_dna_shape():
    ss = gen.make(2, n=10000)
    assert (lengths(ss) == 100).all()
    assert set(ss.text[: ss.text_bytes].tolist()) == set(b"ACGT") | {0}


def test_zipf_shape():
    ss = gen.make(3, n=100000)
    l = lengths(ss)
    assert l.min() >= 1 and l.max() <= 15
    strs = ss.strings()
    top = max(set(strs), key=strs.count) if len(strs) < 5000 else None
    counts = {}
    for s in strs:
        counts[s] = counts.get(s, 0) + 1
    assert max(counts.values()) > 0.05 * len(strs)  # heavy head (rank-1 mass ~6%)
    del top


def test_newline_file_roundtrip():
    ss = gen.make(1, n=1000)
    lines = ss.newline_bytes().split(b"\n")
    assert lines[-1] == b"" and lines[:-1] == ss.strings()


def test_from_strings():
    ss = gen.from_strings([b"b", b"", b"a"])
    assert ss.strings() == [b"b", b"", b"a"]
    assert ss.offsets.tolist() == [0, 2, 3]
