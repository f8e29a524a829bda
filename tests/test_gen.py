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


def test_request_stream_is_seeded_and_prefix_stable():
    """The stream's request table: sizes log-uniform in [1e3, pool], ids 0..3,
    every slice inside its pool; a shorter stream is a prefix of the full one."""
    pools, begin = gen.concat([gen.GENERATORS[g](3000, gen.SEEDS[g]) for g in range(4)], "pools")
    a = gen.request_stream(nreq=50, pool_n=3000, sizes=(10, 3000), pools=pools, pool_begin=begin)
    b = gen.request_stream(nreq=400, pool_n=3000, sizes=(10, 3000), pools=pools, pool_begin=begin)
    assert np.array_equal(a.size, b.size[:50]) and np.array_equal(a.gen_id, b.gen_id[:50])
    assert np.array_equal(a.start, b.start[:50])
    assert b.size.min() >= 10 and b.size.max() <= 3000
    assert set(np.unique(b.gen_id)) == {0, 1, 2, 3}
    assert np.all(b.start >= 0) and np.all(b.start + b.size <= 3000)
    for i in range(b.nreq):
        o = b.offsets(i)
        g = int(b.gen_id[i])
        assert o.shape[0] == b.size[i]
        assert np.array_equal(o, pools.offsets[begin[g] + b.start[i]: begin[g] + b.start[i] + b.size[i]])
    # log-uniform: the median size is near the geometric mean of the range
    med = np.median(b.size)
    assert 100 < med < 300
    t = gen.arrivals(1000, 500.0)
    assert t[0] == 0 and np.all(np.diff(t) >= 0)
    assert abs(t[-1] / 999 - 1 / 500.0) < 0.2 / 500.0
    assert np.all(gen.arrivals(5, float("inf")) == 0)
