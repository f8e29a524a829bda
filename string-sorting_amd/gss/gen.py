"""Seeded synthetic string sets for the sort's workloads.

Every generator returns a :class:`StringSet`: one text buffer in which each
string is followed by a NUL byte (the layout ``sortstring`` builds from a
newline file, src/sortstring.c:310-335), plus the 32-bit offset of every
string in input order.  All randomness comes from ``numpy.random.Generator``
(PCG64) seeded explicitly, so the same call always yields the same bytes.

Workloads (N = default size):

* ``uniform_az``  N=1,000,000; lengths uniform in [5, 20]; bytes i.i.d.
  uniform over 'a'..'z'.
* ``url_like``    N=10,000,000; 2,000 prefixes of length uniform in [20, 60]
  over [a-z0-9./:-]; each string is, with probability 0.3, an exact copy of a
  uniformly chosen earlier string, otherwise a uniformly chosen prefix plus a
  suffix of length uniform in [5, 40] over the same alphabet.
* ``dna``         N=32,000,000; substrings of length 100 starting at uniform
  positions of one 10,000,000-byte i.i.d. text over {A,C,G,T}.
* ``zipf_words``  N=100,000,000; draws from a 2,000,000-word vocabulary by
  Zipf(s=1.1) rank; word lengths min(15, Geometric(mean 7)) >= 1 over 'a'..'z'.
  The vocabulary is NOT de-duplicated: equal words at different ranks simply
  pool their probability.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np

A26 = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz", dtype=np.uint8)
A40 = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz0123456789./:-", dtype=np.uint8)
ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)

# Default seeds and sizes of the four plain workloads.
SEEDS = {0: 20240, 1: 20241, 2: 20242, 3: 20243}
SIZES = {0: 1_000_000, 1: 10_000_000, 2: 32_000_000, 3: 100_000_000}
NAMES = {0: "uniform_az", 1: "url_like", 2: "dna", 3: "zipf_words"}

# Bytes of zero padding after the last string (device loads read whole words).
PAD = 16


@dataclass
class StringSet:
    text: np.ndarray      # uint8, NUL after every string, PAD zero bytes at the end
    offsets: np.ndarray   # uint32 offsets of the strings, input order
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.offsets.shape[0])

    @property
    def text_bytes(self) -> int:
        """Bytes of string data incl. terminators (excludes the padding)."""
        return int(self.text.shape[0] - PAD)

    def strings(self) -> list:
        """The strings as ``bytes`` (small sets only)."""
        t = self.text.tobytes()
        return [t[o:t.index(b"\0", o)] for o in self.offsets.tolist()]

    def md5(self) -> str:
        """Digest of the text and offsets: pins the generator's output."""
        h = hashlib.md5()
        h.update(self.text[: self.text_bytes].tobytes())
        h.update(self.offsets.astype("<u4").tobytes())
        return h.hexdigest()

    def newline_bytes(self) -> bytes:
        """The set as a newline-terminated file, the format sortstring reads."""
        t = self.text[: self.text_bytes].copy()
        t[t == 0] = ord("\n")
        return t.tobytes()


def _excl_cumsum(a: np.ndarray) -> np.ndarray:
    out = np.empty_like(a, dtype=np.int64)
    if a.size:
        out[0] = 0
        np.cumsum(a[:-1], out=out[1:])
    return out


def from_strings(strs, name: str = "") -> StringSet:
    """Builds a StringSet from an iterable of bytes without NUL bytes."""
    strs = [bytes(s) for s in strs]
    for s in strs:
        if b"\0" in s:
            raise ValueError("strings must not contain NUL bytes")
    lens = np.array([len(s) + 1 for s in strs], dtype=np.int64)
    offs = _excl_cumsum(lens)
    total = int(lens.sum())
    if total + PAD >= 2**32:
        raise ValueError("text too large for 32-bit offsets")
    text = np.zeros(total + PAD, dtype=np.uint8)
    if total:
        text[:total] = np.frombuffer(b"\0".join(strs) + b"\0", dtype=np.uint8)
    return StringSet(text, offs.astype(np.uint32), name)


def _segment_gather(src: np.ndarray, starts: np.ndarray, lens: np.ndarray) -> np.ndarray:
    """Concatenation of src[starts[i]:starts[i]+lens[i]] over all i."""
    total = int(lens.sum())
    dst = _excl_cumsum(lens)
    idx = np.arange(total, dtype=np.int64) - np.repeat(dst - starts, lens)
    return src[idx]


def _assemble(n: int, lens: np.ndarray, fill, name: str, chunk: int = 1 << 20) -> StringSet:
    """Allocates the text for strings of length lens[i] and lets
    fill(a, b, out) write strings a..b-1 (with terminators) into out."""
    rec = lens.astype(np.int64) + 1
    offs = _excl_cumsum(rec)
    total = int(rec.sum())
    if total + PAD >= 2**32:
        raise ValueError("text too large for 32-bit offsets")
    text = np.zeros(total + PAD, dtype=np.uint8)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        lo = int(offs[a])
        hi = int(offs[b]) if b < n else total
        fill(a, b, text[lo:hi])
    return StringSet(text, offs.astype(np.uint32), name)


def uniform_az(n: int = SIZES[0], seed: int = SEEDS[0], lmin: int = 5, lmax: int = 20) -> StringSet:
    rng = np.random.default_rng(seed)
    lens = rng.integers(lmin, lmax + 1, size=n, dtype=np.int64)
    source = np.concatenate([A26[rng.integers(0, 26, size=int(lens.sum()), dtype=np.int64)],
                             np.zeros(1, np.uint8)])
    choff = _excl_cumsum(lens)
    zero = source.size - 1

    def fill(a, b, out):
        starts = np.stack([choff[a:b], np.full(b - a, zero)], 1).ravel()
        ls = np.stack([lens[a:b], np.ones(b - a, np.int64)], 1).ravel()
        out[:] = _segment_gather(source, starts, ls)

    return _assemble(n, lens, fill, "uniform_az")


def url_like(n: int = SIZES[1], seed: int = SEEDS[1], nprefix: int = 2000,
             plen=(20, 60), slen=(5, 40), dup: float = 0.3) -> StringSet:
    rng = np.random.default_rng(seed)
    plens = rng.integers(plen[0], plen[1] + 1, size=nprefix, dtype=np.int64)
    ptext = A40[rng.integers(0, 40, size=int(plens.sum()), dtype=np.int64)]
    poff = _excl_cumsum(plens)
    isdup = rng.random(n) < dup
    if n:
        isdup[0] = False
    src = (rng.random(n) * np.arange(n)).astype(np.int64)  # uniform earlier index
    pid = rng.integers(0, nprefix, size=n, dtype=np.int64)
    sl = rng.integers(slen[0], slen[1] + 1, size=n, dtype=np.int64)
    sl_own = np.where(isdup, 0, sl)
    stext = A40[rng.integers(0, 40, size=int(sl_own.sum()), dtype=np.int64)]
    soff = _excl_cumsum(sl_own)
    root = np.where(isdup, src, np.arange(n, dtype=np.int64))
    while True:
        m = isdup[root]
        if not m.any():
            break
        root[m] = src[root[m]]
    rp = pid[root]
    lens = plens[rp] + sl[root]
    source = np.concatenate([ptext, stext, np.zeros(1, np.uint8)])
    base_s = ptext.size
    zero = source.size - 1

    def fill(a, b, out):
        r = root[a:b]
        p = rp[a:b]
        starts = np.stack([poff[p], base_s + soff[r], np.full(b - a, zero)], 1).ravel()
        ls = np.stack([plens[p], sl[r], np.ones(b - a, np.int64)], 1).ravel()
        out[:] = _segment_gather(source, starts, ls)

    return _assemble(n, lens, fill, "url_like")


def dna(n: int = SIZES[2], seed: int = SEEDS[2], base_len: int = 10_000_000, length: int = 100) -> StringSet:
    rng = np.random.default_rng(seed)
    base = ACGT[rng.integers(0, 4, size=base_len, dtype=np.int64)]
    pos = rng.integers(0, base_len - length + 1, size=n, dtype=np.int64)
    total = n * (length + 1)
    if total + PAD >= 2**32:
        raise ValueError("text too large for 32-bit offsets")
    text = np.zeros(total + PAD, dtype=np.uint8)
    grid = text[:total].reshape(n, length + 1)
    win = np.lib.stride_tricks.sliding_window_view(base, length)
    step = 1 << 20
    for a in range(0, n, step):
        b = min(n, a + step)
        grid[a:b, :length] = win[pos[a:b]]
    offs = (np.arange(n, dtype=np.int64) * (length + 1)).astype(np.uint32)
    return StringSet(text, offs, "dna")


def zipf_words(n: int = SIZES[3], seed: int = SEEDS[3], vocab: int = 2_000_000,
               s: float = 1.1, mean_len: float = 7.0, max_len: int = 15) -> StringSet:
    rng = np.random.default_rng(seed)
    wl = np.minimum(rng.geometric(1.0 / mean_len, size=vocab), max_len).astype(np.int64)
    wtext = A26[rng.integers(0, 26, size=int(wl.sum()), dtype=np.int64)]
    woff = _excl_cumsum(wl)
    cdf = np.cumsum(np.arange(1, vocab + 1, dtype=np.float64) ** -s)
    cdf /= cdf[-1]
    ranks = np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), vocab - 1)
    lens = wl[ranks]
    source = np.concatenate([wtext, np.zeros(1, np.uint8)])
    zero = source.size - 1

    def fill(a, b, out):
        r = ranks[a:b]
        starts = np.stack([woff[r], np.full(b - a, zero)], 1).ravel()
        ls = np.stack([wl[r], np.ones(b - a, np.int64)], 1).ravel()
        out[:] = _segment_gather(source, starts, ls)

    return _assemble(n, lens, fill, "zipf_words")


def concat(sets, name: str = "batch"):
    """One text holding several sets one after another; returns the joined
    StringSet and the request table (set r owns offsets [req[r], req[r+1]))."""
    parts, offs, req = [], [], [0]
    base = 0
    for s in sets:
        parts.append(s.text[: s.text_bytes])
        offs.append(s.offsets.astype(np.int64) + base)
        base += s.text_bytes
        req.append(req[-1] + s.n)
    if base + PAD >= 2**32:
        raise ValueError("text too large for 32-bit offsets")
    text = np.zeros(base + PAD, dtype=np.uint8)
    if parts:
        text[:base] = np.concatenate(parts)
    off = np.concatenate(offs).astype(np.uint32) if offs else np.zeros(0, np.uint32)
    return StringSet(text, off, name), req


GENERATORS = {0: uniform_az, 1: url_like, 2: dna, 3: zipf_words}


def make(config: int, n: int | None = None, seed: int | None = None) -> StringSet:
    """Workload `config` (0..3) at its default size and seed unless given."""
    gen = GENERATORS[config]
    return gen(SIZES[config] if n is None else n, SEEDS[config] if seed is None else seed)


# ---------------------------------------------------------------------------
# Request stream: many independent sorts of uneven size.
#
# Payloads come from a seeded pool per generator: pool g is GENERATORS[g] at
# POOL_N strings and its default seed.  Request i draws its generator id
# uniformly from 0..3, its size as floor(exp(U[ln lo, ln hi])) and a uniform
# start in its pool; its strings are the pool's strings [start, start+size)
# in pool order.  Arrivals are a Poisson process: exponential gaps at `rate`
# requests per second.

STREAM_SEED = 20244
STREAM_REQUESTS = 20_000
POOL_N = 4_000_000
STREAM_SIZES = (1_000, 4_000_000)


@dataclass
class RequestStream:
    pools: StringSet        # the four pools joined (gen.concat order 0..3)
    pool_begin: list        # pool g owns offsets [pool_begin[g], pool_begin[g+1])
    gen_id: np.ndarray      # int64 per request
    start: np.ndarray       # int64 start inside the pool
    size: np.ndarray        # int64 strings per request

    @property
    def nreq(self) -> int:
        return int(self.size.shape[0])

    def first(self, i: int) -> int:
        """Index of request i's first string in the joined offset array."""
        return int(self.pool_begin[int(self.gen_id[i])] + self.start[i])

    def offsets(self, i: int) -> np.ndarray:
        a = self.first(i)
        return self.pools.offsets[a:a + int(self.size[i])]


def request_stream(nreq: int = STREAM_REQUESTS, seed: int = STREAM_SEED, pool_n: int = POOL_N,
                   sizes=STREAM_SIZES, pools: StringSet | None = None,
                   pool_begin=None) -> RequestStream:
    """The first `nreq` requests of the seeded stream (pools are built unless
    given, so several streams can share them)."""
    if pools is None:
        pools, pool_begin = concat([GENERATORS[g](pool_n, SEEDS[g]) for g in range(4)], "pools")
    lo, hi = sizes
    hi = min(hi, pool_n)
    rng = np.random.default_rng(seed)
    m = max(nreq, STREAM_REQUESTS)  # drawn in full, so a shorter stream is a prefix
    gen_id = rng.integers(0, 4, size=m, dtype=np.int64)[:nreq]
    size = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=m)))[:nreq].astype(np.int64)
    size = np.clip(size, 1, pool_n)
    start = np.floor(rng.random(m)[:nreq] * (pool_n - size + 1)).astype(np.int64)
    return RequestStream(pools, list(pool_begin), gen_id, start, size)


def arrivals(nreq: int, rate: float, seed: int = STREAM_SEED) -> np.ndarray:
    """Poisson arrival times (seconds from the first arrival) at `rate`/s;
    rate = inf gives all requests at time 0."""
    if not np.isfinite(rate):
        return np.zeros(nreq)
    rng = np.random.default_rng(seed + 1)
    gaps = rng.exponential(1.0 / rate, size=nreq)
    gaps[0] = 0.0
    return np.cumsum(gaps)
