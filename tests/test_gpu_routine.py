"""GPU: the routine.h drop-in (gss_msd / gss_msd_lcp) driven as the reference
drives its routines -- by the reference's own unit test and CLI compiled from
its sources with the GPU routine registered (oracle/_ref), and through the
pointer ABI from Python with the unit test's string families."""

import ctypes as C
import hashlib
import mmap
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu

import gss  # noqa: E402
import gss.gen as gen  # noqa: E402

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref"


def test_unit_test_families():
    """unit-test/main.cpp:220-270, through the C pointer ABI (each string a
    separate allocation, as strdup gives)."""
    r = gss.routine_from_name("gss_msd")
    for k in range(1, 2000, 200):
        assert r.f([b"aaa"] * k) == [b"aaa"] * k
    for k in range(1, 1000, 200):
        out = r.f([b"bb", b"a", b"bbb"] * k)
        assert out == [b"a"] * k + [b"bb"] * k + [b"bbb"] * k
    for k in range(1, 10000, 2000):
        strs = [bytes([ord("a") + (i % 30)]) * 9 for i in range(k)]
        assert r.f(strs) == sorted(strs)


def test_pointer_lcp(oracle_mod):
    ss = gen.make(1, n=20000, seed=9)
    strs = ss.strings()
    out, lcp = gss.sort_strings(strs, lcp=True)
    assert out == sorted(strs)
    want = [0] * (len(out) - 1)
    for i in range(len(out) - 1):
        a, b = out[i], out[i + 1]
        j = 0
        while j < len(a) and j < len(b) and a[j] == b[j]:
            j += 1
        want[i] = j
    assert lcp == want


def test_pointer_span_text(oracle_mod):
    """Pointers into one text buffer (sortstring's layout): the output is the
    input pointer set permuted into msd_CE2's order."""
    ss = gen.make(0, n=100000, seed=4)
    base = ss.text.ctypes.data
    ptrs = (C.c_void_p * ss.n)(*[base + int(o) for o in ss.offsets])
    gss.sort_pointers(ptrs, ss.n)
    got = np.array([p - base for p in ptrs], dtype=np.uint32)
    assert np.array_equal(got, oracle_mod.sort_offsets(ss.text, ss.offsets))


def test_pointer_span_text_staged(oracle_mod):
    """A span of over 32 MB and 32 MB of pointers go through the staged
    copies (pinned chunks filled and drained by the host thread pool) both
    ways, the LCP array too: the output is the input pointers in msd_CE2's
    order and the LCPs are exact."""
    ss = gen.make(1, n=4_200_000, seed=12)
    base = ss.text.ctypes.data
    ptrs = (np.uint64(base) + ss.offsets.astype(np.uint64)).astype(np.uint64)
    lcp = np.zeros(ss.n - 1, dtype=np.uint64)
    log(f"staged pointer sort n={ss.n} text={ss.text_bytes}")
    gss.lib().gss_msd_lcp(ptrs.ctypes.data_as(C.POINTER(C.c_void_p)), ss.n,
                          lcp.ctypes.data_as(C.POINTER(C.c_size_t)))
    got = (ptrs - np.uint64(base)).astype(np.uint32)
    want, want_lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, lcp=True)
    assert np.array_equal(got, want)
    assert np.array_equal(lcp, want_lcp.astype(np.uint64))


def test_pointer_packed_path(oracle_mod):
    """Strings in two mappings separated by an unmapped hole force the packed
    upload; the result must still be the caller's own pointers."""
    libc = C.CDLL(None, use_errno=True)
    libc.mmap.restype = C.c_void_p
    libc.mmap.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_long]
    libc.munmap.argtypes = [C.c_void_p, C.c_size_t]
    page = mmap.PAGESIZE
    size = 64 * page
    addr = libc.mmap(None, 3 * size, mmap.PROT_READ | mmap.PROT_WRITE,
                     mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS, -1, 0)
    assert addr not in (None, C.c_void_p(-1).value)
    assert libc.munmap(addr + size, size) == 0  # the hole
    try:
        rng = np.random.default_rng(1)
        strs, ptrs = [], []
        for region in (addr, addr + 2 * size):
            pos = region
            while pos + 40 < region + size:
                s = bytes(rng.integers(97, 101, size=int(rng.integers(0, 12))).astype(np.uint8))
                C.memmove(pos, s + b"\0", len(s) + 1)
                strs.append(s)
                ptrs.append(pos)
                pos += len(s) + 1
        order = rng.permutation(len(ptrs))
        arr = (C.c_void_p * len(ptrs))(*[ptrs[i] for i in order])
        lcp = gss.sort_pointers(arr, len(ptrs), lcp=True)
        got = [C.string_at(p) for p in arr]
        assert got == sorted(strs)
        assert sorted(arr) == sorted(ptrs)
        # stable: equal strings keep their input order
        rank = {ptrs[i]: r for r, i in enumerate(order)}
        for a, b in zip(arr[:-1], arr[1:]):
            if C.string_at(a) == C.string_at(b):
                assert rank[a] < rank[b]
        assert lcp[0] == len(os.path.commonprefix([got[0], got[1]]))
    finally:
        libc.munmap(addr, size)
        libc.munmap(addr + 2 * size, size)


def _mmap_anon(size):
    libc = C.CDLL(None, use_errno=True)
    libc.mmap.restype = C.c_void_p
    libc.mmap.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_long]
    libc.munmap.argtypes = [C.c_void_p, C.c_size_t]
    addr = libc.mmap(None, size, mmap.PROT_READ | mmap.PROT_WRITE,
                     mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS, -1, 0)
    assert addr not in (None, C.c_void_p(-1).value)
    return libc, addr


def _lcps(strs):
    return [len(os.path.commonprefix([a, b])) for a, b in zip(strs[:-1], strs[1:])]


def test_pointer_sparse_span_is_packed():
    """A few thousand strings spread over one 512 MiB mapping: the span is
    readable and under 4 GiB but mostly empty, so the strings are packed and
    only their own bytes cross to the device."""
    size = 512 << 20
    libc, addr = _mmap_anon(size)
    try:
        rng = np.random.default_rng(5)
        slots = np.sort(rng.choice(size // 4096 - 1, size=3000, replace=False))
        strs, ptrs = [], []
        for k, slot in enumerate(slots):
            s = bytes(rng.integers(97, 100, size=int(rng.integers(0, 20))).astype(np.uint8))
            pos = addr + int(slot) * 4096 + int(rng.integers(0, 2048))
            C.memmove(pos, s + b"\0", len(s) + 1)
            strs.append(s)
            ptrs.append(pos)
        order = rng.permutation(len(ptrs))
        arr = (C.c_void_p * len(ptrs))(*[ptrs[i] for i in order])
        lcp = gss.sort_pointers(arr, len(ptrs), lcp=True)
        info = gss.pointer_info()
        log(f"sparse span: packed={info.packed} uploaded={info.uploaded_text_bytes} "
            f"pieces={info.pieces}")
        assert info.packed == 1 and info.pieces == 1
        assert info.uploaded_text_bytes == sum(len(s) + 1 for s in strs)
        got = [C.string_at(p) for p in arr]
        assert got == sorted(strs)
        # stable: equal strings keep their input order
        assert list(arr) == [ptrs[i] for i in sorted(order, key=lambda i: strs[i])]
        assert lcp == _lcps(got)
    finally:
        libc.munmap(addr, size)


def _deep_dups(n, seed):
    """Strings with a 300-byte shared prefix, exact duplicates, strings that
    end exactly where others go on, and an empty string."""
    rng = np.random.default_rng(seed)
    pre = b"p" * 300
    out = [pre + bytes(rng.integers(97, 100, size=int(rng.integers(0, 6))).astype(np.uint8))
           for _ in range(n)]
    out += out[: n // 4] + [pre, pre[:150], b""]
    return [out[i] for i in rng.permutation(len(out))]


@pytest.mark.parametrize("case", ["url_like", "zipf_words", "deep_dups"])
@pytest.mark.parametrize("limits", [(4096, 1 << 20), (1 << 20, 64), (2048, 32)])
def test_pointer_pieces(case, limits, oracle_mod):
    """With small device-sort limits the pointer path splits the input on
    the host by leading bytes into many device sorts: the output pointers
    and LCPs equal msd_CE2's order and the oracle's LCPs, as one device sort
    gives them (the same code path takes a sortstring input of over 4 GiB
    under the default limits)."""
    if case == "deep_dups":
        strs = _deep_dups(3000, 11)
        ss = gen.from_strings(strs, "deep_dups")
    else:
        ss = gen.make(1 if case == "url_like" else 3, n=20000, seed=21)
    base = ss.text.ctypes.data
    ptrs = (np.uint64(base) + ss.offsets.astype(np.uint64)).astype(np.uint64)
    lcp = np.zeros(ss.n - 1, dtype=np.uint64)
    gss.set_pointer_limits(*limits)
    try:
        log(f"pointer pieces {case} n={ss.n} limits={limits}")
        gss.lib().gss_msd_lcp(ptrs.ctypes.data_as(C.POINTER(C.c_void_p)), ss.n,
                              lcp.ctypes.data_as(C.POINTER(C.c_size_t)))
        info = gss.pointer_info()
    finally:
        gss.set_pointer_limits()
    log(f"  pieces={info.pieces} packed={info.packed} uploaded={info.uploaded_text_bytes}")
    assert info.packed == 1 and info.pieces > 1
    got = (ptrs - np.uint64(base)).astype(np.uint32)
    want, want_lcp = oracle_mod.sort_offsets(ss.text, ss.offsets, lcp=True)
    assert np.array_equal(got, want)
    assert np.array_equal(lcp, want_lcp.astype(np.uint64))


def test_pointer_text_over_4gib():
    """1,120,000 lines of 3,900 bytes in one 4.37 GB buffer (sortstring's
    layout for a file of over 4 GiB): the pointer path splits it into
    device sorts under 4 GiB each, and the order and LCPs equal what the
    strings' first 16 bytes fix (the other 3,884 bytes are one constant
    tail, and 1% of the lines repeat another line whole)."""
    n, width = 1_120_000, 3901
    rng = np.random.default_rng(44)
    text = np.empty((n, width), dtype=np.uint8)
    text[:, 16:] = ord("t")
    text[:, -1] = 0
    head = rng.integers(97, 123, size=(n, 16), dtype=np.uint8)
    dup = rng.choice(n, size=n // 100, replace=False)
    head[dup] = head[rng.choice(n, size=n // 100)]
    text[:, :16] = head
    flat = text.reshape(-1)
    assert flat.size > (1 << 32)
    base = flat.ctypes.data
    ptrs = np.uint64(base) + np.arange(n, dtype=np.uint64) * np.uint64(width)
    lcp = np.zeros(n - 1, dtype=np.uint64)
    log(f"pointer sort over 4 GiB: {flat.size} bytes, {n} lines")
    gss.lib().gss_msd_lcp(ptrs.ctypes.data_as(C.POINTER(C.c_void_p)), n,
                          lcp.ctypes.data_as(C.POINTER(C.c_size_t)))
    info = gss.pointer_info()
    log(f"  pieces={info.pieces} uploaded={info.uploaded_text_bytes}")
    assert info.packed == 1 and info.pieces >= 2 and info.uploaded_text_bytes == flat.size
    be = head.view(">u8")  # (n, 2) big-endian words: their order is the strings' order
    want = np.lexsort((np.arange(n), be[:, 1], be[:, 0]))
    got = ((ptrs - np.uint64(base)) // np.uint64(width)).astype(np.int64)
    assert np.array_equal(got, want)
    h = head[want]
    diff = h[1:] != h[:-1]
    first = np.where(diff.any(axis=1), diff.argmax(axis=1), width - 1)
    assert np.array_equal(lcp, first.astype(np.uint64))


@pytest.mark.skipif(not (REF / "unit-test").exists(), reason="oracle/_ref not built")
def test_reference_unit_test():
    """The reference's own unit test (unit-test/main.cpp), unchanged, with the
    GPU routines registered through routine.h."""
    p = subprocess.run([str(REF / "unit-test")], capture_output=True, text=True, timeout=600)
    log(p.stderr[-2000:])
    assert p.returncode == 0
    assert "[gss_msd]" in p.stderr and "[gss_msd_with_lcp]" in p.stderr
    assert "*** All OK ***" in p.stderr


@pytest.mark.skipif(not (REF / "sortstring").exists(), reason="oracle/_ref not built")
@pytest.mark.parametrize("config", [0, 1])
def test_reference_sortstring(config, tmp_path):
    """sortstring --check --write (src/sortstring.c:424-451): the GPU routine
    reports 'Check: GOOD' and writes the same bytes as msd_CE2."""
    ss = gen.make(config, n=200000, seed=21 + config)
    inp = tmp_path / "input.txt"
    inp.write_bytes(ss.newline_bytes())
    digests = {}
    for alg in ("msd_CE2", "gss_msd", "gss_msd_with_lcp"):
        out = tmp_path / f"{alg}.out"
        p = subprocess.run([str(REF / "sortstring"), "--check", f"--write={out}", alg, str(inp)],
                           capture_output=True, text=True, timeout=600, cwd=tmp_path)
        log(p.stdout[-500:] + p.stderr[-500:])
        assert p.returncode == 0
        assert "Check: GOOD" in p.stderr
        digests[alg] = hashlib.md5(out.read_bytes()).hexdigest()
    assert digests["gss_msd"] == digests["msd_CE2"] == digests["gss_msd_with_lcp"]
