"""Python access to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only to compare results with or as the timed CPU baseline.  The product
library never imports it.

* ``liboracle.so`` -- the plain-C restatement in oracle.c (built by
  ``make -C oracle``).
* ``_ref/libref.so`` -- the reference's own routines compiled from its sources
  (``make -C oracle ref``, only where the reference tree is present); looked
  up through the reference's registry (routine_from_name, src/routines.c:45-53).
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "liboracle.so"
REF_DIR = ORACLE_DIR / "_ref"
REF_LIB = REF_DIR / "libref.so"

ALGS = {"msd_ce2": 0, "multikey_cache8": 1}

_lib = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "liboracle.so"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        l.or_sort_offsets32.restype = C.c_int
        l.or_sort_offsets32.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                        C.c_void_p]
        l.or_check_result.restype = C.c_int
        l.or_check_result.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        l.or_key8.restype = C.c_uint64
        l.or_key8.argtypes = [C.c_char_p, C.c_size_t]
        l.or_lcp.restype = C.c_size_t
        l.or_lcp.argtypes = [C.c_char_p, C.c_char_p]
        l.or_lcp_array.restype = None
        l.or_lcp_array.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
        l.or_insertion_sort.restype = None
        l.or_insertion_sort.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        l.or_ingest32.restype = C.c_size_t
        l.or_ingest32.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t]
        l.or_write32.restype = C.c_size_t
        l.or_write32.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                                 C.c_size_t]
        l.or_compare.restype = C.c_int
        l.or_compare.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        l.or_merge_lcp32.restype = None
        l.or_merge_lcp32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
        for name in ("or_msd_ce2", "or_multikey_cache8"):
            fn = getattr(l, name)
            fn.restype = None
            fn.argtypes = [C.c_void_p, C.c_size_t]
        _lib = l
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def sort_offsets(text: np.ndarray, off: np.ndarray, alg: str = "msd_ce2", lcp: bool = False):
    """Sorted uint32 offsets (and uint64 LCPs) of the strings at `off`."""
    n = int(off.shape[0])
    off = np.ascontiguousarray(off, dtype=np.uint32)
    out = np.empty(n, dtype=np.uint32)
    lc = np.empty(max(1, n - 1), dtype=np.uint64) if lcp else None
    rc = lib().or_sort_offsets32(ALGS[alg], _ptr(text), _ptr(off), n, _ptr(out),
                                 _ptr(lc) if lcp else None)
    if rc:
        raise RuntimeError("oracle sort failed")
    return (out, lc[: max(0, n - 1)]) if lcp else out


def pointers(text: np.ndarray, off: np.ndarray) -> np.ndarray:
    return (np.uint64(text.ctypes.data) + off.astype(np.uint64)).astype(np.uint64)


def check_result(text: np.ndarray, off_sorted: np.ndarray):
    """(identical, invalid, wrong) as counted by src/util/debug.h:41-70."""
    p = pointers(text, off_sorted)
    a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
    lib().or_check_result(_ptr(p), p.shape[0], C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def lcp_array(text: np.ndarray, off_sorted: np.ndarray) -> np.ndarray:
    p = pointers(text, off_sorted)
    n = p.shape[0]
    out = np.zeros(max(1, n - 1), dtype=np.uint64)
    lib().or_lcp_array(_ptr(p), n, _ptr(out))
    return out[: max(0, n - 1)]


def key8(s: bytes, depth: int = 0) -> int:
    return lib().or_key8(bytes(s) + b"\0" * 8, depth)


def ingest(text: np.ndarray, delim: int = ord("\n")):
    """create_strings_delim on a copy of `text` (uint8): returns the uint32
    offsets and the rewritten text."""
    t = np.ascontiguousarray(text, dtype=np.uint8).copy()
    n = lib().or_ingest32(_ptr(t), t.shape[0], int(delim), None, 0)
    t2 = np.ascontiguousarray(text, dtype=np.uint8).copy()
    off = np.empty(max(1, n), dtype=np.uint32)
    lib().or_ingest32(_ptr(t2), t2.shape[0], int(delim), _ptr(off), n)
    return off[:n], t2


def write(text: np.ndarray, off: np.ndarray, length: int | None = None) -> bytes:
    """write_result's bytes for the strings at `off` (in that order)."""
    off = np.ascontiguousarray(off, dtype=np.uint32)
    ln = text.shape[0] if length is None else int(length)
    total = lib().or_write32(_ptr(text), ln, _ptr(off), off.shape[0], None, 0)
    out = np.empty(max(1, total), dtype=np.uint8)
    lib().or_write32(_ptr(text), ln, _ptr(off), off.shape[0], _ptr(out), total)
    return out[:total].tobytes()


def merge_lcp(text: np.ndarray, a, alcp, b, blcp):
    """merge_lcp_2way of two sorted offset runs with their n-1 LCPs."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    alcp = np.ascontiguousarray(alcp if alcp is not None else np.zeros(1), dtype=np.uint64)
    blcp = np.ascontiguousarray(blcp if blcp is not None else np.zeros(1), dtype=np.uint64)
    n = a.shape[0] + b.shape[0]
    out = np.empty(max(1, n), dtype=np.uint32)
    ol = np.empty(max(1, n - 1), dtype=np.uint64)
    lib().or_merge_lcp32(_ptr(text), _ptr(a), _ptr(alcp), a.shape[0], _ptr(b), _ptr(blcp),
                         b.shape[0], _ptr(out), _ptr(ol))
    return out[:n], ol[: max(0, n - 1)]


# ---------------------------------------------------------------------------
# The reference's compiled routines (oracle/_ref/libref.so)

class _Routine(C.Structure):
    _fields_ = [("f", C.c_void_p), ("name", C.c_char_p), ("desc", C.c_char_p),
                ("multicore", C.c_uint, 1)]


_SORT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t)


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise FileNotFoundError(f"{REF_LIB} is not built (make -C oracle ref)")
        l = C.CDLL(str(REF_LIB))
        l.routine_from_name.restype = C.POINTER(_Routine)
        l.routine_from_name.argtypes = [C.c_char_p]
        _ref = l
    return _ref


def ref_routine(name: str):
    r = ref_lib().routine_from_name(name.encode())
    if not r:
        raise KeyError(name)
    return _SORT_FN(r.contents.f)


def ref_sort_pointers(name: str, ptrs: np.ndarray) -> None:
    """Runs reference routine `name` on a uint64 pointer array in place."""
    ref_routine(name)(_ptr(ptrs), ptrs.shape[0])


def ref_sort_offsets(text: np.ndarray, off: np.ndarray, name: str = "msd_CE2") -> np.ndarray:
    p = pointers(text, off)
    ref_sort_pointers(name, p)
    return (p - np.uint64(text.ctypes.data)).astype(np.uint32)
